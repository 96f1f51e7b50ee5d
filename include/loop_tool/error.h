/*
Copyright (c) Facebook, Inc. and its affiliates.

This source code is licensed under the MIT license found in the
LICENSE file in the root directory of this source tree.

The declarations below restate the reference loop_tool's public interface
(include/loop_tool/) so that code written against it compiles unchanged.
*/
// Assertion helper of loop_tool: `ASSERT(cond) << "message"` throws a
// std::runtime_error carrying the condition text, the source location and the
// streamed message when `cond` is false.
// Interface kept from include/loop_tool/error.h:27-61 of the reference
// (same macro name and throwing behaviour, so callers and tests see the same
// exception type and can match on "assertion:").
#pragma once

#include <sstream>
#include <stdexcept>
#include <string>

#define LT_STRINGIFY_(x) #x
#define LT_STRINGIFY(x) LT_STRINGIFY_(x)
#define LOCATION __FILE__ ":" LT_STRINGIFY(__LINE__)

namespace loop_tool {

struct NullStream {
  template <typename T>
  NullStream &operator<<(T const &) {
    return *this;
  }
};

namespace detail {

// Collects the message and throws from its destructor, so that the whole
// `ASSERT(x) << a << b` expression is evaluated before the throw.
class AssertStream {
 public:
  AssertStream(const char *location, const char *cond) {
    msg_ << "assertion: " << cond << " failed @ " << location << " ";
  }
  template <typename T>
  AssertStream &operator<<(const T &v) {
    msg_ << v;
    return *this;
  }
  ~AssertStream() noexcept(false) { throw std::runtime_error(msg_.str()); }

 private:
  std::ostringstream msg_;
};

}  // namespace detail
}  // namespace loop_tool

#ifdef NOEXCEPTIONS
#define ASSERT(x) loop_tool::NullStream()
#else
#define ASSERT(x) \
  if (x) {        \
  } else          \
    ::loop_tool::detail::AssertStream(LOCATION, #x)
#endif
