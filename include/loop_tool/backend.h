/*
Copyright (c) Facebook, Inc. and its affiliates.

This source code is licensed under the MIT license found in the
LICENSE file in the root directory of this source tree.

The declarations below restate the reference loop_tool's public interface
(include/loop_tool/) so that code written against it compiles unchanged.
*/
// Backend plugin interface of loop_tool: a Backend compiles a LoopTree into a
// Compiled object whose run() takes one buffer per IR input and output, in
// IR::inputs() then IR::outputs() order (dense row-major float32).
//
// Interface kept verbatim from include/loop_tool/backend.h:18-92 of the
// reference so that out-of-tree backends (this project's "hip") and callers
// (the lazy frontend, tests) compile against it unchanged.
#pragma once

#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "ir.h"
#include "tensor.h"

namespace loop_tool {

struct Compiled {
  virtual ~Compiled() {}
  virtual void run(const std::vector<void *> &memory, bool sync = true) const = 0;
  virtual std::string dump() const {
    return "[not implemented, override `std::string Compiled::dump() const`]";
  }

  void operator()(const std::vector<Tensor *> &tensors, bool sync = true) const;

  std::vector<void *> allocate(std::vector<int64_t> &sizes) const;

  template <bool sync, typename... Args>
  void run(Args const &...tensors) const {
    std::vector<void *> memory = {tensors.data.address...};
    run(memory, sync);
  }

  template <typename... Args>
  void operator()(Args const &...tensors) const {
    run<true, Args...>(std::forward<Args const &>(tensors)...);
  }

  template <typename... Args>
  void async(Args const &...tensors) const {
    run<false, Args...>(std::forward<Args const &>(tensors)...);
  }

  std::unordered_map<std::string, int> int_properties;
  std::unordered_map<std::string, std::string> string_properties;
  int hardware_requirement = -1;
  std::string name;
};

struct Backend {
  std::string name_;
  explicit Backend(std::string name) : name_(std::move(name)) {}
  virtual ~Backend() {}

  const std::string &name() const { return name_; }

  virtual std::unique_ptr<Compiled> compile_impl(const LoopTree &lt) const = 0;
  virtual int hardware_requirement() const = 0;

  std::unique_ptr<Compiled> compile(const LoopTree &lt) const {
    auto compiled = compile_impl(lt);
    compiled->hardware_requirement = hardware_requirement();
    compiled->name = name();
    return compiled;
  }
};

const std::unordered_map<std::string, std::shared_ptr<Backend>> &getBackends();
void registerBackend(std::shared_ptr<Backend> backend);

std::shared_ptr<Backend> &getDefaultBackend();
void setDefaultBackend(std::string backend);

struct ScopedBackend {
  std::string old_backend_name;
  explicit ScopedBackend(std::string backend_name) {
    old_backend_name = getDefaultBackend()->name();
    setDefaultBackend(backend_name);
  }
  ~ScopedBackend() { setDefaultBackend(old_backend_name); }
};

struct RegisterBackend {
  explicit RegisterBackend(std::shared_ptr<Backend> backend) {
    registerBackend(backend);
  }
};

// dlopen a shared library with RTLD_GLOBAL so its static registrations run.
void loadLibrary(std::string lib_name);

}  // namespace loop_tool
