/*
Copyright (c) Facebook, Inc. and its affiliates.

This source code is licensed under the MIT license found in the
LICENSE file in the root directory of this source tree.

The declarations below restate the reference loop_tool's public interface
(include/loop_tool/) so that code written against it compiles unchanged.
*/
// dlopen/dlsym wrapper (interface of include/loop_tool/dynlib.h:16-52 of the
// reference).
#pragma once
#include <dlfcn.h>

#include <string>

#include "error.h"

namespace loop_tool {

struct DynamicLibrary {
  explicit DynamicLibrary(const char *name, bool expose_symbols = false)
      : name_(name) {
    lib_ = dlopen(name, (expose_symbols ? RTLD_GLOBAL : RTLD_LOCAL) | RTLD_NOW);
    ASSERT(lib_) << "Couldn't load library " << name_ << " dlerror: " << dlerror();
  }
  DynamicLibrary(const DynamicLibrary &) = delete;
  DynamicLibrary &operator=(const DynamicLibrary &) = delete;

  static bool exists(const char *name, bool expose_symbols = false) {
    return dlopen(name, (expose_symbols ? RTLD_GLOBAL : RTLD_LOCAL) | RTLD_NOW) != nullptr;
  }

  void *sym(const char *sym_name) const {
    ASSERT(lib_) << "Library " << name_ << " not loaded for symbol " << sym_name;
    void *s = dlsym(lib_, sym_name);
    ASSERT(s) << "Couldn't find " << sym_name << " in " << name_;
    return s;
  }

  template <typename F>
  F sym(const char *sym_name) const {
    F fn;
    reinterpret_cast<void *&>(fn) = sym(sym_name);
    return fn;
  }

  ~DynamicLibrary() {
    if (lib_) dlclose(lib_);
  }

 private:
  void *lib_ = nullptr;
  std::string name_;
};

#define DYNLIB(lib, name) reinterpret_cast<decltype(&name)>(lib->sym(#name))

}  // namespace loop_tool
