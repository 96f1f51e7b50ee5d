/*
Copyright (c) Facebook, Inc. and its affiliates.

This source code is licensed under the MIT license found in the
LICENSE file in the root directory of this source tree.

The declarations below restate the reference loop_tool's public interface
(include/loop_tool/) so that code written against it compiles unchanged.
*/
// A raw float buffer of N elements allocated on one registered hardware.
// Interface kept from include/loop_tool/tensor.h:15-24 of the reference.
#pragma once
#include <cstddef>

#include "hardware.h"

namespace loop_tool {

struct Tensor {
  Tensor(size_t N, int hardware = 0);
  Tensor() = delete;
  Tensor(const Tensor &) = delete;
  Tensor(Tensor &&) = default;
  ~Tensor();
  int hardware_id = -1;
  Memory data;
  size_t numel;
};

}  // namespace loop_tool
