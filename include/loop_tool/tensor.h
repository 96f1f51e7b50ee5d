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
