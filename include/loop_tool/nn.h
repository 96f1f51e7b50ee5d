/*
Copyright (c) Facebook, Inc. and its affiliates.

This source code is licensed under the MIT license found in the
LICENSE file in the root directory of this source tree.

The declarations below restate the reference loop_tool's public interface
(include/loop_tool/) so that code written against it compiles unchanged.
*/
// Neural-network helpers built from lazy views: convolution as a window view
// times weights summed over the window and shared channels, and max-pooling
// as a window view max-reduced.  Interface of include/loop_tool/nn.h:14-19.
#pragma once

#include "lazy.h"

namespace loop_tool {
namespace nn {

lazy::Tensor convolve(lazy::Tensor X, lazy::Tensor W,
                      std::vector<symbolic::Symbol> spatial_dims,
                      std::vector<symbolic::Symbol> window_dims, int stride = 1);
lazy::Tensor maxpool(lazy::Tensor X, std::vector<symbolic::Symbol> spatial_dims,
                     int k, int stride = 1);

}  // namespace nn
}  // namespace loop_tool
