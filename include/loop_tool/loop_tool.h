/*
Copyright (c) Facebook, Inc. and its affiliates.

This source code is licensed under the MIT license found in the
LICENSE file in the root directory of this source tree.

The declarations below restate the reference loop_tool's public interface
(include/loop_tool/) so that code written against it compiles unchanged.
*/
// Umbrella header of the loop_tool API restated by this project
// (include/loop_tool/loop_tool.h of the reference).
#pragma once

#include "backend.h"
#include "compile.h"
#include "hardware.h"
#include "ir.h"
#include "serialization.h"
#include "tensor.h"
// default frontends
#include "lazy.h"
#include "mutate.h"
#include "nn.h"
