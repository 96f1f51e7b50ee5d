/*
Copyright (c) Facebook, Inc. and its affiliates.

This source code is licensed under the MIT license found in the
LICENSE file in the root directory of this source tree.

The declarations below restate the reference loop_tool's public interface
(include/loop_tool/) so that code written against it compiles unchanged.
*/
// Text serialization of an IR (nodes, constraints, orders, annotations,
// reuse flags, inputs/outputs) so that a tuned schedule can be stored and
// reloaded.  Same entry points as include/loop_tool/serialization.h of the
// reference; the text format is this project's own (one record per line,
// expressions in prefix form) and round-trips losslessly.
#pragma once

#include <string>

#include "ir.h"

namespace loop_tool {

std::string serialize(const IR &ir);
IR deserialize(const std::string &str);

}  // namespace loop_tool
