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
