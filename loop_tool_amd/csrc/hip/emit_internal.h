// Internal emitter interfaces of the hip backend.
#pragma once

#include <map>
#include <memory>
#include <string>

#include "emit.h"

namespace loop_tool {
namespace hip {

// Emits one-lane-per-element kernels (see emit_pointwise.cpp).
class PointwiseEmitter {
 public:
  // materialized: node -> scratch number (argument name "s<k>");
  // input_args: IR input index -> argument name
  PointwiseEmitter(const Dataflow &df, const std::map<IR::NodeRef, int> &materialized,
                   const std::map<int, std::string> &input_args);
  ~PointwiseEmitter();
  // A kernel computing every element of `target` (a write or a materialised
  // compute node) into `out_arg`.  Returns the source; *threads = elements.
  std::string kernel(const std::string &name, IR::NodeRef target,
                     const std::string &signature, const std::string &out_arg,
                     int block, int64_t *threads, int vec, bool int32);

  struct Impl;

 private:
  std::unique_ptr<Impl> impl_;
};

// Device helpers shared by every generated module.
const char *kernel_prelude();

// Row-cooperative lowering (emit_rowreduce.cpp); returns false when the
// nest does not have the row-reduction shape.
bool try_rowreduce(const Dataflow &df, const Knobs &knobs, Plan &plan);
// Contraction lowering on MFMA (emit_contract.cpp).
bool try_contract(const Dataflow &df, const Knobs &knobs, Plan &plan);

}  // namespace hip
}  // namespace loop_tool
