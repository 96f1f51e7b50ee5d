// Dataflow view of a LoopTree used by the HIP kernel emitters.
//
// Every compute node is treated as a pure function of its vars: pointwise
// nodes combine their inputs, reductions fold their inputs over the reduction
// vars in the order given by the reduction loops above the node's leaf (the
// same order the reference C backend accumulates in,
// src/backends/cpu/cpp.cpp:295-351), and unscheduled views are folded into the
// affine index expressions of the accesses (src/core/compile.cpp:611-888).
// Out-of-range view accesses read as 0.
#pragma once

#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "loop_tool/compile.h"

namespace loop_tool {
namespace hip {

struct RedLoop {
  IR::VarRef var;
  int64_t size;
  int64_t tail;
  int64_t inner;  // extent of the same var covered by loops below this one
  bool unroll = false;  // the loop carries the "unroll" annotation
};

struct DFInput {
  IR::NodeRef src;                    // resolved producer (read or compute)
  std::vector<symbolic::Expr> coords;  // one per src var, over loop-var symbols
};

struct DFNode {
  IR::NodeRef nr = -1;
  Operation op = Operation::constant;
  std::vector<IR::VarRef> vars;       // dims of the node's value
  std::vector<IR::VarRef> red_vars;   // reduced vars (loop vars not in vars)
  std::vector<DFInput> inputs;
  std::vector<RedLoop> red_loops;     // reduction loop chain, outer to inner
  int input_index = -1;               // position in IR::inputs() for reads
  int output_index = -1;              // position in IR::outputs() for writes
  int consumers = 0;                  // number of uses by other DF nodes
  // the node's loops in the LoopTree: per var symbol id, the sizes of the
  // loops over that var above the node's leaf, outer to inner (two or more
  // when the schedule split the var)
  std::map<int, std::vector<int64_t>> loops;
};

struct Dataflow {
  std::unique_ptr<Compiler> compiler;
  std::map<IR::NodeRef, DFNode> nodes;  // reads, computes and writes
  std::vector<IR::NodeRef> topo;        // compute/write nodes in dependence order
  std::vector<IR::NodeRef> writes;      // in IR::outputs() order
  int num_inputs = 0;

  const IR &ir() const { return compiler->lt.ir; }
  int64_t size(IR::VarRef v) const { return compiler->var_sizes.at(v); }
  const symbolic::Symbol &sym(IR::VarRef v) const { return compiler->var_to_sym.at(v); }
  int64_t numel(const std::vector<IR::VarRef> &vars) const {
    int64_t n = 1;
    for (auto v : vars) n *= size(v);
    return n;
  }
  // Where the schedule split var `sym` into nested loops above node nr's
  // leaf: the size of the innermost of them (level 0), or of the one above
  // it (level 1, when there are three or more); 0 where it did not.
  int64_t split_extent(IR::NodeRef nr, int sym, int level = 0) const {
    auto n = nodes.find(nr);
    if (n == nodes.end()) return 0;
    auto l = n->second.loops.find(sym);
    if (l == n->second.loops.end() || (int)l->second.size() < 2 + level) return 0;
    return l->second[l->second.size() - 1 - level];
  }
};

// Throws (ASSERT) on trees the dataflow model cannot express.
std::unique_ptr<Dataflow> analyze(const LoopTree &lt);

// Smallest/largest value of an integer expression when each symbol ranges
// over [0, size-1]; `ok` is false for non-affine expressions.
struct Range {
  int64_t lo = 0, hi = 0;
  bool ok = false;
};
Range expr_range(const symbolic::Expr &e,
                 const std::map<int, int64_t> &sym_sizes);

}  // namespace hip
}  // namespace loop_tool
