/*
Copyright (c) Facebook, Inc. and its affiliates.

This source code is licensed under the MIT license found in the
LICENSE file in the root directory of this source tree.

The declarations below restate the reference loop_tool's public interface
(include/loop_tool/) so that code written against it compiles unchanged.
*/
// Lowering analysis shared by every backend: variable sizes, per-loop inner
// strides, buffer allocations (inputs | outputs | scratch) with their LCA
// scope, and per-access index equations with bound predicates.
//
// Interface kept from include/loop_tool/compile.h:24-120 of the reference (the
// analysis half of `Compiler`).  The tree interpreter that the reference also
// hangs on this class (gen_exec & co., compile.cpp:163-424) is test-only here
// and lives with the CPU oracle.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "backend.h"
#include "ir.h"

#define MAX_DEPTH 8

namespace loop_tool {

class Compiler {
 public:
  struct Allocation {
    Allocation() = default;
    Allocation(int memory_idx, IR::NodeRef node_ref_)
        : mem_idx(memory_idx), node_ref(node_ref_) {}
    Allocation(int memory_idx, IR::NodeRef node_ref_,
               const std::vector<int64_t> &sizes_, LoopTree::TreeRef lca_)
        : mem_idx(memory_idx), sizes(sizes_), node_ref(node_ref_), lca(lca_) {}
    int mem_idx = -1;
    std::vector<int64_t> sizes;  // scoped sizes, one per node var
    std::vector<int64_t> strides;
    int64_t size(int start_idx = 0) const {
      int64_t s = 1;
      for (size_t i = start_idx; i < sizes.size(); ++i)
        s *= sizes[i] > 1 ? sizes[i] : 1;
      return s;
    }
    IR::NodeRef node_ref = -1;
    LoopTree::TreeRef lca = -1;
  };

  struct Access {
    explicit Access(const Allocation &a) : alloc(a) {}
    Allocation alloc;
    std::vector<symbolic::Expr> full_exprs;    // per buffer dim, all loop syms
    std::vector<symbolic::Expr> scoped_exprs;  // loop syms below alloc.lca only
    std::vector<std::pair<int64_t, int64_t>> bounds;  // (min, max or -1)
  };

  mutable bool set_called = false;
  size_t count;
  LoopTree lt;
  std::unordered_map<LoopTree::TreeRef, int64_t> inner_sizes;
  std::unordered_map<IR::NodeRef, Allocation> allocations;
  std::unordered_map<IR::NodeRef, IR::NodeRef> resolved_reads;
  std::unordered_map<IR::NodeRef, IR::NodeRef> resolved_writes;
  std::unordered_map<IR::VarRef, int64_t> var_sizes;
  std::unordered_map<IR::VarRef, symbolic::Symbol> var_to_sym;
  std::unordered_map<symbolic::Symbol, IR::VarRef, symbolic::Hash<symbolic::Symbol>>
      sym_to_var;

  explicit Compiler(const LoopTree &lt_);
  virtual ~Compiler() = default;

  Allocation gen_alloc(IR::NodeRef node_ref) const;

  std::pair<std::vector<symbolic::Expr>, std::vector<symbolic::Expr>>
  gen_index_equations(IR::NodeRef read_node_ref, IR::NodeRef write_node_ref,
                      LoopTree::TreeRef ref) const;
  // access information for `node` used at schedule point `ref`
  Access gen_access(IR::NodeRef node, LoopTree::TreeRef ref) const;

  symbolic::Expr get_scoped_expr(const Access &access) const;
  std::unordered_map<symbolic::Symbol,
                     std::vector<std::pair<LoopTree::TreeRef, int64_t>>,
                     symbolic::Hash<symbolic::Symbol>>
  get_symbol_strides(LoopTree::TreeRef ref, LoopTree::TreeRef root) const;
  std::vector<std::pair<symbolic::Expr, int64_t>> get_constraints(
      const Access &access) const;

  symbolic::Expr reify_sizes(const symbolic::Expr &expr) const;
  int64_t get_expr_max(const symbolic::Expr &) const;
  int64_t get_expr_min(const symbolic::Expr &) const;

  virtual std::string gen_string() const;

  std::vector<void *> allocate() const;
  std::vector<int64_t> memory_sizes(bool include_io = false) const;
};

}  // namespace loop_tool
