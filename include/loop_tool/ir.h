/*
Copyright (c) Facebook, Inc. and its affiliates.

This source code is licensed under the MIT license found in the
LICENSE file in the root directory of this source tree.

The declarations below restate the reference loop_tool's public interface
(include/loop_tool/) so that code written against it compiles unchanged.
*/
// Dataflow IR of loop_tool and the LoopTree derived from it.
//
// Interface kept from include/loop_tool/ir.h:19-422 of the reference: the
// Operation list (and its order, which the serialized text format relies on),
// IR node/var creation, per-node loop orders with {size, tail}, priorities,
// reuse switches, annotations, and the LoopTree (loops + leaves, LCA, walk,
// dump).  The LoopTree constructor reproduces the reference's greedy loop
// reuse (src/core/ir.cpp:531-728) because the loop structure decides reduction
// order and scratch scoping, which every backend must agree on.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "error.h"
#include "symbolic.h"

namespace loop_tool {

#define LT_OPS(_) \
  _(constant)     \
  _(write)        \
  _(read)         \
  _(view)         \
  _(copy)         \
  _(add)          \
  _(subtract)     \
  _(multiply)     \
  _(divide)       \
  _(min)          \
  _(max)          \
  _(log)          \
  _(exp)          \
  _(sqrt)         \
  _(abs)          \
  _(negate)       \
  _(reciprocal)   \
  _(name)

enum struct Operation {
#define LT_X(op) op,
  LT_OPS(LT_X)
#undef LT_X
};

std::string dump(const Operation &op);

template <typename T, template <typename _> class H = std::hash>
std::unordered_set<T, H<T>> to_set(std::vector<T> v) {
  // element-wise insertion: iteration order of the resulting set is part of
  // the observable behaviour (toposort tie-breaking), so build it the same
  // way every time
  std::unordered_set<T, H<T>> out;
  for (const auto &e : v) out.insert(e);
  return out;
}

template <typename T, template <typename _> class H = std::hash>
std::unordered_set<T, H<T>> intersection(const std::unordered_set<T, H<T>> &a,
                                         const std::unordered_set<T, H<T>> &b) {
  std::unordered_set<T, H<T>> c;
  for (const auto &e : a)
    if (b.count(e)) c.insert(e);
  return c;
}

class Node;
class Var;

class IR {
 public:
  using NodeRef = int32_t;
  using VarRef = int32_t;
  struct LoopSize {
    int64_t size;
    int64_t tail;
  };

  IR() {}

  NodeRef create_node(Operation op, std::vector<NodeRef> inputs,
                      std::vector<VarRef> vars,
                      std::vector<symbolic::Constraint> constraints = {},
                      std::unordered_map<int, VarRef> sym_var_map = {});
  VarRef create_var(std::string name);

  void delete_node(const NodeRef &node_ref);
  void replace_all_uses(NodeRef old_node, NodeRef new_node);
  void update_inputs(NodeRef node_ref, std::vector<NodeRef> inputs);
  void update_vars(NodeRef node_ref, std::vector<VarRef> vars);

  std::vector<VarRef> vars() const;
  std::vector<NodeRef> nodes() const;

  const Node &node(NodeRef ref) const;
  Node &node(NodeRef ref);
  const Var &var(VarRef ref) const;
  Var &var(VarRef ref);

  float priority(NodeRef ref) const { return priorities_[ref]; }
  const std::vector<std::pair<VarRef, LoopSize>> &order(NodeRef ref) const {
    return orders_[ref];
  }
  const std::unordered_set<int> not_reusable(NodeRef ref) const {
    return reuse_disabled_[ref];
  }

  const std::vector<NodeRef> &inputs() const { return inputs_; }
  const std::vector<NodeRef> &outputs() const { return outputs_; }
  void add_input(NodeRef input) { inputs_.emplace_back(input); }
  void add_output(NodeRef output) { outputs_.emplace_back(output); }
  void set_inputs(std::vector<NodeRef> inputs) { inputs_ = std::move(inputs); }
  void set_outputs(std::vector<NodeRef> outputs) { outputs_ = std::move(outputs); }

  void reset_aux(NodeRef node_ref);
  void set_priority(NodeRef ref, float priority) { priorities_[ref] = priority; }
  void set_order(NodeRef ref, std::vector<std::pair<VarRef, LoopSize>> order) {
    loop_annotations_[ref].assign(order.size(), "");
    orders_[ref] = std::move(order);
  }
  void set_order(NodeRef ref, std::vector<std::pair<VarRef, LoopSize>> order,
                 std::vector<std::string> annotations) {
    ASSERT(annotations.size() == order.size());
    orders_[ref] = std::move(order);
    annotate_loops(ref, std::move(annotations));
  }
  void disable_reuse(NodeRef ref, int order_ref) {
    reuse_disabled_[ref].insert(order_ref);
  }
  void enable_reuse(NodeRef ref, int order_ref) {
    reuse_disabled_[ref].erase(order_ref);
  }
  void annotate_loop(NodeRef ref, int idx, std::string annot) {
    loop_annotations_[ref].at(idx) = std::move(annot);
  }
  void annotate_loops(NodeRef ref, std::vector<std::string> annots) {
    loop_annotations_[ref] = std::move(annots);
  }
  std::vector<std::string> loop_annotations(NodeRef ref) const {
    return loop_annotations_.at(ref);
  }
  void annotate(NodeRef ref, std::string annot) {
    annotations_[ref] = std::move(annot);
  }
  std::string annotation(NodeRef ref) const { return annotations_.at(ref); }

  std::string dump(NodeRef ref) const;
  std::vector<VarRef> pointwise_vars(NodeRef ref) const;
  std::vector<VarRef> reduction_vars(NodeRef ref) const;
  std::vector<VarRef> view_reduction_vars(NodeRef idx) const;
  std::vector<VarRef> loop_vars(NodeRef ref) const;
  std::vector<VarRef> all_vars(NodeRef ref) const;
  void reify_deletions();

 private:
  std::vector<Node> nodes_;
  std::unordered_set<NodeRef> deleted_;
  std::vector<Var> vars_;
  std::vector<float> priorities_;
  std::vector<std::vector<std::pair<VarRef, LoopSize>>> orders_;
  std::vector<std::unordered_set<int>> reuse_disabled_;
  std::vector<std::vector<std::string>> loop_annotations_;
  std::vector<std::string> annotations_;
  std::vector<NodeRef> inputs_;
  std::vector<NodeRef> outputs_;
};

// Graphviz text of the IR (for inspection only).
std::string dot(const IR &ir);
// Priority-ordered topological sort starting from the IR inputs.
std::vector<IR::NodeRef> toposort(const IR &ir);

class Node {
 protected:
  friend class IR;
  Node(Operation op, std::vector<IR::NodeRef> inputs,
       std::vector<IR::VarRef> vars,
       std::vector<symbolic::Constraint> constraints,
       std::unordered_map<int, IR::VarRef> sym_var_map)
      : op_(op),
        inputs_(std::move(inputs)),
        vars_(std::move(vars)),
        constraints_(std::move(constraints)),
        sym_var_map_(std::move(sym_var_map)) {}

  void replace_input(IR::NodeRef old_node, IR::NodeRef new_node);
  void update_inputs(std::vector<IR::NodeRef> inputs) { inputs_ = std::move(inputs); }
  void update_outputs(std::vector<IR::NodeRef> outputs) {
    outputs_ = std::move(outputs);
  }
  void update_vars(std::vector<IR::VarRef> vars) { vars_ = std::move(vars); }

 public:
  const std::vector<IR::NodeRef> &inputs() const { return inputs_; }
  const std::vector<IR::NodeRef> &outputs() const { return outputs_; }
  const std::vector<symbolic::Constraint> &constraints() const {
    return constraints_;
  }
  IR::VarRef var(symbolic::Symbol sym) const {
    auto it = sym_var_map_.find(sym.id());
    ASSERT(it != sym_var_map_.end())
        << "symbol " << sym.name() << "#" << sym.id()
        << " is not mapped to a variable";
    return it->second;
  }
  bool has_sym(symbolic::Symbol sym) const { return sym_var_map_.count(sym.id()); }
  const std::unordered_map<int, IR::VarRef> &sym_to_var() const {
    return sym_var_map_;
  }
  const Operation &op() const { return op_; }
  const std::vector<IR::VarRef> &vars() const { return vars_; }
  std::vector<IR::VarRef> &vars() { return vars_; }
  void remap_refs(const std::unordered_map<IR::NodeRef, IR::NodeRef> &);

 private:
  Operation op_;
  std::vector<IR::NodeRef> inputs_;
  std::vector<IR::VarRef> vars_;
  std::vector<symbolic::Constraint> constraints_;
  std::unordered_map<int, IR::VarRef> sym_var_map_;

 protected:
  std::vector<IR::NodeRef> outputs_;
};

class Var {
 public:
  Var(std::string name, int version) : name_(std::move(name)), version_(version) {}
  const std::string &name() const { return name_; }
  const int &version() const { return version_; }

 private:
  std::string name_;
  int version_;
};

struct LoopTree {
  using TreeRef = int;
  enum { NODE = 0, LOOP = 1 };

  struct Loop {
    IR::VarRef var;
    int var_depth;
    int64_t size;
    int64_t tail;
    bool operator==(const Loop &o) const {
      return var == o.var && var_depth == o.var_depth && size == o.size &&
             tail == o.tail;
    }
  };

  struct LoopTreeNode {
    TreeRef parent = -1;
    TreeRef idx = -1;
    int depth = 0;
    int annotation = -1;
    bool kind;  // NODE or LOOP
    IR::NodeRef node = -1;
    Loop loop{-1, 0, 0, 0};
    std::vector<int> children;

    LoopTreeNode(int p, int i, IR::NodeRef n) : parent(p), idx(i), kind(NODE), node(n) {}
    LoopTreeNode(int p, int i, const Loop &l) : parent(p), idx(i), kind(LOOP), loop(l) {}
  };

  TreeRef add_leaf(TreeRef parent, IR::NodeRef n);
  TreeRef add_loop(TreeRef parent, const Loop &l);

  const LoopTreeNode &tree_node(TreeRef ref) const {
    ASSERT(ref >= 0 && ref < (TreeRef)nodes.size()) << "bad tree ref " << ref;
    return nodes[ref];
  }
  LoopTreeNode &tree_node(TreeRef ref) {
    ASSERT(ref >= 0 && ref < (TreeRef)nodes.size()) << "bad tree ref " << ref;
    return nodes[ref];
  }
  bool kind(TreeRef ref) const { return tree_node(ref).kind; }
  const IR::NodeRef node(TreeRef ref) const {
    ASSERT(kind(ref) == LoopTree::NODE) << "tree ref " << ref << " is a loop";
    return tree_node(ref).node;
  }
  Loop loop(TreeRef ref) const {
    ASSERT(kind(ref) == LoopTree::LOOP) << "tree ref " << ref << " is a node";
    return tree_node(ref).loop;
  }
  std::unordered_set<IR::VarRef> scope_vars(TreeRef ref) const;
  const TreeRef parent(TreeRef ref) const { return tree_node(ref).parent; }
  const std::vector<TreeRef> &children(TreeRef ref) const {
    if (ref == -1) return roots;
    return tree_node(ref).children;
  }
  void annotate_(TreeRef ref, std::string annot);
  std::string annotation(TreeRef ref) const {
    auto i = tree_node(ref).annotation;
    return i > -1 ? annotations[i] : "";
  }
  int depth(TreeRef ref) const { return tree_node(ref).depth; }
  TreeRef lca(TreeRef a, TreeRef b) const;
  std::vector<LoopTree::Loop> loop_order(IR::NodeRef ref) const;

  void walk(const std::function<void(LoopTree::TreeRef, int)> &fn,
            TreeRef start = -1) const;
  std::string dump(
      const std::function<std::string(LoopTree::TreeRef)> &fn = {}) const;

  IR ir;
  explicit LoopTree(const IR &ir_);

  std::vector<LoopTreeNode> nodes;
  std::vector<TreeRef> roots;
  std::vector<std::string> annotations;
  std::unordered_map<IR::NodeRef, TreeRef> scheduled;
};

}  // namespace loop_tool
