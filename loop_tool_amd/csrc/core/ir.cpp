// IR and LoopTree (see include/loop_tool/ir.h).
// Reference behaviour: src/core/ir.cpp (create_node :33-68, dump :114-166,
// pointwise/reduction/view-reduction vars :168-305, toposort :380-410,
// LoopTree :420-728, dot :730-787).
#include "loop_tool/ir.h"

#include <algorithm>
#include <sstream>
#include <tuple>

namespace loop_tool {

using symbolic::Expr;
using symbolic::Symbol;

std::string dump(const Operation &op) {
  switch (op) {
#define LT_X(o)       \
  case Operation::o: \
    return #o;
    LT_OPS(LT_X)
#undef LT_X
  }
  ASSERT(0) << "unknown op code " << (int)op;
  return "unknown";
}

// ------------------------------------------------------------------- IR ----
IR::VarRef IR::create_var(std::string name) {
  int version = 0;
  for (const auto &v : vars_)
    if (v.name() == name) version++;
  vars_.emplace_back(name, version);
  return (VarRef)vars_.size() - 1;
}

IR::NodeRef IR::create_node(Operation op, std::vector<NodeRef> inputs,
                            std::vector<VarRef> vars,
                            std::vector<symbolic::Constraint> constraints,
                            std::unordered_map<int, VarRef> sym_var_map) {
  NodeRef idx = (NodeRef)nodes_.size();
  if (!constraints.empty()) {
    ASSERT(op == Operation::view) << "Can only specify constraints with views";
    for (const auto &c : constraints)
      for (const auto &e : {c.first, c.second})
        for (const auto &s : e.symbols())
          ASSERT(sym_var_map.count(s.id()))
              << "Unmapped constraint passed in: " << c.first.dump() << ": "
              << c.second.dump();
  }
  nodes_.emplace_back(Node(op, inputs, std::move(vars), std::move(constraints),
                           std::move(sym_var_map)));
  priorities_.emplace_back(0);
  orders_.emplace_back();
  reuse_disabled_.emplace_back();
  annotations_.emplace_back();
  loop_annotations_.emplace_back();
  reset_aux(idx);
  for (auto i : inputs) node(i).outputs_.emplace_back(idx);
  return idx;
}

const Node &IR::node(NodeRef ref) const {
  ASSERT(!deleted_.count(ref)) << "attempting to access deleted node";
  ASSERT(ref >= 0 && ref < (NodeRef)nodes_.size()) << "node ref '" << ref << "' not valid";
  return nodes_[ref];
}
Node &IR::node(NodeRef ref) {
  ASSERT(!deleted_.count(ref)) << "attempting to access deleted node";
  ASSERT(ref >= 0 && ref < (NodeRef)nodes_.size()) << "node ref '" << ref << "' not valid";
  return nodes_[ref];
}
const Var &IR::var(VarRef ref) const {
  ASSERT(ref >= 0 && ref < (VarRef)vars_.size()) << "var ref '" << ref << "' not valid";
  return vars_[ref];
}
Var &IR::var(VarRef ref) {
  ASSERT(ref >= 0 && ref < (VarRef)vars_.size()) << "var ref '" << ref << "' not valid";
  return vars_[ref];
}

void IR::delete_node(const NodeRef &node_ref) {
  replace_all_uses(node_ref, -1);
  if (node_ref == (NodeRef)nodes_.size() - 1) {
    nodes_.pop_back();
    priorities_.pop_back();
    orders_.pop_back();
    reuse_disabled_.pop_back();
    annotations_.pop_back();
    loop_annotations_.pop_back();
  } else {
    deleted_.insert(node_ref);
  }
}

void IR::reset_aux(NodeRef node_ref) {
  ASSERT(!deleted_.count(node_ref)) << "invalid node reference";
  priorities_[node_ref] = 0;
  reuse_disabled_[node_ref].clear();
}

void IR::replace_all_uses(NodeRef old_node, NodeRef new_node) {
  auto outs = node(old_node).outputs();
  for (auto o : outs) node(o).replace_input(old_node, new_node);
  if (new_node >= 0) node(new_node).update_outputs(outs);
  node(old_node).update_outputs({});
}

void IR::update_inputs(NodeRef node_ref, std::vector<NodeRef> inputs) {
  ASSERT(node(node_ref).inputs().empty()) << "node already has inputs";
  node(node_ref).update_inputs(inputs);
  for (auto i : inputs) {
    auto outs = node(i).outputs();
    outs.emplace_back(node_ref);
    node(i).update_outputs(outs);
  }
}

void IR::update_vars(NodeRef node_ref, std::vector<VarRef> vars) {
  node(node_ref).update_vars(std::move(vars));
}

std::string IR::dump(NodeRef idx) const {
  const auto &n = node(idx);
  std::stringstream ss;
  auto var_list = [&](const std::vector<VarRef> &vs) {
    for (size_t i = 0; i < vs.size(); ++i) {
      if (i) ss << ", ";
      ss << var(vs[i]).name();
    }
  };
  ss << "%" << idx << "[";
  var_list(n.vars());
  ss << "] <- ";
  bool is_view = n.op() == Operation::view;
  if (!is_view) ss << loop_tool::dump(n.op()) << "(";
  for (size_t i = 0; i < n.inputs().size(); ++i) {
    if (i) ss << ", ";
    auto inp = n.inputs()[i];
    ss << "%" << inp;
    if (!n.constraints().empty()) {
      ss << "[";
      var_list(node(inp).vars());
      ss << "]";
    }
  }
  if (!is_view) ss << ")";
  if (!n.constraints().empty()) {
    auto in_vars = to_set(node(n.inputs().at(0)).vars());
    ss << "{";
    bool first = true;
    for (const auto &c : n.constraints()) {
      if (c.first.type() != Expr::Type::symbol) continue;
      if (!n.has_sym(c.first.symbol())) continue;
      if (!in_vars.count(n.var(c.first.symbol()))) continue;
      if (!first) ss << ", ";
      first = false;
      ss << c.first.dump(true) << "=" << c.second.dump(true);
    }
    ss << "}";
  }
  return ss.str();
}

std::vector<IR::VarRef> IR::pointwise_vars(NodeRef idx) const {
  const auto &n = node(idx);
  std::unordered_set<VarRef> only_here(n.vars().begin(), n.vars().end());
  for (auto inp : n.inputs())
    for (auto v : node(inp).vars()) only_here.erase(v);
  std::vector<VarRef> out;
  for (auto v : n.vars())
    if (!only_here.count(v)) out.emplace_back(v);
  return out;
}

// A view "reduces" a set of its input vars when several of them map onto one
// output coordinate with unit-or-zero strides (or several non-unit strides
// remain).  Same decision procedure as src/core/ir.cpp:193-263.
std::vector<IR::VarRef> IR::view_reduction_vars(NodeRef idx) const {
  const auto &n = node(idx);
  std::unordered_set<VarRef> out_vars(n.vars().begin(), n.vars().end());
  std::vector<VarRef> result;
  auto is_out_constraint = [&](const symbolic::Constraint &c) {
    return c.first.type() == Expr::Type::symbol &&
           out_vars.count(n.var(c.first.symbol()));
  };
  std::vector<Symbol> rhs_syms;
  std::unordered_set<int> rhs_ids;
  for (const auto &c : n.constraints()) {
    if (!is_out_constraint(c)) continue;
    for (const auto &s : c.second.symbols()) {
      if (out_vars.count(n.var(s))) continue;
      if (rhs_ids.insert(s.id()).second) rhs_syms.emplace_back(s);
    }
  }
  for (const auto &c : n.constraints()) {
    if (!is_out_constraint(c)) continue;
    std::vector<std::pair<Symbol, Expr>> strides;
    for (const auto &s : rhs_syms)
      strides.emplace_back(s, symbolic::differentiate(c.second, s));
    for (const auto &s : rhs_syms) {
      std::vector<Symbol> contiguous;
      for (const auto &p : strides)
        if (p.second.can_evaluate() && p.second.evaluate() <= 1)
          contiguous.emplace_back(p.first);
      strides.erase(std::remove_if(strides.begin(), strides.end(),
                                   [&](const std::pair<Symbol, Expr> &p) {
                                     for (const auto &q : contiguous)
                                       if (q == p.first) return true;
                                     return false;
                                   }),
                    strides.end());
      for (auto &p : strides)
        if (p.second.contains(s))
          p.second = (p.second / Expr::size(Expr(s))).simplify();
      if (contiguous.size() > 1)
        for (const auto &q : contiguous) result.emplace_back(n.var(q));
    }
    if (strides.size() > 1)
      for (const auto &p : strides) result.emplace_back(n.var(p.first));
  }
  return result;
}

std::vector<IR::VarRef> IR::reduction_vars(NodeRef idx) const {
  const auto &n = node(idx);
  if (n.op() == Operation::view) return view_reduction_vars(idx);
  std::unordered_set<VarRef> seen(n.vars().begin(), n.vars().end());
  std::vector<VarRef> out;
  for (auto inp : n.inputs())
    for (auto v : node(inp).vars())
      if (seen.insert(v).second) out.emplace_back(v);
  return out;
}

std::vector<IR::VarRef> IR::all_vars(NodeRef node_ref) const {
  const auto &n = node(node_ref);
  std::unordered_set<VarRef> seen(n.vars().begin(), n.vars().end());
  std::vector<VarRef> out = n.vars();
  for (auto inp : n.inputs())
    for (auto v : node(inp).vars())
      if (seen.insert(v).second) out.emplace_back(v);
  std::sort(out.begin(), out.end());
  return out;
}

std::vector<IR::VarRef> IR::loop_vars(NodeRef node_ref) const {
  ASSERT(node(node_ref).op() != Operation::view)
      << "loop vars are undefined with views";
  return all_vars(node_ref);
}

std::vector<IR::VarRef> IR::vars() const {
  std::vector<VarRef> out(vars_.size());
  for (size_t i = 0; i < vars_.size(); ++i) out[i] = (VarRef)i;
  return out;
}

std::vector<IR::NodeRef> IR::nodes() const {
  std::vector<NodeRef> out;
  for (size_t i = 0; i < nodes_.size(); ++i)
    if (!deleted_.count((NodeRef)i)) out.emplace_back((NodeRef)i);
  return out;
}

void IR::reify_deletions() {
  std::unordered_map<NodeRef, NodeRef> remap;
  NodeRef next = 0;
  for (auto nr : nodes()) remap[nr] = next++;
  for (auto nr : nodes()) node(nr).remap_refs(remap);
  for (auto &nr : inputs_)
    if (remap.count(nr)) nr = remap.at(nr);
  for (auto &nr : outputs_)
    if (remap.count(nr)) nr = remap.at(nr);
  auto keep = [&](auto &vec) {
    using V = std::decay_t<decltype(vec)>;
    V out;
    for (size_t i = 0; i < vec.size(); ++i)
      if (!deleted_.count((NodeRef)i)) out.emplace_back(vec[i]);
    vec = std::move(out);
  };
  keep(nodes_);
  keep(priorities_);
  keep(orders_);
  keep(reuse_disabled_);
  keep(annotations_);
  keep(loop_annotations_);
  deleted_.clear();
}

void Node::replace_input(IR::NodeRef old_node, IR::NodeRef new_node) {
  for (auto &n : inputs_)
    if (n == old_node) n = new_node;
}

void Node::remap_refs(const std::unordered_map<IR::NodeRef, IR::NodeRef> &remap) {
  for (auto &nr : inputs_)
    if (remap.count(nr)) nr = remap.at(nr);
  for (auto &nr : outputs_)
    if (remap.count(nr)) nr = remap.at(nr);
}

// Kahn's algorithm from the IR inputs; among ready nodes the highest
// priority goes first, ties keep discovery order (stable).
std::vector<IR::NodeRef> toposort(const IR &ir) {
  std::vector<IR::NodeRef> sorted;
  std::vector<std::pair<IR::NodeRef, float>> frontier;
  for (auto i : ir.inputs()) frontier.emplace_back(i, ir.priority(i));
  std::unordered_set<IR::NodeRef> done;
  while (!frontier.empty()) {
    std::stable_sort(frontier.begin(), frontier.end(),
                     [](const std::pair<IR::NodeRef, float> &a,
                        const std::pair<IR::NodeRef, float> &b) {
                       return a.second > b.second;
                     });
    auto cur = frontier.front().first;
    frontier.erase(frontier.begin());
    sorted.emplace_back(cur);
    done.insert(cur);
    // each consumer considered once per producer, in hash-set order
    for (auto dep : to_set(ir.node(cur).outputs())) {
      const auto &ins = ir.node(dep).inputs();
      if (std::all_of(ins.begin(), ins.end(),
                      [&](IR::NodeRef i) { return done.count(i) > 0; }))
        frontier.emplace_back(dep, ir.priority(dep));
    }
  }
  return sorted;
}

// ------------------------------------------------------------- LoopTree ----
std::unordered_set<IR::VarRef> LoopTree::scope_vars(TreeRef ref) const {
  std::unordered_set<IR::VarRef> out;
  while (ref != -1) {
    if (kind(ref) == LOOP) out.insert(loop(ref).var);
    ref = parent(ref);
  }
  return out;
}

LoopTree::TreeRef LoopTree::add_leaf(TreeRef parent, IR::NodeRef n) {
  TreeRef idx = (TreeRef)nodes.size();
  nodes.emplace_back(parent, idx, n);
  if (parent == -1) {
    roots.emplace_back(idx);
  } else {
    nodes[idx].depth = nodes[parent].depth + 1;
    nodes[parent].children.emplace_back(idx);
  }
  scheduled[n] = idx;
  return idx;
}

LoopTree::TreeRef LoopTree::add_loop(TreeRef parent, const Loop &l) {
  TreeRef idx = (TreeRef)nodes.size();
  nodes.emplace_back(parent, idx, l);
  if (parent == -1) {
    roots.emplace_back(idx);
  } else {
    nodes[idx].depth = nodes[parent].depth + 1;
    nodes[parent].children.emplace_back(idx);
  }
  return idx;
}

void LoopTree::annotate_(TreeRef ref, std::string annot) {
  auto it = std::find(annotations.begin(), annotations.end(), annot);
  if (it != annotations.end()) {
    tree_node(ref).annotation = (int)(it - annotations.begin());
    return;
  }
  tree_node(ref).annotation = (int)annotations.size();
  annotations.emplace_back(std::move(annot));
}

LoopTree::TreeRef LoopTree::lca(TreeRef a, TreeRef b) const {
  if (a == -1 || b == -1) return -1;
  while (depth(a) > depth(b)) a = parent(a);
  while (depth(b) > depth(a)) b = parent(b);
  while (a != b) {
    a = parent(a);
    b = parent(b);
  }
  return a;
}

std::vector<LoopTree::Loop> LoopTree::loop_order(IR::NodeRef ref) const {
  std::vector<Loop> out;
  std::unordered_map<IR::VarRef, int> depth_of;
  for (const auto &p : ir.order(ref)) {
    out.push_back(Loop{p.first, depth_of[p.first]++, p.second.size, p.second.tail});
  }
  return out;
}

void LoopTree::walk(const std::function<void(LoopTree::TreeRef, int)> &fn,
                    TreeRef start) const {
  std::function<void(TreeRef, int)> rec = [&](TreeRef t, int d) {
    fn(t, d);
    for (auto c : tree_node(t).children) rec(c, d + 1);
  };
  if (start == -1) {
    for (auto r : roots) rec(r, 0);
  } else {
    rec(start, 0);
  }
}

std::string LoopTree::dump(
    const std::function<std::string(LoopTree::TreeRef)> &fn) const {
  std::stringstream ss;
  walk([&](TreeRef t, int d) {
    ss << std::string(d, ' ');
    const auto &tn = tree_node(t);
    if (tn.kind == NODE) {
      ss << ir.dump(tn.node);
    } else {
      ss << "for " << ir.var(tn.loop.var).name()
         << std::string(tn.loop.var_depth, '\'');
      if (tn.loop.size > -1) ss << " in " << tn.loop.size;
      if (tn.loop.tail > 0) ss << " r " << tn.loop.tail;
      ss << " : L" << t;
    }
    if (tn.annotation > -1) ss << " " << annotations[tn.annotation];
    if (fn) ss << " " << fn(t);
    ss << "\n";
  });
  return ss.str();
}

namespace {

// base var of every var reachable through views: a view that reduces some of
// its input vars ties the remaining vars of each constraint to that base.
std::unordered_map<IR::VarRef, IR::VarRef> view_bases(
    const IR &ir, const std::vector<IR::NodeRef> &sorted) {
  std::unordered_map<IR::VarRef, IR::VarRef> base;
  for (auto nr : sorted) {
    const auto &n = ir.node(nr);
    if (n.op() != Operation::view) {
      for (auto v : n.vars())
        if (!base.count(v)) base[v] = v;
      continue;
    }
    auto reduced = to_set(ir.reduction_vars(nr));
    for (const auto &c : n.constraints()) {
      std::vector<Symbol> syms = c.first.symbols();
      for (const auto &s : c.second.symbols()) {
        bool dup = false;
        for (const auto &t : syms) dup |= (t == s);
        if (!dup) syms.emplace_back(s);
      }
      IR::VarRef base_var = -1;
      std::vector<IR::VarRef> viewed;
      for (const auto &s : syms) {
        auto v = n.var(s);
        if (reduced.count(v)) {
          ASSERT(base.count(v)) << "view base unknown for " << ir.var(v).name();
          if (base_var != -1) {
            ASSERT(base.at(v) == base_var)
                << "found impossible lowering condition in view! "
                << ir.var(v).name() << " and " << ir.var(base_var).name()
                << " are both base variables to a view in node " << ir.dump(nr);
          }
          base_var = base.at(v);
          continue;
        }
        viewed.emplace_back(v);
      }
      if (base_var == -1) continue;
      for (auto v : viewed) base[v] = base_var;
    }
  }
  return base;
}

}  // namespace

// Greedy loop reuse: nodes are placed in priority-topological order; each
// node reuses the longest compatible chain of currently open loops and opens
// the rest of its order beneath it (ir.cpp:531-728 semantics).
LoopTree::LoopTree(const IR &ir_) : ir(ir_) {
  struct Open {
    Loop loop;
    TreeRef ref;
    std::string annot;
  };
  std::vector<Open> open;  // the current chain of reusable loops, outer first
  auto sorted = toposort(ir);
  auto base = view_bases(ir, sorted);
  auto base_of = [&](IR::VarRef v) {
    auto it = base.find(v);
    ASSERT(it != base.end()) << "no view base for var " << ir.var(v).name();
    return it->second;
  };
  auto truncate_at_first = [&](const std::function<bool(const Open &)> &pred) {
    auto it = std::find_if(open.begin(), open.end(), pred);
    open.erase(it, open.end());
  };

  for (auto nr : sorted) {
    const auto &n = ir.node(nr);
    auto order = loop_order(nr);
    auto annots = ir.loop_annotations(nr);
    if (!n.vars().empty() && order.empty()) continue;  // unscheduled

    // 1. a loop over the same var at the same depth but another split ends reuse
    for (const auto &l : order)
      truncate_at_first([&](const Open &o) {
        return o.loop.var == l.var && o.loop.var_depth == l.var_depth &&
               !(o.loop == l);
      });
    // 2. a loop over a view-related var that this node does not schedule
    std::unordered_set<IR::VarRef> sched_vars;
    for (const auto &l : order) sched_vars.insert(l.var);
    for (const auto &l : order)
      truncate_at_first([&](const Open &o) {
        return base_of(o.loop.var) == base_of(l.var) && !sched_vars.count(o.loop.var);
      });
    // 3. differing annotations end reuse
    for (size_t i = 0; i < order.size(); ++i)
      truncate_at_first([&](const Open &o) {
        return o.loop.var == order[i].var &&
               o.loop.var_depth == order[i].var_depth && o.annot != annots[i];
      });

    // 4. longest prefix of the order found in increasing positions of `open`
    std::vector<std::pair<int, int>> cand;  // (order position, open position)
    for (size_t i = 0; i < order.size(); ++i) {
      int pos = (int)open.size();
      for (size_t j = 0; j < open.size(); ++j)
        if (open[j].loop == order[i]) {
          pos = (int)j;
          break;
        }
      cand.emplace_back((int)i, pos);
    }
    std::stable_sort(cand.begin(), cand.end(),
                     [](const std::pair<int, int> &a, const std::pair<int, int> &b) {
                       return a.second < b.second;
                     });
    size_t reuse_end = 0;  // open[0, reuse_end) is reused
    size_t first_new = 0;
    for (size_t i = 0; i < order.size(); ++i) {
      if (cand[i].first != (int)i || cand[i].second == (int)open.size()) break;
      reuse_end = cand[i].second + 1;
      first_new++;
    }

    for (size_t i = first_new; i < order.size(); ++i) {
      TreeRef parent_ref = reuse_end > 0 ? open[reuse_end - 1].ref : -1;
      TreeRef loc = add_loop(parent_ref, order[i]);
      annotate_(loc, annots[i]);
      open.erase(open.begin() + reuse_end, open.end());
      open.push_back(Open{order[i], loc, annots[i]});
      reuse_end = open.size();
    }
    TreeRef leaf_parent = open.empty() ? -1 : open.back().ref;
    auto leaf = add_leaf(leaf_parent, nr);
    annotate_(leaf, ir.annotation(nr));

    // 5. loops over vars this node reduces cannot be shared with later nodes
    std::unordered_set<IR::VarRef> reduced;
    for (auto inp : n.inputs())
      for (auto v : ir.node(inp).vars()) reduced.insert(v);
    for (auto v : n.vars()) reduced.erase(v);
    truncate_at_first([&](const Open &o) { return reduced.count(o.loop.var) > 0; });

    // 6. explicit reuse barriers
    for (auto idx : ir.not_reusable(nr)) {
      ASSERT(idx >= 0 && idx < (int)order.size()) << "reuse index out of range";
      truncate_at_first([&](const Open &o) { return o.loop == order[idx]; });
    }
  }
}

std::string dot(const IR &ir) {
  std::stringstream ss;
  ss << "digraph G {\n node [fontname = \"courier\", fontsize=12];\n";
  ss << " { rank=sink; vars[shape=record,label=\"";
  auto vs = ir.vars();
  for (size_t i = 0; i < vs.size(); ++i) {
    if (i) ss << "|";
    ss << "<" << vs[i] << ">" << ir.var(vs[i]).name();
  }
  ss << "\"]; }\n";
  for (auto n : toposort(ir)) {
    ss << " " << n << "[shape=record,label=\"{" << dump(ir.node(n).op()) << " : [";
    const auto &nv = ir.node(n).vars();
    for (size_t i = 0; i < nv.size(); ++i) {
      if (i) ss << ", ";
      ss << ir.var(nv[i]).name();
    }
    ss << "]|{";
    const auto &order = ir.order(n);
    for (size_t i = 0; i < order.size(); ++i) {
      if (i) ss << "|";
      ss << "<" << i << ">" << ir.var(order[i].first).name();
      if (order[i].second.size > 0) ss << ":" << order[i].second.size;
      if (order[i].second.tail > 0) ss << "r" << order[i].second.tail;
    }
    ss << "}}\"];\n";
    for (auto o : ir.node(n).outputs()) ss << " " << n << " -> " << o << ";\n";
    for (size_t i = 0; i < order.size(); ++i)
      ss << " \"vars\":" << order[i].first << " -> \"" << n << "\":" << i
         << "[style=dotted,arrowhead=none,weight=0];\n";
  }
  ss << "}\n";
  return ss.str();
}

}  // namespace loop_tool
