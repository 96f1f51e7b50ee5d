// Schedule transforms (see include/loop_tool/mutate.h).
// Reference behaviour: src/frontends/mutate.cpp (split_node :14-31, subtree
// :95-200, split :226-272, merge :274-335, swaps :350-481, reuse :483-551,
// navigation :553-620, flops :622-657, is_trivially_parallel :659-692,
// annotate :694-770, map_ref :772-815, maximize_reuse :817-834,
// unroll_inner_loops :836-863).
#include "loop_tool/mutate.h"

#include <algorithm>
#include <iostream>

namespace loop_tool {

using Order = std::vector<std::pair<IR::VarRef, IR::LoopSize>>;

IR split_node(const IR &ir, IR::NodeRef node_ref, std::vector<IR::VarRef> injected) {
  IR out = ir;
  auto loop_vars = to_set(out.loop_vars(node_ref));
  for (auto v : injected) {
    ASSERT(loop_vars.count(v)) << "injected var not a loop var of the node";
    loop_vars.erase(v);
  }
  ASSERT(!loop_vars.empty()) << "nothing left to split off";
  auto op = out.node(node_ref).op();
  auto vars = out.node(node_ref).vars();
  auto second = out.create_node(op, {}, vars);
  out.replace_all_uses(node_ref, second);
  out.update_vars(node_ref, injected);
  out.update_inputs(second, {node_ref});
  out.reset_aux(node_ref);
  out.reset_aux(second);
  return out;
}

IR split_var(const IR &ir, IR::VarRef v) {
  ASSERT(0) << "split_var is not implemented (as upstream)";
  return ir;
}

IR swap_vars(const IR &ir, IR::NodeRef node_ref, IR::VarRef a, IR::VarRef b) {
  IR out = ir;
  ASSERT(a != b) << "cannot swap var with itself";
  auto &vars = out.node(node_ref).vars();
  auto ia = std::find(vars.begin(), vars.end(), a);
  auto ib = std::find(vars.begin(), vars.end(), b);
  ASSERT(ia != vars.end()) << "cannot find var " << out.var(a).name() << " in node "
                           << out.dump(node_ref);
  ASSERT(ib != vars.end()) << "cannot find var " << out.var(b).name() << " in node "
                           << out.dump(node_ref);
  std::iter_swap(ia, ib);
  return out;
}

namespace {

std::vector<IR::NodeRef> nodes_under(const LoopTree &lt, LoopTree::TreeRef ref) {
  ASSERT(lt.kind(ref) == LoopTree::LOOP) << "can only collect nodes within loops";
  std::vector<IR::NodeRef> out;
  lt.walk(
      [&](LoopTree::TreeRef t, int) {
        if (lt.kind(t) == LoopTree::NODE) out.emplace_back(lt.node(t));
      },
      ref);
  return out;
}

// extent of the loop's var covered by the loops nested inside it
int64_t inner_extent(const LoopTree &lt, LoopTree::TreeRef ref) {
  auto var = lt.loop(ref).var;
  int64_t best = 0;
  lt.walk(
      [&](LoopTree::TreeRef t, int) {
        if (lt.kind(t) == LoopTree::LOOP) return;
        int64_t cur = 1;
        for (auto p = lt.parent(t); p != ref; p = lt.parent(p)) {
          auto l = lt.loop(p);
          if (l.var == var) cur = cur * l.size + l.tail;
        }
        best = std::max(best, cur);
      },
      ref);
  return best;
}

}  // namespace

LoopTree subtree(const LoopTree &lt, LoopTree::TreeRef ref,
                 std::unordered_map<IR::NodeRef, IR::NodeRef> node_map,
                 std::unordered_map<IR::VarRef, IR::VarRef> var_map) {
  std::unordered_set<IR::NodeRef> keep;
  if (lt.kind(ref) == LoopTree::NODE) {
    keep.insert(lt.node(ref));
  } else {
    keep = to_set(nodes_under(lt, ref));
  }
  IR sub;
  auto all = lt.ir.nodes();
  for (auto nr : all) {
    const auto &n = lt.ir.node(nr);
    if (!lt.scheduled.count(nr)) {
      bool only_kept_users = true;
      for (auto o : n.outputs()) only_kept_users &= keep.count(o) > 0;
      if (only_kept_users) keep.insert(nr);
    }
    if (!keep.count(nr)) continue;
    std::vector<IR::NodeRef> inputs;
    for (auto i : n.inputs())
      if (keep.count(i)) inputs.emplace_back(node_map.at(i));
    std::vector<IR::VarRef> vars;
    for (auto v : n.vars()) {
      if (!var_map.count(v)) var_map[v] = sub.create_var(lt.ir.var(v).name());
      vars.emplace_back(var_map.at(v));
    }
    if (inputs.empty() && n.op() != Operation::read) {
      auto r = sub.create_node(Operation::read, {}, vars);
      inputs.emplace_back(r);
      sub.add_input(r);
    }
    std::unordered_map<int, IR::VarRef> sym_var_map;
    for (const auto &p : n.sym_to_var()) {
      if (!var_map.count(p.second))
        var_map[p.second] = sub.create_var(lt.ir.var(p.second).name());
      sym_var_map[p.first] = var_map.at(p.second);
    }
    node_map[nr] = sub.create_node(n.op(), inputs, vars, n.constraints(), sym_var_map);
    if (n.op() == Operation::read) sub.add_input(node_map[nr]);
    bool needs_write = true;
    if (n.op() == Operation::write) {
      sub.add_output(node_map[nr]);
      needs_write = false;
    }
    for (auto o : n.outputs())
      if (keep.count(o)) needs_write = false;
    if (needs_write) sub.add_output(sub.create_node(Operation::write, {node_map[nr]}, vars));
  }
  ASSERT(!sub.outputs().empty()) << "subtree has no outputs";
  ASSERT(!sub.inputs().empty()) << "subtree has no inputs";
  for (auto nr : all) {
    if (!keep.count(nr)) continue;
    Order order;
    std::vector<std::string> annots;
    const auto &old_order = lt.ir.order(nr);
    const auto &old_annots = lt.ir.loop_annotations(nr);
    for (size_t i = 0; i < old_order.size(); ++i) {
      if (!var_map.count(old_order[i].first)) continue;
      order.emplace_back(var_map.at(old_order[i].first), old_order[i].second);
      annots.emplace_back(old_annots.at(i));
    }
    auto m = node_map.at(nr);
    sub.set_order(m, order, annots);
    sub.annotate(m, lt.ir.annotation(nr));
    for (auto i : lt.ir.not_reusable(nr)) sub.disable_reuse(m, i);
  }
  return LoopTree(sub);
}

LoopTree split(const LoopTree &lt, LoopTree::TreeRef ref, int64_t size) {
  ASSERT(lt.kind(ref) == LoopTree::LOOP) << "can only split loops";
  auto loop = lt.loop(ref);
  ASSERT(size > 0 && loop.size / size > 0)
      << "attempting to split a loop of size " << loop.size << " by " << size;
  auto inner = inner_extent(lt, ref);
  auto total = loop.size * inner + loop.tail;
  auto outer_size = total / (size * inner);
  auto outer_tail = total % (size * inner);
  ASSERT(outer_size > 0) << "invalid split size";

  IR ir = lt.ir;
  for (auto nr : nodes_under(lt, ref)) {
    auto lo = lt.loop_order(nr);
    for (size_t i = 0; i < lo.size(); ++i) {
      if (!(lo[i] == loop)) continue;
      const auto &old_order = lt.ir.order(nr);
      const auto &old_annot = lt.ir.loop_annotations(nr);
      Order order;
      std::vector<std::string> annots;
      for (size_t j = 0; j < old_order.size(); ++j) {
        if (j == i) {
          order.emplace_back(loop.var, IR::LoopSize{outer_size, outer_tail});
          order.emplace_back(loop.var, IR::LoopSize{size, 0});
          annots.emplace_back();
          annots.emplace_back();
        } else {
          order.emplace_back(old_order[j]);
          annots.emplace_back(old_annot[j]);
        }
      }
      ir.set_order(nr, order, annots);
    }
  }
  return LoopTree(ir);
}

LoopTree merge(const LoopTree &lt, LoopTree::TreeRef ref) {
  ASSERT(lt.kind(ref) == LoopTree::LOOP) << "can only merge loops";
  auto loop = lt.loop(ref);
  auto p = lt.parent(ref);
  while (p != -1 && lt.loop(p).var != loop.var) p = lt.parent(p);
  if (p == -1) return lt;
  auto outer = lt.loop(p);
  auto inner = inner_extent(lt, ref);
  auto total = outer.size * (loop.size * inner + loop.tail) + outer.tail;
  IR::LoopSize merged{total / inner, total % inner};

  IR ir = lt.ir;
  for (auto nr : nodes_under(lt, p)) {
    auto lo = lt.loop_order(nr);
    int mi = -1, di = -1;
    for (size_t i = 0; i < lo.size(); ++i) {
      if (lo[i] == outer)
        mi = (int)i;
      else if (lo[i] == loop)
        di = (int)i;
    }
    if (mi < 0 || di < 0) continue;
    const auto &old_order = lt.ir.order(nr);
    const auto &old_annot = lt.ir.loop_annotations(nr);
    Order order;
    std::vector<std::string> annots;
    for (size_t i = 0; i < old_order.size(); ++i) {
      if ((int)i == mi) {
        order.emplace_back(loop.var, merged);
        annots.emplace_back();
      } else if ((int)i != di) {
        order.emplace_back(old_order[i]);
        annots.emplace_back(old_annot[i]);
      }
    }
    ir.set_order(nr, order, annots);
  }
  return LoopTree(ir);
}

LoopTree copy_input(const LoopTree &lt, LoopTree::TreeRef ref, int idx) {
  ASSERT(lt.kind(ref) == LoopTree::NODE) << "copy_input needs a node";
  ASSERT(idx >= 0) << "cannot use negatively indexed input";
  auto nr = lt.node(ref);
  IR ir = lt.ir;
  ASSERT((int)ir.node(nr).inputs().size() > idx) << "cannot get input at index " << idx;
  auto inp = ir.node(nr).inputs().at(idx);
  auto copy = ir.create_node(Operation::copy, {}, ir.node(inp).vars());
  ir.set_order(copy, ir.order(nr), ir.loop_annotations(nr));
  ir.replace_all_uses(inp, copy);
  ir.update_inputs(copy, {inp});
  return LoopTree(ir);
}

LoopTree delete_copy(const LoopTree &lt, LoopTree::TreeRef ref) {
  ASSERT(lt.kind(ref) == LoopTree::NODE) << "delete_copy needs a node";
  auto nr = lt.node(ref);
  IR ir = lt.ir;
  ASSERT(ir.node(nr).op() == Operation::copy) << "not a copy";
  auto inp = ir.node(nr).inputs().at(0);
  ir.replace_all_uses(nr, inp);
  ir.delete_node(nr);
  return LoopTree(ir);
}

LoopTree swap_nodes(const LoopTree &lt, LoopTree::TreeRef a, LoopTree::TreeRef b) {
  ASSERT(lt.kind(a) == LoopTree::NODE && lt.kind(b) == LoopTree::NODE)
      << "swap_nodes needs two nodes";
  auto na = lt.node(a), nb = lt.node(b);
  IR ir = lt.ir;
  auto pa = ir.priority(na), pb = ir.priority(nb);
  if (pa == pb) pa = pb + 0.1f;
  ir.set_priority(na, pb);
  ir.set_priority(nb, pa);
  return LoopTree(ir);
}

LoopTree swap_vars(const LoopTree &lt, IR::NodeRef node_ref, IR::VarRef a, IR::VarRef b) {
  return LoopTree(swap_vars(lt.ir, node_ref, a, b));
}

LoopTree remove_loop(const LoopTree &lt, LoopTree::TreeRef ref, LoopTree::TreeRef rem) {
  ASSERT(lt.kind(ref) == LoopTree::NODE) << "remove_loop needs a node";
  ASSERT(lt.kind(rem) == LoopTree::LOOP) << "remove_loop needs a loop";
  auto nr = lt.node(ref);
  auto loop = lt.loop(rem);
  if (lt.ir.node(nr).op() != Operation::view) {
    ASSERT(!to_set(lt.ir.loop_vars(nr)).count(loop.var))
        << "attempting to deschedule a necessary loop";
  }
  auto lo = lt.loop_order(nr);
  int idx = -1;
  for (size_t i = 0; i < lo.size(); ++i)
    if (lo[i] == loop) idx = (int)i;
  IR ir = lt.ir;
  if (idx >= 0) {
    auto order = ir.order(nr);
    order.erase(order.begin() + idx);
    ir.set_order(nr, order);
  }
  return LoopTree(ir);
}

LoopTree add_loop(const LoopTree &lt, LoopTree::TreeRef ref, LoopTree::TreeRef add) {
  ASSERT(lt.kind(ref) == LoopTree::NODE) << "add_loop needs a node";
  ASSERT(lt.kind(add) == LoopTree::LOOP) << "add_loop needs a loop";
  auto nr = lt.node(ref);
  auto loop = lt.loop(add);
  auto annot = lt.annotation(add);
  ASSERT(lt.ir.reduction_vars(nr).empty())
      << "cannot add orthogonal inner loop to reduction";
  IR ir = lt.ir;
  auto order = ir.order(nr);
  order.emplace_back(loop.var, IR::LoopSize{loop.size, loop.tail});
  ir.set_order(nr, order);
  LoopTree new_lt(ir);
  auto new_ref = map_ref(new_lt, ref, lt);
  return annotate(new_lt, new_ref, annot);
}

LoopTree swap_loops(const LoopTree &lt, LoopTree::TreeRef a, LoopTree::TreeRef b) {
  ASSERT(lt.kind(a) == LoopTree::LOOP && lt.kind(b) == LoopTree::LOOP)
      << "swap_loops needs two loops";
  auto contains = [&](LoopTree::TreeRef outer, LoopTree::TreeRef inner) {
    bool found = false;
    lt.walk([&](LoopTree::TreeRef t, int) { found |= (t == inner); }, outer);
    return found;
  };
  bool a_outer = contains(a, b), b_outer = contains(b, a);
  ASSERT(a_outer ^ b_outer) << "loops must be nested to be swapped";
  if (b_outer) std::swap(a, b);
  auto la = lt.loop(a), lb = lt.loop(b);
  ASSERT(la.var != lb.var) << "swapping the same var is not supported, resplit instead";
  IR ir = lt.ir;
  for (auto nr : nodes_under(lt, a)) {
    auto lo = lt.loop_order(nr);
    int ia = -1, ib = -1;
    for (size_t i = 0; i < lo.size(); ++i) {
      if (lo[i] == la) ia = (int)i;
      if (lo[i] == lb) ib = (int)i;
    }
    if (ia < 0 || ib < 0) continue;
    auto order = lt.ir.order(nr);
    auto annots = lt.ir.loop_annotations(nr);
    std::swap(order[ia], order[ib]);
    std::swap(annots[ia], annots[ib]);
    ir.set_order(nr, order, annots);
  }
  return LoopTree(ir);
}

LoopTree try_swap(const LoopTree &lt, LoopTree::TreeRef a, LoopTree::TreeRef b) {
  bool al = lt.kind(a) == LoopTree::LOOP, bl = lt.kind(b) == LoopTree::LOOP;
  if (!al && !bl) return swap_nodes(lt, a, b);
  if (al && bl) {
    if (lt.loop(a).var != lt.loop(b).var) return swap_loops(lt, a, b);
  } else if (!al && bl) {
    if (lt.parent(a) == b) return remove_loop(lt, a, b);
    return add_loop(lt, a, b);
  }
  return lt;
}

namespace {
LoopTree toggle_reuse(const LoopTree &lt, LoopTree::TreeRef ref, IR::NodeRef n,
                      bool enable) {
  IR ir = lt.ir;
  auto lo = lt.loop_order(n);
  for (size_t i = 0; i < lo.size(); ++i) {
    if (lo[i] == lt.loop(ref)) {
      if (enable)
        ir.enable_reuse(n, (int)i);
      else
        ir.disable_reuse(n, (int)i);
      break;
    }
  }
  return LoopTree(ir);
}
}  // namespace

LoopTree disable_reuse(const LoopTree &lt, LoopTree::TreeRef loop, IR::NodeRef n) {
  return toggle_reuse(lt, loop, n, false);
}

LoopTree enable_reuse(const LoopTree &lt, LoopTree::TreeRef loop, IR::NodeRef n) {
  return toggle_reuse(lt, loop, n, true);
}

LoopTree decrease_reuse(const LoopTree &lt, LoopTree::TreeRef ref) {
  ASSERT(lt.kind(ref) == LoopTree::NODE) << "decrease_reuse needs a node";
  auto nr = lt.node(ref);
  IR ir = lt.ir;
  auto nr_set = ir.not_reusable(nr);
  const auto &order = ir.order(nr);
  if (!nr_set.empty()) {
    for (size_t i = 0; i + 1 < order.size(); ++i) {
      if (nr_set.count((int)i + 1)) {
        ir.disable_reuse(nr, (int)i);
        break;
      }
    }
  } else if (!order.empty()) {
    ir.disable_reuse(nr, (int)order.size() - 1);
  }
  return LoopTree(ir);
}

LoopTree increase_reuse(const LoopTree &lt, LoopTree::TreeRef ref) {
  ASSERT(lt.kind(ref) == LoopTree::NODE) << "increase_reuse needs a node";
  auto nr = lt.node(ref);
  IR ir = lt.ir;
  auto nr_set = ir.not_reusable(nr);
  for (size_t i = 0; i < ir.order(nr).size(); ++i) {
    if (nr_set.count((int)i)) {
      ir.enable_reuse(nr, (int)i);
      break;
    }
  }
  return LoopTree(ir);
}

namespace {
LoopTree::TreeRef next_impl(const LoopTree &lt, LoopTree::TreeRef ref, bool descend) {
  if (ref == -1) return -1;
  const auto &kids = lt.children(ref);
  if (descend && !kids.empty()) return kids[0];
  const auto &sib = lt.children(lt.parent(ref));
  auto it = std::find(sib.begin(), sib.end(), ref);
  if (it != sib.end() && it + 1 != sib.end()) return *(it + 1);
  return next_impl(lt, lt.parent(ref), false);
}
}  // namespace

LoopTree::TreeRef next_ref(const LoopTree &lt, LoopTree::TreeRef ref) {
  auto n = next_impl(lt, ref, true);
  return n == -1 ? ref : n;
}

LoopTree::TreeRef previous_ref(const LoopTree &lt, LoopTree::TreeRef ref) {
  if (ref == -1) return ref;
  const auto &sib = lt.children(lt.parent(ref));
  auto it = std::find(sib.begin(), sib.end(), ref);
  if (it == sib.begin()) {
    auto p = lt.parent(ref);
    return p == -1 ? ref : p;
  }
  // last node visited (pre-order) before `ref`
  auto cur = *(it - 1);
  auto trailing = cur;
  while (cur != ref) {
    trailing = cur;
    cur = next_ref(lt, cur);
  }
  return trailing;
}

LoopTree annotate(const LoopTree &lt, LoopTree::TreeRef ref, std::string annot) {
  return annotate(lt, std::unordered_map<LoopTree::TreeRef, std::string>{{ref, annot}});
}

LoopTree annotate(const LoopTree &lt,
                  std::unordered_map<LoopTree::TreeRef, std::string> annotations) {
  IR ir = lt.ir;
  for (const auto &p : annotations) {
    if (lt.kind(p.first) == LoopTree::NODE) {
      ir.annotate(lt.node(p.first), p.second);
      continue;
    }
    auto loop = lt.loop(p.first);
    for (auto nr : nodes_under(lt, p.first)) {
      auto lo = lt.loop_order(nr);
      for (size_t i = 0; i < lo.size(); ++i)
        if (lo[i] == loop) ir.annotate_loop(nr, (int)i, p.second);
    }
  }
  return LoopTree(ir);
}

LoopTree::TreeRef map_ref(const LoopTree &new_lt, LoopTree::TreeRef old_ref,
                          const LoopTree &old_lt) {
  if (old_lt.kind(old_ref) == LoopTree::NODE) {
    auto nr = old_lt.node(old_ref);
    if (new_lt.scheduled.count(nr)) return new_lt.scheduled.at(nr);
    return new_lt.roots.at(0);
  }
  auto loop = old_lt.loop(old_ref);
  // which occurrence (pre-order) of this var/depth the old ref was
  int version = 0, count = 0;
  old_lt.walk([&](LoopTree::TreeRef t, int) {
    if (old_lt.kind(t) != LoopTree::LOOP) return;
    if (t == old_ref) version = count;
    auto l = old_lt.loop(t);
    if (l.var == loop.var && l.var_depth == loop.var_depth) count++;
  });
  std::vector<LoopTree::TreeRef> same, shallower;
  new_lt.walk([&](LoopTree::TreeRef t, int) {
    if (new_lt.kind(t) != LoopTree::LOOP) return;
    auto l = new_lt.loop(t);
    if (l.var != loop.var) return;
    if (l.var_depth == loop.var_depth) same.emplace_back(t);
    if (l.var_depth + 1 == loop.var_depth) shallower.emplace_back(t);
  });
  if (same.empty()) {
    if (shallower.empty()) return new_lt.roots.at(0);
    return shallower.at(std::min(version, (int)shallower.size() - 1));
  }
  return same.at(std::min(version, (int)same.size() - 1));
}

LoopTree maximize_reuse(const LoopTree &lt) {
  IR ir = lt.ir;
  // pointwise vars weigh up, reduced vars weigh down: reductions go inward
  std::unordered_map<IR::VarRef, float> weight;
  for (auto nr : ir.nodes()) {
    auto pw = to_set(ir.pointwise_vars(nr));
    auto red = to_set(ir.reduction_vars(nr));
    for (auto v : ir.node(nr).vars()) {
      if (pw.count(v))
        weight[v] += 1;
      else if (red.count(v))
        weight[v] -= 1;
    }
  }
  for (auto nr : ir.nodes()) {
    auto order = ir.order(nr);
    std::stable_sort(order.begin(), order.end(),
                     [&](const std::pair<IR::VarRef, IR::LoopSize> &a,
                         const std::pair<IR::VarRef, IR::LoopSize> &b) {
                       return weight[a.first] > weight[b.first];
                     });
    ir.set_order(nr, order);
  }
  return LoopTree(ir);
}

LoopTree unroll_inner_loops(const LoopTree &lt, int32_t unroll_amount) {
  std::unordered_map<LoopTree::TreeRef, std::string> annots;
  std::vector<LoopTree::TreeRef> stop_at;
  for (auto nr : lt.ir.nodes()) {
    if (!lt.scheduled.count(nr)) continue;
    auto p = lt.parent(lt.scheduled.at(nr));
    int64_t size = p != -1 ? lt.loop(p).size : 1;
    while (p != -1 && size < unroll_amount) {
      annots[p] = "unroll";
      p = lt.parent(p);
      size *= p != -1 ? lt.loop(p).size : 1;
    }
    stop_at.emplace_back(p);
  }
  for (auto p : stop_at)
    for (; p != -1; p = lt.parent(p)) annots.erase(p);
  return annotate(lt, annots);
}

int64_t flops(const LoopTree &lt) {
  int64_t total = 0;
  std::vector<LoopTree::Loop> running;
  lt.walk([&](LoopTree::TreeRef ref, int depth) {
    running.resize(depth);
    if (lt.kind(ref) == LoopTree::LOOP) {
      running.emplace_back(lt.loop(ref));
      return;
    }
    switch (lt.ir.node(lt.node(ref)).op()) {
      case Operation::copy:
      case Operation::read:
      case Operation::write:
      case Operation::view:
        return;
      default:
        break;
    }
    std::unordered_map<IR::VarRef, int64_t> extent;
    for (auto it = running.rbegin(); it != running.rend(); ++it) {
      auto &e = extent.emplace(it->var, 1).first->second;
      e = e * it->size + it->tail;
    }
    int64_t iters = 1;
    for (const auto &p : extent) iters *= p.second;
    total += iters;
  });
  return total;
}

bool is_trivially_parallel(const LoopTree &lt, LoopTree::TreeRef ref) {
  if (lt.kind(ref) == LoopTree::NODE) return false;
  auto var = lt.loop(ref).var;
  bool ok = true;
  lt.walk(
      [&](LoopTree::TreeRef t, int) {
        if (lt.kind(t) == LoopTree::LOOP) return;
        auto nr = lt.node(t);
        if (lt.ir.node(nr).op() == Operation::view) return;
        auto lv = lt.ir.loop_vars(nr);
        if (std::find(lv.begin(), lv.end(), var) == lv.end()) return;
        auto pw = lt.ir.pointwise_vars(nr);
        if (std::find(pw.begin(), pw.end(), var) == pw.end()) ok = false;
      },
      ref);
  return ok;
}

std::vector<IR::NodeRef> find(const IR &ir, Operation op) {
  std::vector<IR::NodeRef> out;
  for (auto n : ir.nodes())
    if (ir.node(n).op() == op) out.emplace_back(n);
  return out;
}

}  // namespace loop_tool
