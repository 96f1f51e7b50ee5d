// Lowering analysis (see include/loop_tool/compile.h).
// Reference behaviour: src/core/compile.cpp (constructor :33-157, memory_sizes
// :482-499, gen_alloc :502-589, gen_index_equations :611-756, reify/bounds
// :758-803, gen_access :805-888, scoped expr/strides/constraints :433-480).
#include "loop_tool/compile.h"

#include <algorithm>
#include <cstdlib>
#include <functional>
#include <sstream>

namespace loop_tool {

using symbolic::Expr;
using symbolic::Op;
using symbolic::Symbol;

namespace {
size_t next_compile_count() {
  static size_t count = 0;
  return count++;
}
}  // namespace

Compiler::Compiler(const LoopTree &lt_) : lt(lt_) {
  count = next_compile_count();
  const auto &ir = lt.ir;

  // Unscheduled views are transparent: reads look through them toward the
  // buffer they view, writes toward the node they feed.
  auto resolve_view = [&](IR::NodeRef n, bool as_write) {
    if (ir.node(n).op() != Operation::view) return n;
    while (!lt.scheduled.count(n)) {
      const auto &node = ir.node(n);
      if (node.op() == Operation::write || node.op() == Operation::read) return n;
      ASSERT(node.op() == Operation::view) << "non-view in view chain";
      if (as_write) {
        ASSERT(node.outputs().size() == 1) << "view written through must have one use";
        n = node.outputs()[0];
      } else {
        ASSERT(node.inputs().size() == 1) << "view must have one input";
        n = node.inputs()[0];
      }
    }
    return n;
  };

  for (auto nr : ir.nodes()) {
    resolved_reads[nr] = resolve_view(nr, false);
    resolved_writes[nr] = resolve_view(nr, true);
    const auto &node = ir.node(nr);
    for (const auto &c : node.constraints()) {
      for (const auto &e : {c.first, c.second})
        for (const auto &s : e.symbols())
          if (node.has_sym(s)) {
            var_to_sym[node.var(s)] = s;
            sym_to_var[s] = node.var(s);
          }
      if (c.first.type() == Expr::Type::function && c.first.op() == Op::size) {
        auto a = c.first.arg(0);
        if (a.type() == Expr::Type::symbol && c.second.can_evaluate())
          var_sizes[sym_to_var.at(a.symbol())] = c.second.evaluate();
      }
    }
    // sizes implied by a user loop order (inner loops first)
    std::unordered_map<IR::VarRef, int64_t> implied;
    const auto &order = ir.order(nr);
    for (auto it = order.rbegin(); it != order.rend(); ++it) {
      auto &s = implied.emplace(it->first, 1).first->second;
      s = s * it->second.size + it->second.tail;
    }
    for (const auto &p : implied) {
      auto f = var_sizes.find(p.first);
      var_sizes[p.first] = f == var_sizes.end() ? p.second : std::max(p.second, f->second);
    }
  }
  for (auto v : ir.vars()) {
    if (var_to_sym.count(v)) continue;
    Symbol s(ir.var(v).name() + "_gen");
    var_to_sym.emplace(v, s);
    sym_to_var.emplace(s, v);
  }
  for (auto v : ir.vars())
    ASSERT(var_sizes.count(v)) << "size could not be deduced for var "
                               << ir.var(v).name();

  // inner size of a loop = extent of the same var in the loops below it
  std::unordered_set<LoopTree::TreeRef> seen;
  for (auto n : ir.nodes()) {
    if (!lt.scheduled.count(n)) continue;
    auto p = lt.parent(lt.scheduled.at(n));
    if (seen.count(p)) continue;
    seen.insert(p);
    std::unordered_map<IR::VarRef, int64_t> extent;
    while (p != -1) {
      auto loop = lt.loop(p);
      int64_t inner = extent.count(loop.var) ? extent.at(loop.var) : 1;
      if (inner_sizes.count(p)) inner = std::max(inner, inner_sizes.at(p));
      inner_sizes[p] = inner;
      int64_t size = loop.size * inner + loop.tail;
      extent[loop.var] = size;
      if (var_sizes.count(loop.var)) size = std::max(var_sizes.at(loop.var), size);
      var_sizes[loop.var] = size;
      p = lt.parent(p);
    }
  }

  for (auto nr : ir.nodes()) allocations[nr] = gen_alloc(nr);
}

Compiler::Allocation Compiler::gen_alloc(IR::NodeRef node_ref) const {
  const auto &ir = lt.ir;
  const auto &ins = ir.inputs();
  const auto &outs = ir.outputs();
  int n_io = (int)(ins.size() + outs.size());
  int mem_idx = -1;
  for (size_t i = 0; i < ins.size(); ++i)
    if (ins[i] == node_ref) mem_idx = (int)i;
  for (size_t i = 0; i < outs.size(); ++i)
    if (outs[i] == node_ref) mem_idx = (int)(i + ins.size());
  if (mem_idx == -1) {
    mem_idx = n_io;
    for (const auto &p : allocations)
      if (p.second.mem_idx >= n_io) mem_idx++;
  }

  const auto &node = ir.node(node_ref);
  bool is_io_op = node.op() == Operation::write || node.op() == Operation::read;
  if (!lt.scheduled.count(node_ref)) {
    std::vector<int64_t> sizes;
    if (is_io_op)
      for (auto v : node.vars()) sizes.emplace_back(var_sizes.at(v));
    return Allocation(mem_idx, node_ref, sizes, -1);
  }

  // tree refs of the scheduled consumers (looking through unscheduled nodes;
  // an unscheduled write counts as the root)
  std::function<std::vector<LoopTree::TreeRef>(IR::NodeRef)> consumers =
      [&](IR::NodeRef nr) {
        std::vector<LoopTree::TreeRef> out;
        for (auto o : ir.node(nr).outputs()) {
          if (lt.scheduled.count(o)) {
            out.emplace_back(lt.scheduled.at(o));
          } else if (ir.node(o).op() == Operation::write) {
            out.emplace_back(-1);
          } else {
            for (auto t : consumers(o)) out.emplace_back(t);
          }
        }
        return out;
      };
  auto ref = lt.parent(lt.scheduled.at(node_ref));
  auto lca = ref;
  for (auto t : consumers(node_ref)) lca = lt.lca(lca, t);
  if (is_io_op) lca = -1;

  std::unordered_map<IR::VarRef, int64_t> extent;
  while (ref != lca) {
    auto loop = lt.loop(ref);
    ref = lt.parent(ref);
    auto &e = extent.emplace(loop.var, 1).first->second;
    e = e * loop.size + loop.tail;
  }
  std::vector<int64_t> sizes;
  for (auto v : node.vars()) {
    if (extent.count(v)) {
      ASSERT(extent.at(v) > 0) << "empty extent for " << ir.var(v).name();
      sizes.emplace_back(extent.at(v));
    } else {
      sizes.emplace_back(1);
    }
  }
  return Allocation(mem_idx, node_ref, sizes, lca);
}

std::pair<std::vector<Expr>, std::vector<Expr>> Compiler::gen_index_equations(
    IR::NodeRef read_node_ref, IR::NodeRef write_node_ref,
    LoopTree::TreeRef ref) const {
  const auto &ir = lt.ir;
  // views between the accessed node and the buffer it resolves to
  std::vector<IR::NodeRef> rev_chain;
  for (auto n = write_node_ref; n != read_node_ref; n = ir.node(n).inputs().at(0)) {
    ASSERT(ir.node(n).op() == Operation::view) << "non-view in chain!";
    rev_chain.emplace_back(n);
  }
  std::vector<IR::NodeRef> chain(rev_chain.rbegin(), rev_chain.rend());

  // index of `var` (an input var of the view when !toward_input, an output
  // var otherwise) as an expression of the other side's vars
  auto get_expr = [&](IR::NodeRef node_ref, IR::VarRef var, bool toward_input) {
    const auto &sym = var_to_sym.at(var);
    const auto &node = ir.node(node_ref);
    auto out_vars = to_set(node.vars());
    std::unordered_set<IR::VarRef> in_vars;
    if (!node.inputs().empty()) in_vars = to_set(ir.node(node.inputs()[0]).vars());
    for (const auto &c : node.constraints()) {
      if (!(c.first == Expr(sym))) continue;
      bool only_out = true, only_in = true;
      for (const auto &s : c.second.symbols(false)) {
        auto v = sym_to_var.at(s);
        only_out &= out_vars.count(v) > 0;
        only_in &= in_vars.count(v) > 0;
      }
      if (!toward_input && only_out) return c.second;
      if (toward_input && only_in) return c.second;
      // bijective views: drop foreign vars, wrapping by their stride
      if (!toward_input && ir.reduction_vars(node_ref).empty()) {
        auto expr = c.second;
        for (const auto &s : expr.symbols(false)) {
          if (out_vars.count(sym_to_var.at(s))) continue;
          auto mod = (-symbolic::differentiate(expr, s)).simplify();
          expr = expr.replace(s, Expr(0)).simplify();
          if (mod.can_evaluate() && mod.evaluate() > 1) expr = expr % mod;
        }
        bool ok = true;
        for (const auto &s : expr.symbols(false)) ok &= out_vars.count(sym_to_var.at(s)) > 0;
        if (ok) return expr;
      }
    }
    return Expr(sym);
  };

  std::unordered_set<Symbol, symbolic::Hash<Symbol>> avail;
  for (auto p = lt.parent(ref); p != -1; p = lt.parent(p)) {
    auto v = lt.loop(p).var;
    ASSERT(var_to_sym.count(v)) << "Cannot find symbolicated variable "
                                << ir.var(v).name();
    avail.insert(var_to_sym.at(v));
  }

  std::vector<Expr> read_exprs;
  for (auto v : ir.node(read_node_ref).vars())
    read_exprs.emplace_back(get_expr(read_node_ref, v, false));
  std::vector<Expr> write_exprs;
  for (auto v : ir.node(write_node_ref).vars())
    write_exprs.emplace_back(get_expr(write_node_ref, v, false));

  auto substitute = [&](std::vector<Expr> &exprs, IR::VarRef v, const Expr &e) {
    const auto &sym = var_to_sym.at(v);
    for (auto &x : exprs) {
      x = x.replace(sym, e).simplify();
      x = x.replace(Expr::size(Expr(sym)), var_sizes.at(v)).simplify();
    }
  };
  for (auto nr : chain) {
    for (auto v : ir.node(ir.node(nr).inputs().at(0)).vars()) {
      if (avail.count(var_to_sym.at(v))) continue;
      substitute(read_exprs, v, get_expr(nr, v, false));
    }
  }
  for (auto nr : rev_chain) {
    for (auto v : ir.node(nr).vars()) {
      if (avail.count(var_to_sym.at(v))) continue;
      substitute(write_exprs, v, get_expr(nr, v, true));
    }
  }
  return {read_exprs, write_exprs};
}

Expr Compiler::reify_sizes(const Expr &expr) const {
  return expr
      .walk([&](const Expr &e) {
        if (e.type() == Expr::Type::function && e.op() == Op::size)
          return Expr(var_sizes.at(sym_to_var.at(e.arg(0).symbol())));
        return e;
      })
      .simplify();
}

int64_t Compiler::get_expr_max(const Expr &expr) const {
  auto m = reify_sizes(expr)
               .walk([&](const Expr &e) {
                 if (e.type() == Expr::Type::symbol)
                   return Expr(std::max(var_sizes.at(sym_to_var.at(e.symbol())) - 1,
                                        (int64_t)0));
                 return e;
               })
               .simplify();
  ASSERT(m.can_evaluate()) << "Couldn't derive explicit upper bound for expr "
                           << expr.dump() << " (simplified to " << m.dump() << ")";
  return m.evaluate() + 1;
}

int64_t Compiler::get_expr_min(const Expr &expr) const {
  auto m = reify_sizes(expr)
               .walk([&](const Expr &e) {
                 if (e.type() == Expr::Type::symbol) return Expr(0);
                 return e;
               })
               .simplify();
  ASSERT(m.type() == Expr::Type::value)
      << "Couldn't derive explicit lower bound for expr " << expr.dump()
      << " (simplified to " << m.dump() << ")";
  return m.value();
}

Compiler::Access Compiler::gen_access(IR::NodeRef node_ref,
                                      LoopTree::TreeRef ref) const {
  auto read_node_ref = resolved_reads.at(node_ref);
  auto view_exprs = gen_index_equations(read_node_ref, node_ref, ref);
  bool is_write = node_ref == lt.node(ref);
  for (const auto &e : view_exprs.second)
    ASSERT(!is_write || e.type() == Expr::Type::symbol)
        << "viewed writes not yet supported, found expr: " << e.dump();

  const auto &alloc = allocations.at(read_node_ref);
  const auto &read_exprs = view_exprs.first;
  ASSERT(alloc.sizes.size() == read_exprs.size())
      << "access rank mismatch for " << lt.ir.dump(read_node_ref);

  Access access(alloc);
  std::unordered_set<Symbol, symbolic::Hash<Symbol>> in_scope;
  for (auto p = lt.parent(ref); p != alloc.lca; p = lt.parent(p))
    in_scope.insert(var_to_sym.at(lt.loop(p).var));

  for (size_t i = 0; i < read_exprs.size(); ++i) {
    auto expr = reify_sizes(read_exprs[i]);
    auto lo = get_expr_min(expr);
    auto hi = get_expr_max(expr);
    if (hi <= alloc.sizes[i]) {
      hi = -1;  // never exceeds the buffer
    } else {
      hi = alloc.sizes[i];
      ASSERT(hi > 0);
    }
    access.full_exprs.emplace_back(expr);
    access.bounds.emplace_back(lo, hi);
  }
  for (const auto &e : access.full_exprs) {
    access.scoped_exprs.emplace_back(
        e.walk([&](const Expr &x) {
           if (x.type() == Expr::Type::symbol && !in_scope.count(x.symbol()))
             return Expr(0);
           return x;
         }).simplify());
  }
  return access;
}

Expr Compiler::get_scoped_expr(const Access &access) const {
  Expr full(0);
  for (size_t i = 0; i < access.scoped_exprs.size(); ++i)
    full = full + access.scoped_exprs[i] * Expr(access.alloc.size((int)i + 1));
  return full.simplify();
}

std::unordered_map<Symbol, std::vector<std::pair<LoopTree::TreeRef, int64_t>>,
                   symbolic::Hash<Symbol>>
Compiler::get_symbol_strides(LoopTree::TreeRef ref, LoopTree::TreeRef root) const {
  std::unordered_map<Symbol, std::vector<std::pair<LoopTree::TreeRef, int64_t>>,
                     symbolic::Hash<Symbol>>
      out;
  for (auto p = lt.parent(ref); p != root; p = lt.parent(p))
    out[var_to_sym.at(lt.loop(p).var)].emplace_back(p, inner_sizes.at(p));
  return out;
}

std::vector<std::pair<Expr, int64_t>> Compiler::get_constraints(
    const Access &access) const {
  std::vector<std::pair<Expr, int64_t>> out;
  for (size_t i = 0; i < access.bounds.size(); ++i) {
    const auto &b = access.bounds[i];
    const auto &e = access.scoped_exprs[i];
    if (b.first == 0 && b.second == -1) continue;
    if (e.can_evaluate() && e.evaluate() == 0) continue;
    out.emplace_back(e, b.second);
  }
  return out;
}

std::vector<int64_t> Compiler::memory_sizes(bool include_io) const {
  std::vector<int64_t> mem(allocations.size(), 0);
  auto n_io = lt.ir.inputs().size() + lt.ir.outputs().size();
  for (const auto &p : allocations) {
    const auto &a = p.second;
    if (!include_io && a.mem_idx < (int)n_io) continue;
    int64_t s = 1;
    for (auto x : a.sizes) s *= x > 0 ? x : 1;
    mem.at(a.mem_idx) = s;
  }
  return mem;
}

std::vector<void *> Compiler::allocate() const {
  auto sizes = memory_sizes();
  std::vector<void *> mem(sizes.size(), nullptr);
  for (size_t i = 0; i < sizes.size(); ++i)
    if (sizes[i] > 0) mem[i] = calloc(sizes[i], sizeof(float));
  return mem;
}

std::string Compiler::gen_string() const {
  std::stringstream ss;
  ss << "[interpreted]\n" << lt.dump();
  return ss.str();
}

}  // namespace loop_tool
