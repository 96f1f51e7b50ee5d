// Dataflow analysis of a LoopTree (see nest.h).
#include "nest.h"

#include <algorithm>
#include <functional>
#include <sstream>

namespace loop_tool {
namespace hip {

using symbolic::Expr;
using symbolic::Op;
using symbolic::Symbol;

Range expr_range(const Expr &e, const std::map<int, int64_t> &sizes) {
  Range r;
  auto s = e.simplify();
  // affine form: walk the sum of terms coef*sym or constants
  std::function<bool(const Expr &, int64_t, int64_t &, int64_t &)> rec =
      [&](const Expr &x, int64_t scale, int64_t &lo, int64_t &hi) -> bool {
    if (x.type() == Expr::Type::value) {
      lo += scale * x.value();
      hi += scale * x.value();
      return true;
    }
    if (x.type() == Expr::Type::symbol) {
      auto it = sizes.find(x.symbol().id());
      if (it == sizes.end()) return false;
      int64_t m = std::max<int64_t>(it->second - 1, 0);
      if (scale >= 0) {
        hi += scale * m;
      } else {
        lo += scale * m;
      }
      return true;
    }
    if (x.op() == Op::add) return rec(x.arg(0), scale, lo, hi) && rec(x.arg(1), scale, lo, hi);
    if (x.op() == Op::negate) return rec(x.arg(0), -scale, lo, hi);
    if (x.op() == Op::multiply) {
      if (x.arg(1).type() == Expr::Type::value)
        return rec(x.arg(0), scale * x.arg(1).value(), lo, hi);
      if (x.arg(0).type() == Expr::Type::value)
        return rec(x.arg(1), scale * x.arg(0).value(), lo, hi);
      return false;
    }
    return false;
  };
  int64_t lo = 0, hi = 0;
  r.ok = rec(s, 1, lo, hi);
  r.lo = lo;
  r.hi = hi;
  return r;
}

std::unique_ptr<Dataflow> analyze(const LoopTree &lt) {
  auto df = std::make_unique<Dataflow>();
  df->compiler = std::make_unique<Compiler>(lt);
  const auto &c = *df->compiler;
  const auto &ir = c.lt.ir;
  df->num_inputs = (int)ir.inputs().size();

  for (size_t i = 0; i < ir.inputs().size(); ++i) {
    DFNode n;
    n.nr = ir.inputs()[i];
    n.op = Operation::read;
    n.vars = ir.node(n.nr).vars();
    n.input_index = (int)i;
    df->nodes[n.nr] = n;
  }

  for (auto nr : ir.nodes()) {
    const auto &node = ir.node(nr);
    auto op = node.op();
    if (op == Operation::read || op == Operation::view) {
      if (op == Operation::view)
        ASSERT(!c.lt.scheduled.count(nr))
            << "the hip backend folds views into accesses; scheduled view "
            << ir.dump(nr) << " is not supported";
      continue;
    }
    ASSERT(op != Operation::copy) << "copy nodes are not supported by the hip backend";
    ASSERT(c.lt.scheduled.count(nr)) << "unscheduled compute node " << ir.dump(nr);
    auto ref = c.lt.scheduled.at(nr);
    DFNode n;
    n.nr = nr;
    n.op = op;
    n.vars = node.vars();
    auto lv = ir.all_vars(nr);
    auto vset = to_set(n.vars);
    for (auto v : lv)
      if (!vset.count(v)) n.red_vars.emplace_back(v);
    if (op == Operation::write) {
      for (size_t j = 0; j < ir.outputs().size(); ++j)
        if (ir.outputs()[j] == nr) n.output_index = (int)j;
      ASSERT(n.output_index >= 0) << "write node not among IR outputs";
      ASSERT(n.red_vars.empty()) << "reducing writes are not supported";
    }
    // symbols a node may index with: its loop vars; any other symbol left in
    // an access is out of scope and reads as index 0 (cpp.cpp:33-60)
    std::set<int> allowed;
    for (auto v : lv) allowed.insert(c.var_to_sym.at(v).id());
    for (auto inp : node.inputs()) {
      DFInput in;
      in.src = c.resolved_reads.at(inp);
      auto acc = c.gen_access(inp, ref);
      for (auto e : acc.full_exprs) {
        e = e.walk([&](const Expr &x) {
               if (x.type() == Expr::Type::symbol && !allowed.count(x.symbol().id()))
                 return Expr(0);
               return x;
             }).simplify();
        in.coords.emplace_back(e);
      }
      const auto &src = ir.node(in.src);
      ASSERT(src.op() != Operation::view)
          << "access resolves to a scheduled view in " << ir.dump(nr);
      ASSERT(in.coords.size() == src.vars().size())
          << "rank mismatch reading " << ir.dump(in.src) << " from " << ir.dump(nr);
      n.inputs.emplace_back(in);
    }
    // the reduction loop chain: ancestors of the leaf over reduced vars
    auto red = to_set(n.red_vars);
    std::vector<RedLoop> chain;
    for (auto p = c.lt.parent(ref); p != -1; p = c.lt.parent(p)) {
      auto l = c.lt.loop(p);
      auto &sizes = n.loops[c.var_to_sym.at(l.var).id()];
      sizes.insert(sizes.begin(), l.size);
      if (!red.count(l.var)) continue;
      std::stringstream annot(c.lt.annotation(p));
      bool unroll = false;
      for (std::string tok; annot >> tok;) unroll |= tok == "unroll";
      chain.push_back(RedLoop{l.var, l.size, l.tail, c.inner_sizes.at(p), unroll});
    }
    std::reverse(chain.begin(), chain.end());
    for (auto v : n.red_vars) {
      bool covered = false;
      for (const auto &l : chain) covered |= l.var == v;
      ASSERT(covered) << "reduction var " << ir.var(v).name() << " of "
                      << ir.dump(nr) << " has no loop";
    }
    n.red_loops = chain;
    df->nodes[nr] = n;
  }

  for (auto &p : df->nodes)
    for (const auto &in : p.second.inputs) df->nodes.at(in.src).consumers++;

  // dependence order among compute/write nodes
  std::set<IR::NodeRef> done;
  std::function<void(IR::NodeRef)> visit = [&](IR::NodeRef nr) {
    if (done.count(nr)) return;
    done.insert(nr);
    const auto &n = df->nodes.at(nr);
    for (const auto &in : n.inputs) visit(in.src);
    if (n.op != Operation::read) df->topo.emplace_back(nr);
  };
  for (auto o : ir.outputs()) {
    ASSERT(df->nodes.count(o)) << "output is not a scheduled write";
    visit(o);
    df->writes.emplace_back(o);
  }
  return df;
}

}  // namespace hip
}  // namespace loop_tool
