// Pointwise lowering: one lane per element of the target value (an IR output
// or a materialised reduction), the contiguous (innermost) var on the lane
// axis so loads and stores of the target coalesce.
//
// Each lane evaluates the dataflow DAG for its element: pointwise nodes are
// inlined, reads become predicated loads (out-of-range view accesses read 0,
// as `cond ? mem[idx] : 0` in src/backends/cpu/cpp.cpp:64-106), and reductions
// replay the LoopTree's reduction loop chain, tails included
// (cpp.cpp:295-351), accumulating `acc = acc OP x` in the same order as the
// reference C backend.  Together with -ffp-contract=off this makes the path
// bit-exact with the C oracle for + - * / min max and for the index-only
// nests; exp/log/sqrt are evaluated in double and rounded to float as the C
// backend does through libm.  One deliberate difference: a tail block inside
// another loop's tail block starts at the iterations actually run there, not
// at the loop's full size as in the C backend, whose index then runs past the
// buffers (such schedules come from splitting one var three or more times).
#include <algorithm>
#include <functional>
#include <memory>
#include <sstream>

#include "emit.h"
#include "emit_internal.h"

namespace loop_tool {
namespace hip {

using symbolic::Expr;
using symbolic::Symbol;

std::string float_literal(double v) {
  std::stringstream s;
  s.precision(9);
  s << "(float)(" << v << ")";
  return s.str();
}

std::string unary_expr(Operation op, const std::string &x) {
  switch (op) {
    case Operation::exp:
      return "(float)exp((double)" + x + ")";
    case Operation::log:
      return "(float)log((double)" + x + ")";
    case Operation::sqrt:
      return "(float)sqrt((double)" + x + ")";
    case Operation::abs:
      return "fabsf(" + x + ")";
    case Operation::negate:
      return "(-" + x + ")";
    case Operation::reciprocal:
      return "(1.0f / " + x + ")";
    default:
      ASSERT(0) << "not a unary op: " << dump(op);
  }
  return "";
}

namespace {

// -------------------------------------------------------------- blocks ----
struct Block {
  std::string open, close;
  struct Item {
    std::string text;
    Block *child = nullptr;
  };
  std::vector<Item> items;
  std::vector<std::unique_ptr<Block>> owned;
  Block *parent = nullptr;
  Block *open_child = nullptr;
  std::set<int> bound;                    // symbol ids usable in this block
  std::map<std::string, std::string> memo;  // value cache: key -> variable

  // new items go just before the child that is currently being filled
  size_t insert_pos() const {
    if (!open_child) return items.size();
    for (size_t i = 0; i < items.size(); ++i)
      if (items[i].child == open_child) return i;
    return items.size();
  }
  void add(const std::string &line) {
    items.insert(items.begin() + insert_pos(), Item{line, nullptr});
  }
  Block *child(const std::string &o, const std::string &c) {
    auto b = std::make_unique<Block>();
    b->open = o;
    b->close = c;
    b->parent = this;
    b->bound = bound;
    Block *raw = b.get();
    items.insert(items.begin() + insert_pos(), Item{"", raw});
    owned.emplace_back(std::move(b));
    return raw;
  }
  void render(std::stringstream &ss, int indent) const {
    std::string pad(indent * 2, ' ');
    if (!open.empty()) ss << pad << open << "\n";
    int inner = open.empty() ? indent : indent + 1;
    for (const auto &it : items) {
      if (it.child)
        it.child->render(ss, inner);
      else
        ss << std::string(inner * 2, ' ') << it.text << "\n";
    }
    if (!close.empty()) ss << pad << close << "\n";
  }
};

}  // namespace

// ----------------------------------------------------------- evaluator ----
struct PointwiseEmitter::Impl {
  const Dataflow &df;
  const std::map<IR::NodeRef, int> &materialized;  // node -> kernel arg position
  const std::map<int, std::string> &input_args;     // input index -> arg name
  std::map<int, int64_t> extents;                    // symbol id -> extent
  std::map<int, std::string> names;                  // symbol id -> C name
  int counter = 0;
  std::string ity = "long";                          // index type

  Impl(const Dataflow &d, const std::map<IR::NodeRef, int> &m,
       const std::map<int, std::string> &a)
      : df(d), materialized(m), input_args(a) {}

  std::string fresh(const std::string &p) { return p + std::to_string(counter++); }

  std::string print(const Expr &e) {
    std::unordered_map<Symbol, std::string, symbolic::Hash<Symbol>> repl;
    for (const auto &s : e.symbols()) {
      auto it = names.find(s.id());
      ASSERT(it != names.end()) << "unbound symbol " << s.name() << " in " << e.dump();
      repl[s] = it->second;
    }
    return "(" + e.simplify().dump(false, repl) + ")";
  }

  Block *target_block(Block *cur, const std::vector<Expr> &coords) {
    std::set<int> syms;
    for (const auto &e : coords)
      for (const auto &s : e.symbols()) syms.insert(s.id());
    std::vector<Block *> path;
    for (Block *b = cur; b; b = b->parent) path.push_back(b);
    std::reverse(path.begin(), path.end());
    for (auto *b : path) {
      bool ok = true;
      for (auto id : syms) ok &= b->bound.count(id) > 0;
      if (ok) return b;
    }
    return cur;
  }

  std::string lookup(Block *b, const std::string &key) {
    for (; b; b = b->parent) {
      auto it = b->memo.find(key);
      if (it != b->memo.end()) return it->second;
    }
    return "";
  }

  // predicate that `coords` lie inside the dims of `vars`
  std::string bounds_check(const std::vector<Expr> &coords,
                           const std::vector<IR::VarRef> &vars) {
    std::vector<std::string> conds;
    for (size_t d = 0; d < coords.size(); ++d) {
      auto size = df.size(vars[d]);
      auto r = expr_range(coords[d], extents);
      bool need_lo = !r.ok || r.lo < 0;
      bool need_hi = !r.ok || r.hi > size - 1;
      auto s = print(coords[d]);
      if (need_lo) conds.push_back(s + " >= 0");
      if (need_hi) conds.push_back(s + " < " + std::to_string(size));
    }
    std::string out;
    for (size_t i = 0; i < conds.size(); ++i) out += (i ? " && " : "") + conds[i];
    return out;
  }

  std::string linear_index(const std::vector<Expr> &coords,
                           const std::vector<IR::VarRef> &vars) {
    Expr idx(0);
    int64_t stride = 1;
    for (int d = (int)vars.size() - 1; d >= 0; --d) {
      idx = idx + coords[d] * Expr(stride);
      stride *= df.size(vars[d]);
    }
    return print(idx.simplify());
  }

  std::string load(Block *b, const std::string &buf, const std::vector<Expr> &coords,
                   const std::vector<IR::VarRef> &vars) {
    auto v = fresh("v");
    auto pred = bounds_check(coords, vars);
    auto idx = linear_index(coords, vars);
    if (pred.empty())
      b->add("const float " + v + " = " + buf + "[" + idx + "];");
    else
      b->add("const float " + v + " = (" + pred + ") ? " + buf + "[" + idx + "] : 0.0f;");
    return v;
  }

  // value of node `nr` at `coords` (one Expr per node var)
  std::string value(Block *cur, IR::NodeRef nr, const std::vector<Expr> &coords) {
    const auto &n = df.nodes.at(nr);
    Block *b = target_block(cur, coords);
    std::string key = std::to_string(nr) + "@";
    for (const auto &c : coords) key += print(c) + ",";
    auto hit = lookup(b, key);
    if (!hit.empty()) return hit;

    std::string out;
    if (n.op == Operation::read) {
      out = load(b, input_args.at(n.input_index), coords, n.vars);
    } else if (materialized.count(nr)) {
      out = load(b, "s" + std::to_string(materialized.at(nr)), coords, n.vars);
    } else {
      auto raw = compute(b, n, coords);
      auto pred = bounds_check(coords, n.vars);
      if (pred.empty()) {
        out = raw;
      } else {
        out = fresh("v");
        b->add("const float " + out + " = (" + pred + ") ? " + raw + " : 0.0f;");
      }
    }
    b->memo[key] = out;
    return out;
  }

  std::vector<Expr> input_coords(const DFNode &n, const DFInput &in,
                                 const std::map<int, Expr> &subst) {
    std::vector<Expr> out;
    for (auto e : in.coords) {
      e = e.walk([&](const Expr &x) {
        if (x.type() == Expr::Type::symbol) {
          auto it = subst.find(x.symbol().id());
          if (it != subst.end()) return it->second;
        }
        return x;
      });
      out.emplace_back(e.simplify());
    }
    return out;
  }

  static std::string combine(Operation op, const std::vector<std::string> &xs) {
    std::string infix;
    switch (op) {
      case Operation::add: infix = " + "; break;
      case Operation::subtract: infix = " - "; break;
      case Operation::multiply: infix = " * "; break;
      case Operation::divide: infix = " / "; break;
      case Operation::max:
      case Operation::min: {
        // select with the left operand first, nested to the right
        const char *cmp = op == Operation::max ? " > " : " < ";
        std::function<std::string(size_t)> nest = [&](size_t i) -> std::string {
          if (i + 1 == xs.size()) return xs[i];
          auto rhs = nest(i + 1);
          return "lt_sel" + std::string(op == Operation::max ? "max" : "min") + "(" +
                 xs[i] + ", " + rhs + ")";
        };
        (void)cmp;
        return nest(0);
      }
      default:
        ASSERT(xs.size() == 1) << "unary op with several inputs";
        return unary_expr(op, xs[0]);
    }
    std::string s = "(";
    for (size_t i = 0; i < xs.size(); ++i) s += (i ? infix : "") + xs[i];
    return s + ")";
  }

  static std::string reduction_init(Operation op) {
    switch (op) {
      case Operation::add: return "0.0f";
      case Operation::multiply: return "1.0f";
      case Operation::max: return float_literal(-3.40282e+38);
      case Operation::min: return float_literal(3.40282e+38);
      default:
        ASSERT(0) << "cannot reduce with " << dump(op);
    }
    return "";
  }

  std::string compute(Block *b, const DFNode &n, const std::vector<Expr> &coords) {
    std::map<int, Expr> subst;
    for (size_t i = 0; i < n.vars.size(); ++i) subst[df.sym(n.vars[i]).id()] = coords[i];
    if (n.red_vars.empty()) {
      std::vector<std::string> xs;
      for (const auto &in : n.inputs)
        xs.push_back(value(b, in.src, input_coords(n, in, subst)));
      auto v = fresh("v");
      b->add("const float " + v + " = " + combine(n.op, xs) + ";");
      return v;
    }
    auto acc = fresh("acc");
    b->add("float " + acc + " = " + reduction_init(n.op) + ";");
    std::map<IR::VarRef, int> overrides;
    chain(b, n, 0, overrides, subst, acc);
    return acc;
  }

  // the reduction loop chain of `n`, mirroring the C backend's loop/tail
  // structure (overrides shrink inner loops of the same var in tail blocks)
  void chain(Block *b, const DFNode &n, size_t i, std::map<IR::VarRef, int> overrides,
             std::map<int, Expr> subst, const std::string &acc) {
    if (i == n.red_loops.size()) {
      // reduced var index = sum of loop iterators times their inner extents
      std::map<int, Expr> full = subst;
      for (auto v : n.red_vars) full[df.sym(v).id()] = Expr(0);
      for (const auto &p : iter_terms) {
        auto id = df.sym(p.first).id();
        full[id] = (full[id] + p.second).simplify();
      }
      std::vector<std::string> xs{acc};
      for (const auto &in : n.inputs)
        xs.push_back(value(b, in.src, input_coords(n, in, full)));
      b->add(acc + " = " + combine(n.op, xs) + ";");
      return;
    }
    const auto &L = n.red_loops[i];
    int64_t size = L.size, tail = L.tail;
    auto ov = overrides.find(L.var);
    if (ov != overrides.end()) {
      size = ov->second / L.inner;
      tail = ov->second % L.inner;
      overrides.erase(ov);
    }
    if (size > 0) {
      Symbol it("it");
      auto name = fresh("i");
      names[it.id()] = name;
      extents[it.id()] = size;
      // a loop the schedule marked "unroll" is unrolled completely
      auto *loop = b->child(std::string(L.unroll ? "#pragma unroll\n" : "") + "for (" + ity +
                                " " + name + " = 0; " + name + " < " + std::to_string(size) +
                                "; ++" + name + ") {",
                            "}");
      loop->bound.insert(it.id());
      Block *saved = b->open_child;
      b->open_child = loop;
      iter_terms.emplace_back(L.var, Expr(it) * Expr(L.inner));
      chain(loop, n, i + 1, overrides, subst, acc);
      iter_terms.pop_back();
      b->open_child = saved;
    }
    if (tail > 0) {
      auto *blk = b->child("{", "}");
      Block *saved = b->open_child;
      b->open_child = blk;
      auto o = overrides;
      o[L.var] = (int)tail;
      // a tail block starts after the `size` full iterations run in this
      // context.  The C backend pins it to the loop's own size instead
      // (cpp.cpp:338-346), which differs inside another loop's tail block:
      // there it indexes past the buffers and skips the last elements.
      iter_terms.emplace_back(L.var, Expr(size * L.inner));
      chain(blk, n, i + 1, o, subst, acc);
      iter_terms.pop_back();
      b->open_child = saved;
    }
  }

  std::vector<std::pair<IR::VarRef, Expr>> iter_terms;
};

PointwiseEmitter::PointwiseEmitter(const Dataflow &df,
                                   const std::map<IR::NodeRef, int> &materialized,
                                   const std::map<int, std::string> &input_args)
    : impl_(new Impl(df, materialized, input_args)) {}
PointwiseEmitter::~PointwiseEmitter() = default;

std::string PointwiseEmitter::kernel(const std::string &name, IR::NodeRef target,
                                     const std::string &signature,
                                     const std::string &out_arg, int block,
                                     int64_t *threads, int vec, bool int32) {
  auto &I = *impl_;
  const auto &df = I.df;
  const auto &n = df.nodes.at(target);
  I.ity = int32 ? "int" : "long";
  const std::string ity = I.ity;
  // the target's element space: a write covers its vars, a materialised node
  // its own vars.  Each lane owns E consecutive elements of the contiguous
  // (last) var and stores them with one vector store.
  std::vector<IR::VarRef> vars = n.vars;
  int64_t total = df.numel(vars);
  int64_t last = vars.empty() ? 1 : df.size(vars.back());
  int E = 1;
  for (int cand : {4, 2}) {
    if (vec > 0 && cand > vec) continue;
    if (last % cand == 0) {
      E = cand;
      break;
    }
  }
  const int64_t lanes = total / E;
  *threads = lanes;

  Block root;
  std::stringstream pre;
  pre << "const " << ity << " o = (" << ity << ")blockIdx.x * " << block
      << " + (" << ity << ")threadIdx.x;\n";
  pre << "if (o >= " << lanes << ") return;\n";
  pre << ity << " t = o;\n";
  std::vector<Expr> base;  // per var: its coordinate symbol, or the lane's chunk
  Symbol q("q");
  for (int d = (int)vars.size() - 1; d >= 0; --d) {
    bool is_last = d == (int)vars.size() - 1;
    Symbol s = is_last ? q : df.sym(vars[d]);
    int64_t extent = is_last ? last / E : df.size(vars[d]);
    auto cn = "c" + std::to_string(d);
    I.names[s.id()] = cn;
    I.extents[s.id()] = extent;
    root.bound.insert(s.id());
    if (d > 0)
      pre << "const " << ity << " " << cn << " = t % " << extent << "; t /= " << extent
          << ";\n";
    else
      pre << "const " << ity << " " << cn << " = t;\n";
  }
  for (size_t d = 0; d < vars.size(); ++d) base.emplace_back(Expr(df.sym(vars[d])));

  std::vector<std::string> results;
  for (int e = 0; e < E; ++e) {
    std::vector<Expr> coords = base;
    if (!coords.empty()) coords.back() = (Expr(q) * Expr(E) + Expr(e)).simplify();
    if (n.op == Operation::write) {
      const auto &in = n.inputs.at(0);
      std::map<int, Expr> subst;
      for (size_t i = 0; i < vars.size(); ++i) subst[df.sym(vars[i]).id()] = coords[i];
      results.push_back(I.value(&root, in.src, I.input_coords(n, in, subst)));
    } else {
      results.push_back(I.compute(&root, n, coords));
    }
  }
  if (E == 1) {
    root.add(out_arg + "[o] = " + results[0] + ";");
  } else {
    std::string vt = E == 4 ? "lt_f4" : "lt_f2";
    std::string lit = "(" + vt + "){";
    for (int e = 0; e < E; ++e) lit += (e ? ", " : "") + results[e];
    lit += "}";
    root.add("__builtin_nontemporal_store(" + lit + ", (" + vt + "*)(" + out_arg + ") + o);");
  }

  std::stringstream ss;
  ss << "extern \"C\" __global__ void __launch_bounds__(" << block << ") " << name << "("
     << signature << ") {\n";
  std::stringstream body;
  root.render(body, 1);
  std::stringstream prelines(pre.str());
  std::string line;
  while (std::getline(prelines, line)) ss << "  " << line << "\n";
  ss << body.str() << "}\n";
  return ss.str();
}

}  // namespace hip
}  // namespace loop_tool
