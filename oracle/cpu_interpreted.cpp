// TEST INFRASTRUCTURE ONLY -- the secondary CPU oracle of this repository.
//
// A restatement of the reference's interpreted CPU backend, registered as
// "cpu_interpreted" (include/loop_tool/compile.h:122-126,
// src/core/compile.cpp:163-424 and 890-1159): the LoopTree becomes a tree of
// closures over an iterator array indexed by loop depth.  Semantics followed
// from the reference (each is cited where it is implemented):
//   * a loop resets the allocations scoped to it before every iteration and
//     once more after the loop, then runs its tail with the inner loops of the
//     same var sized by the override;
//   * an access is sum(indices[depth] * stride) + offset, strides from the
//     derivative of the scoped index expression per loop symbol; a failing
//     bound predicate makes the access absent;
//   * binary nodes fold their inputs in order with float arithmetic and the
//     select-style min/max; an absent first input writes 0 (reductions too);
//   * unary nodes use the float overloads (std::exp(float), std::abs(float));
//   * a loop annotated "[backend]" runs its subtree on that backend with the
//     buffers offset by the enclosing loop indices (compile.cpp:273-386).
// The reference computes every predicate's offset from the intercept of the
// access expression instead of the predicate's own (compile.cpp:912-914);
// that is kept, so padded views evaluate as the reference's interpreter
// evaluates them (differently from its C backend), and the tests pin this
// oracle bit-for-bit against the reference build's own interpreter.
#include <cmath>
#include <functional>
#include <limits>
#include <sstream>
#include <unordered_set>

#include "loop_tool/compile.h"
#include "loop_tool/ir.h"
#include "loop_tool/mutate.h"

namespace loop_tool {
namespace oracle {

using symbolic::Expr;
using symbolic::Symbol;
using Fn = std::function<void(const std::vector<void *> &, int64_t *)>;
using AccessFn = std::function<float *(const std::vector<void *> &, const int64_t *)>;

class Interpreter : public Compiler {
 public:
  explicit Interpreter(const LoopTree &lt) : Compiler(lt) {}

  // compile.cpp:359-416
  Fn exec(LoopTree::TreeRef ref, std::unordered_map<IR::VarRef, int> overrides) const {
    if (ref == -1) {
      std::vector<Fn> roots;
      for (auto c : lt.roots) roots.push_back(exec(c, overrides));
      Fn rst = reset(-1);
      return [roots, rst](const std::vector<void *> &mem, int64_t *idx) {
        rst(mem, idx);
        for (const auto &f : roots) f(mem, idx);
      };
    }
    std::stringstream ss(lt.annotation(ref));
    std::string tok;
    while (ss >> tok)
      if (tok.size() > 2 && tok.front() == '[' && tok.back() == ']')
        return backend_exec(ref, tok.substr(1, tok.size() - 2));
    if (lt.kind(ref) == LoopTree::NODE) return node(ref);
    return loop(ref, overrides);
  }

 private:
  // compile.cpp:163-212
  Fn loop(LoopTree::TreeRef ref, std::unordered_map<IR::VarRef, int> overrides) const {
    const auto l = lt.loop(ref);
    int64_t size = l.size, tail = l.tail;
    if (overrides.count(l.var)) {
      int64_t o = overrides.at(l.var), inner = inner_sizes.at(ref);
      size = o / inner;
      tail = o % inner;
      overrides.erase(l.var);
    }
    std::vector<Fn> body, tail_body;
    for (auto c : lt.children(ref)) body.push_back(exec(c, overrides));
    if (tail > 0) {
      overrides[l.var] = (int)tail;
      for (auto c : lt.children(ref)) tail_body.push_back(exec(c, overrides));
    }
    Fn rst = reset(ref);
    const int d = lt.depth(ref);
    ASSERT(d < MAX_DEPTH) << "loop nest deeper than " << MAX_DEPTH;
    return [=](const std::vector<void *> &mem, int64_t *idx) {
      for (int64_t i = 0; i < size; ++i) {
        rst(mem, idx);
        idx[d] = i;
        for (const auto &f : body) f(mem, idx);
      }
      rst(mem, idx);
      idx[d] = size;
      for (const auto &f : tail_body) f(mem, idx);
    };
  }

  // compile.cpp:214-271
  Fn reset(LoopTree::TreeRef ref) const {
    std::vector<std::tuple<int, int64_t, float>> fills;
    for (const auto &p : allocations) {
      const auto &a = p.second;
      if (a.lca != ref) continue;
      float v = 0;
      switch (lt.ir.node(a.node_ref).op()) {
        case Operation::add:
        case Operation::subtract: v = 0; break;
        case Operation::multiply:
        case Operation::divide: v = 1; break;
        case Operation::min: v = std::numeric_limits<float>::max(); break;
        case Operation::max: v = -std::numeric_limits<float>::max(); break;
        case Operation::read:
        case Operation::write:
        case Operation::view:
        case Operation::exp:
        case Operation::sqrt:
        case Operation::negate:
        case Operation::reciprocal: continue;
        default:
          ASSERT(0) << "cannot generate reset for op: " << lt.ir.dump(a.node_ref);
      }
      fills.emplace_back(a.mem_idx, a.size(), v);
    }
    return [fills](const std::vector<void *> &mem, int64_t *) {
      for (const auto &f : fills) {
        float *p = (float *)mem[std::get<0>(f)];
        for (int64_t i = 0; i < std::get<1>(f); ++i) p[i] = std::get<2>(f);
      }
    };
  }

  // compile.cpp:890-937
  AccessFn access_fn(const Access &acc, LoopTree::TreeRef ref) const {
    const Expr expr = get_scoped_expr(acc);
    const auto conds = get_constraints(acc);
    const int depth = lt.depth(ref);
    auto sym_strides = get_symbol_strides(ref, acc.alloc.lca);
    const int64_t offset = symbolic::intercept(expr).evaluate();  // shared, see header
    auto strides_of = [&](const Expr &e) {
      std::vector<int64_t> s(depth, 0);
      for (const auto &sym : e.symbols()) {
        auto dx = symbolic::differentiate(e, sym).simplify();
        ASSERT(dx.can_evaluate()) << "Invalid indexing expr";
        const int64_t base = dx.evaluate();
        for (const auto &p : sym_strides.at(sym)) s[lt.depth(p.first)] = p.second * base;
      }
      return s;
    };
    std::vector<std::pair<std::vector<int64_t>, int64_t>> checks;
    for (const auto &c : conds) checks.emplace_back(strides_of(c.first), c.second);
    const auto strides = strides_of(expr);
    const int mem_idx = acc.alloc.mem_idx;
    auto eval = [offset](const std::vector<int64_t> &s, const int64_t *idx) {
      int64_t v = 0;
      for (size_t i = 0; i < s.size(); ++i) v += idx[i] * s[i];
      return v + offset;
    };
    return [=](const std::vector<void *> &mem, const int64_t *idx) -> float * {
      for (const auto &c : checks) {
        const int64_t v = eval(c.first, idx);
        if (v < 0 || (c.second != -1 && v >= c.second)) return nullptr;
      }
      return (float *)mem[mem_idx] + eval(strides, idx);
    };
  }

  // compile.cpp:939-966
  Fn mem_node(LoopTree::TreeRef ref) const {
    const auto nr = lt.node(ref);
    const auto &n = lt.ir.node(nr);
    if (n.op() == Operation::read) return [](const std::vector<void *> &, int64_t *) {};
    ASSERT(n.inputs().size() == 1) << "memory node with several inputs";
    auto in = access_fn(gen_access(n.inputs()[0], ref), ref);
    auto out = access_fn(gen_access(nr, ref), ref);
    return [in, out](const std::vector<void *> &mem, int64_t *idx) {
      float *o = out(mem, idx);
      if (!o) return;
      float *i = in(mem, idx);
      *o = i ? *i : 0.0f;
    };
  }

  // compile.cpp:968-1032
  Fn binary_node(LoopTree::TreeRef ref) const {
    const auto nr = lt.node(ref);
    const auto &n = lt.ir.node(nr);
    std::vector<AccessFn> ins;
    for (auto i : n.inputs()) ins.push_back(access_fn(gen_access(i, ref), ref));
    auto out = access_fn(gen_access(nr, ref), ref);
    std::function<float(float, float)> f;
    switch (n.op()) {
      case Operation::add: f = [](float a, float b) { return a + b; }; break;
      case Operation::subtract: f = [](float a, float b) { return a - b; }; break;
      case Operation::multiply: f = [](float a, float b) { return a * b; }; break;
      case Operation::divide: f = [](float a, float b) { return a / b; }; break;
      case Operation::min: f = [](float a, float b) { return a < b ? a : b; }; break;
      case Operation::max: f = [](float a, float b) { return a > b ? a : b; }; break;
      default: ASSERT(0) << lt.ir.dump(nr) << " is not a binary operation";
    }
    const bool reduction = !lt.ir.reduction_vars(nr).empty();
    return [ins, out, f, reduction](const std::vector<void *> &mem, int64_t *idx) {
      float *o = out(mem, idx);
      if (!o) return;
      float *x = ins[0](mem, idx);
      if (x && reduction)
        *o = f(*o, *x);
      else
        *o = x ? *x : 0.0f;
      for (size_t i = 1; i < ins.size(); ++i)
        if (float *y = ins[i](mem, idx)) *o = f(*o, *y);
    };
  }

  // compile.cpp:1034-1074
  Fn unary_node(LoopTree::TreeRef ref) const {
    const auto nr = lt.node(ref);
    const auto &n = lt.ir.node(nr);
    ASSERT(n.inputs().size() == 1) << "unary op with several inputs";
    auto in = access_fn(gen_access(n.inputs()[0], ref), ref);
    auto out = access_fn(gen_access(nr, ref), ref);
    std::function<float(float)> f;
    switch (n.op()) {
      case Operation::log: f = [](float a) { return std::log(a); }; break;
      case Operation::exp: f = [](float a) { return std::exp(a); }; break;
      case Operation::sqrt: f = [](float a) { return std::sqrt(a); }; break;
      case Operation::abs: f = [](float a) { return std::abs(a); }; break;
      case Operation::negate: f = [](float a) { return -a; }; break;
      case Operation::reciprocal: f = [](float a) { return 1 / a; }; break;
      default: ASSERT(0) << lt.ir.dump(nr) << " is not a unary operation";
    }
    return [in, out, f](const std::vector<void *> &mem, int64_t *idx) {
      float *o = out(mem, idx);
      if (!o) return;
      float *i = in(mem, idx);
      *o = i ? f(*i) : 0.0f;
    };
  }

  // compile.cpp:1076-1107
  Fn node(LoopTree::TreeRef ref) const {
    switch (lt.ir.node(lt.node(ref)).op()) {
      case Operation::read:
      case Operation::view:
      case Operation::write: return mem_node(ref);
      case Operation::add:
      case Operation::subtract:
      case Operation::multiply:
      case Operation::divide:
      case Operation::min:
      case Operation::max: return binary_node(ref);
      case Operation::log:
      case Operation::exp:
      case Operation::sqrt:
      case Operation::reciprocal:
      case Operation::negate:
      case Operation::abs: return unary_node(ref);
      default: ASSERT(0) << "Cannot generate node: " << lt.ir.dump(lt.node(ref));
    }
    return {};
  }

  // compile.cpp:273-357: the subtree under `ref` runs on `backend`; its
  // buffers are the full buffers offset by the enclosing loops' indices
  // times the strides of the vars the subtree drops
  Fn backend_exec(LoopTree::TreeRef ref, const std::string &backend) const {
    const auto &backends = getBackends();
    ASSERT(backends.count(backend)) << "Can't find backend " << backend;
    std::unordered_map<IR::NodeRef, IR::NodeRef> node_map;
    std::unordered_map<IR::VarRef, IR::VarRef> var_map;
    LoopTree sub = subtree(lt, ref, node_map, var_map);
    const int depth = lt.depth(ref);
    std::unordered_map<int, int> mem_of;  // sub-tree buffer -> buffer
    std::unordered_map<int, std::unordered_map<int, int64_t>> offsets;  // per sub buffer
    auto induced = [&](IR::NodeRef old_nr, IR::NodeRef new_nr, int sub_idx) {
      const auto &vars = lt.ir.node(old_nr).vars();
      std::unordered_set<IR::VarRef> kept;
      for (auto v : sub.ir.node(new_nr).vars()) kept.insert(v);
      auto sym_strides = get_symbol_strides(ref, -1);
      for (size_t i = 0; i < vars.size(); ++i) {
        int64_t stride = 1;
        for (size_t j = i + 1; j < vars.size(); ++j) stride *= var_sizes.at(vars[j]);
        if (var_map.count(vars[i]) && kept.count(var_map.at(vars[i]))) stride = 0;
        for (const auto &p : sym_strides.at(var_to_sym.at(vars[i]))) {
          if (p.first >= depth) continue;
          offsets[sub_idx][p.first] = p.second * stride;
        }
      }
    };
    int k = 0;
    for (auto sn : sub.ir.inputs()) {
      int i = 0;
      bool found = false;
      for (auto n : lt.ir.inputs()) {
        if (n == node_map[sn]) {
          mem_of[k] = i;
          induced(n, sn, k);
          found = true;
          break;
        }
        ++i;
      }
      ASSERT(found) << "sub-tree input not found";
      ++k;
    }
    const int n_in = (int)lt.ir.inputs().size();
    for (auto sn : sub.ir.outputs()) {
      int i = 0;
      bool found = false;
      for (auto n : lt.ir.outputs()) {
        if (n == node_map[sn]) {
          mem_of[k] = n_in + i;
          induced(n, sn, k);
          found = true;
          break;
        }
        ++i;
      }
      ASSERT(found) << "sub-tree output not found";
      ++k;
    }
    std::shared_ptr<Compiled> cc = backends.at(backend)->compile(sub);
    const int n = k;
    return [cc, mem_of, offsets, n](const std::vector<void *> &mem, int64_t *idx) {
      std::vector<void *> sub_mem(n);
      for (int i = 0; i < n; ++i) {
        float *p = (float *)mem[mem_of.at(i)];
        auto it = offsets.find(i);
        if (it != offsets.end())
          for (const auto &o : it->second) p += idx[o.first] * o.second;
        sub_mem[i] = p;
      }
      cc->run(sub_mem);
    };
  }
};

// compile.cpp:1109-1151
struct InterpretedCompiled : public Compiled {
  Fn fn;
  std::string text;
  std::vector<int64_t> sizes;
  mutable std::vector<void *> mem;

  explicit InterpretedCompiled(const LoopTree &lt) {
    Interpreter in(lt);
    fn = in.exec(-1, {});
    text = "[interpreted]\n" + lt.dump();
    sizes = in.memory_sizes();
    mem.assign(sizes.size(), nullptr);
    for (size_t i = 0; i < sizes.size(); ++i)
      if (sizes[i] > 0) mem[i] = calloc(sizes[i], sizeof(float));
  }
  ~InterpretedCompiled() override {
    for (size_t i = 0; i < sizes.size(); ++i)
      if (sizes[i] > 0) free(mem[i]);
  }
  void run(const std::vector<void *> &memory, bool) const override {
    int64_t idx[MAX_DEPTH] = {0};
    std::vector<void *> m = mem;
    for (size_t i = 0; i < memory.size() && i < m.size(); ++i) m[i] = memory[i];
    fn(m, idx);
  }
  std::string dump() const override { return text; }
};

struct InterpretedBackend : public Backend {
  InterpretedBackend() : Backend("cpu_interpreted") {}
  std::unique_ptr<Compiled> compile_impl(const LoopTree &lt) const override {
    return std::make_unique<InterpretedCompiled>(lt);
  }
  int hardware_requirement() const override { return 1 << 0; }
};

static RegisterBackend register_interpreted(std::make_shared<InterpretedBackend>());

}  // namespace oracle
}  // namespace loop_tool
