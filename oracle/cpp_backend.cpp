// TEST INFRASTRUCTURE ONLY -- the CPU oracle of this repository.
//
// A restatement of the reference's default C backend, registered as "cpp"
// (src/backends/cpu/cpp.cpp:17-514): it prints a C99 function
// `void fn_<n>(void** memory)` from a LoopTree (int64 loop iterators, tail
// blocks, LCA-scoped resets, `cond ? mem[idx] : 0` predicates, infix ops,
// select-style max/min macros, double libm exp/log/sqrt) and builds it with
// the reference's flags `cc -Wall -Werror -O3 -std=c99 -fpic -shared`
// (-std=c99 keeps FP contraction off, so every product and sum rounds
// separately).  It is loaded by tests, smoke() and the benchmark's CPU
// baseline through loop_tool::loadLibrary; the product never links it.
//
// Each function cites the reference code it follows.  Parity of this oracle
// is pinned in tests against the reference library built from
// /root/reference (oracle/_ref), against closed forms, and against the
// reference's known-answer tests (test/test_lazy.cpp, test/test_cpp.cpp).
#include <unistd.h>

#include <cstdio>
#include <fstream>
#include <functional>
#include <limits>
#include <sstream>

#include "loop_tool/compile.h"
#include "loop_tool/dynlib.h"
#include "loop_tool/ir.h"

namespace loop_tool {
namespace oracle {

using symbolic::Expr;
using symbolic::Symbol;

class CCodegen : public Compiler {
 public:
  explicit CCodegen(const LoopTree &lt) : Compiler(lt) {}
  std::string gen_string() const override { return gen(-1, {}); }

 private:
  bool is_io(IR::NodeRef nr) const {
    for (auto i : lt.ir.inputs())
      if (i == nr) return true;
    for (auto o : lt.ir.outputs())
      if (o == nr) return true;
    return false;
  }

  std::string indent(LoopTree::TreeRef ref, int extra = 0) const {
    int depth = ref == -1 ? 0 : lt.depth(ref) + 1;
    return std::string((depth + extra) * 2, ' ');
  }

  // cpp.cpp:33-108
  std::string access(IR::NodeRef node_ref, LoopTree::TreeRef ref) const {
    auto acc = gen_access(node_ref, ref);
    std::unordered_map<Symbol, std::string, symbolic::Hash<Symbol>> names;
    for (auto p = lt.parent(ref); p != acc.alloc.lca; p = lt.parent(p)) {
      auto sym = var_to_sym.at(lt.loop(p).var);
      std::stringstream s;
      if (names.count(sym)) s << names.at(sym) << "+";
      s << "i_" << p;
      auto stride = inner_sizes.at(p);
      if (stride > 1) s << "*" << stride;
      names[sym] = s.str();
    }
    for (auto &p : names) p.second = "(" + p.second + ")";

    std::vector<std::pair<Expr, int64_t>> conds;
    if (node_ref != lt.node(ref)) conds = get_constraints(acc);
    std::stringstream ss;
    if (!conds.empty()) ss << "(";
    for (size_t i = 0; i < conds.size(); ++i) {
      auto str = conds[i].first.dump(false, names);
      if (i) ss << " && ";
      ss << "((" << str << ") >= 0)";
      if (conds[i].second != -1) ss << " && ((" << str << ") < " << conds[i].second << ")";
    }
    if (!conds.empty()) ss << " ? ";
    if (acc.alloc.size() > 1 || is_io(acc.alloc.node_ref)) {
      ss << "((float*)memory[" << acc.alloc.mem_idx << "])[";
      for (size_t i = 0; i < acc.scoped_exprs.size(); ++i) {
        if (!acc.alloc.sizes.at(i)) continue;
        ss << (acc.scoped_exprs[i] * Expr(acc.alloc.size((int)i + 1))).simplify().dump(false, names);
        if (i != acc.scoped_exprs.size() - 1) ss << "+";
      }
      if (acc.scoped_exprs.empty()) ss << "0";
      ss << "]";
    } else {
      ss << "v" << acc.alloc.mem_idx;
    }
    if (!conds.empty()) ss << " : 0)";
    return ss.str();
  }

  // cpp.cpp:110-120
  std::string mem_node(LoopTree::TreeRef ref) const {
    auto nr = lt.node(ref);
    const auto &n = lt.ir.node(nr);
    ASSERT(n.inputs().size() == 1) << "memory node with several inputs";
    return access(nr, ref) + " = " + access(n.inputs()[0], ref) + ";";
  }

  static std::string float_literal(float v) {
    // same text as `stream << float` (6 significant digits)
    std::stringstream s;
    s << v;
    return s.str();
  }

  // cpp.cpp:122-170: declare/initialise the allocations whose LCA is `ref`
  std::string resets(LoopTree::TreeRef ref) const {
    std::stringstream ss;
    auto prefix = indent(ref, 1);
    auto init = [&](const Node &node) -> float {
      switch (node.op()) {
        case Operation::add:
          return 0;
        case Operation::multiply:
          return 1;
        case Operation::min:
          return std::numeric_limits<float>::max();
        case Operation::max:
          return -std::numeric_limits<float>::max();
        case Operation::write:
        case Operation::view:
          return 0;
        default:
          ASSERT(0) << "cannot find default value for " << dump(node.op());
      }
      return -1;
    };
    // allocation map iteration as the reference: unordered_map over NodeRef
    for (const auto &p : allocations) {
      const auto &alloc = p.second;
      if (alloc.lca != ref) continue;
      const auto &node = lt.ir.node(alloc.node_ref);
      bool needs_set = !lt.ir.reduction_vars(alloc.node_ref).empty() &&
                       node.op() != Operation::view;
      for (auto inp : node.inputs())
        if (lt.ir.node(inp).op() == Operation::view && !lt.scheduled.count(inp))
          needs_set = true;
      if (!lt.scheduled.count(alloc.node_ref)) continue;
      if (alloc.size() == 1 && !is_io(alloc.node_ref)) {
        ss << prefix << "float v" << alloc.mem_idx;
        if (needs_set) ss << " = " << float_literal(init(node));
        ss << ";\n";
      } else if (needs_set) {
        set_called = true;
        ss << prefix << "set((float*)memory[" << alloc.mem_idx << "], "
           << float_literal(init(node)) << ", " << alloc.size() << ");\n";
      }
    }
    return ss.str();
  }

  // cpp.cpp:172-270
  std::string compute_node(LoopTree::TreeRef ref) const {
    auto nr = lt.node(ref);
    const auto &node = lt.ir.node(nr);
    auto op = node.op();
    bool infix = op == Operation::add || op == Operation::multiply ||
                 op == Operation::subtract || op == Operation::divide;
    bool binary = infix || op == Operation::min || op == Operation::max;
    const char *sym = "";
    switch (op) {
      case Operation::add: sym = "+"; break;
      case Operation::multiply: sym = "*"; break;
      case Operation::subtract: sym = "-"; break;
      case Operation::divide: sym = "/"; break;
      case Operation::max: sym = "max"; break;
      case Operation::min: sym = "min"; break;
      case Operation::log: sym = "log"; break;
      case Operation::exp: sym = "exp"; break;
      case Operation::sqrt: sym = "sqrt"; break;
      case Operation::negate: sym = "-"; break;
      case Operation::abs: sym = "abs"; break;
      case Operation::reciprocal: sym = "1 / "; break;
      default:
        ASSERT(0) << "can't emit code for " << dump(op);
    }
    std::vector<std::string> operands;
    if (!lt.ir.reduction_vars(nr).empty()) operands.emplace_back(access(nr, ref));
    for (auto inp : node.inputs()) operands.emplace_back(access(inp, ref));
    std::stringstream ss;
    ss << access(nr, ref) << " = ";
    if (infix) {
      for (size_t i = 0; i < operands.size(); ++i) {
        if (i) ss << " " << sym << " ";
        ss << operands[i];
      }
    } else if (binary) {
      std::function<void(size_t)> nest = [&](size_t i) {
        if (i + 1 == operands.size()) {
          ss << operands[i];
          return;
        }
        ss << sym << "(" << operands[i] << ", ";
        nest(i + 1);
        ss << ")";
      };
      nest(0);
    } else {
      ASSERT(operands.size() == 1) << "unary op with several operands";
      ss << sym << "(" << operands[0] << ")";
    }
    ss << ";";
    return ss.str();
  }

  std::string node_string(LoopTree::TreeRef ref) const {
    std::stringstream ss;
    const auto &node = lt.ir.node(lt.node(ref));
    if (lt.children(lt.parent(ref)).at(0) == ref) ss << resets(lt.parent(ref));
    ss << indent(ref);
    switch (node.op()) {
      case Operation::write:
      case Operation::view:
        ss << mem_node(ref);
        break;
      case Operation::read:
        break;
      default:
        ss << compute_node(ref);
    }
    ss << " // %" << lt.node(ref) << " (" << dump(node.op()) << ")\n";
    return ss.str();
  }

  // cpp.cpp:295-351: loop with optional tail block; a tail re-emits the body
  // with the inner loops of the same var sized by the override
  std::string loop_string(LoopTree::TreeRef ref,
                          std::unordered_map<IR::VarRef, int> overrides) const {
    std::stringstream ss;
    auto prefix = indent(ref);
    auto loop = lt.loop(ref);
    std::string it = "i_" + std::to_string(ref);
    int64_t size = loop.size, tail = loop.tail;
    if (overrides.count(loop.var)) {
      auto o = overrides.at(loop.var);
      auto inner = inner_sizes.at(ref);
      size = o / inner;
      tail = o % inner;
      overrides.erase(loop.var);
    }
    std::vector<std::string> body, tail_body;
    for (auto c : lt.children(ref)) body.emplace_back(gen(c, overrides));
    if (tail > 0) {
      overrides[loop.var] = (int)tail;
      for (auto c : lt.children(ref)) tail_body.emplace_back(gen(c, overrides));
    }
    if (lt.children(lt.parent(ref)).at(0) == ref) ss << resets(lt.parent(ref));
    ss << prefix << "for (int64_t " << it << " = 0L; " << it << " < " << size
       << "L; ++" << it << ") { // " << lt.ir.var(loop.var).name() << "\n";
    for (const auto &b : body) ss << b;
    ss << prefix << "}\n";
    if (tail > 0) {
      ss << prefix << "{ // " << lt.ir.var(loop.var).name() << " tail\n";
      // pinned to the loop's own size even where an override shrank it, as
      // cpp.cpp:340-341 does: kept, since this file restates the reference
      // (inside another tail block that index runs past the buffers)
      ss << indent(ref, 1) << "int64_t " << it << " = " << loop.size << "L;\n";
      for (const auto &b : tail_body) ss << b;
      ss << prefix << "}\n";
    }
    return ss.str();
  }

  // cpp.cpp:353-448
  std::string gen(LoopTree::TreeRef ref,
                  std::unordered_map<IR::VarRef, int> overrides) const {
    if (ref != -1) {
      if (lt.kind(ref) == LoopTree::NODE) return node_string(ref);
      return loop_string(ref, overrides);
    }
    std::stringstream body;
    for (auto c : lt.roots) body << gen(c, {});
    bool need_max = false, need_min = false;
    for (auto n : lt.ir.nodes()) {
      need_max |= lt.ir.node(n).op() == Operation::max;
      need_min |= lt.ir.node(n).op() == Operation::min;
    }
    std::stringstream ss;
    ss << "#include <math.h>\n#include <stdint.h>\n#include <stdio.h>\n"
          "#include <stdlib.h>\n\n";
    // statement-expression select: the accumulator is `a`, so a NaN or -0.0
    // operand behaves exactly as `a > b ? a : b` in that order
    if (need_max)
      ss << "\n#define max(a,b) \\\n   ({ __typeof__ (a) _a = (a); \\\n"
            "       __typeof__ (b) _b = (b); \\\n     _a > _b ? _a : _b; })\n";
    if (need_min)
      ss << "\n#define min(a,b) \\\n   ({ __typeof__ (a) _a = (a); \\\n"
            "       __typeof__ (b) _b = (b); \\\n     _a < _b ? _a : _b; })\n";
    if (set_called)
      ss << "\nstatic inline void set(float* mem, float val, int64_t length) {\n"
            "  for (int64_t i = 0; i < length; ++i) {\n    mem[i] = val;\n  }\n}\n";
    ss << "\n";
    auto sizes = memory_sizes(true);
    auto n_in = lt.ir.inputs().size(), n_out = lt.ir.outputs().size();
    auto entry = [&](size_t idx, int64_t s, bool force) {
      ss << idx << ":";
      if (s <= 1 && !force)
        ss << "nullptr";
      else
        ss << "float[" << s << "]";
      ss << ", ";
    };
    ss << "// memory: {\n//   ";
    size_t i = 0;
    for (; i < n_in; ++i) entry(i, sizes.at(i), true);
    ss << "// inputs\n//   ";
    for (; i < n_in + n_out; ++i) entry(i, sizes.at(i), true);
    ss << "// outputs\n//   ";
    for (; i < sizes.size(); ++i) entry(i, sizes[i], false);
    ss << "// scratch\n// }\n";
    ss << "void fn_" << count << "(void** memory) {\n" << body.str() << "}\n";
    return ss.str();
  }
};

// cpp.cpp:450-503: compile with the system C compiler and dlopen.  Files go
// to a per-process directory so concurrent test processes do not collide.
struct CCompiled : public Compiled {
  std::string code;
  std::shared_ptr<DynamicLibrary> dll;
  void (*fn)(void **) = nullptr;
  mutable std::vector<void *> mem;
  std::vector<int64_t> mem_sizes;

  explicit CCompiled(const LoopTree &lt) {
    CCodegen gen(lt);
    code = gen.gen_string();
    std::string fn_name = "fn_" + std::to_string(gen.count);
    std::string dir = "/tmp/lt_oracle_" + std::to_string(getpid());
    std::string mk = "mkdir -p " + dir;
    ASSERT(!std::system(mk.c_str())) << "cannot create " << dir;
    std::string src = dir + "/" + fn_name + ".c";
    std::string lib = dir + "/" + fn_name + ".so";
    std::ofstream(src, std::ios::trunc) << code;
    std::string cmd =
        "cc -Wall -Wno-unused-function -Wno-unused-variable -Werror -O3 "
        "-std=c99 -fpic -shared -o " + lib + " " + src;
    ASSERT(!std::system(cmd.c_str())) << "C oracle failed to compile:\n" << code;
    dll = std::make_shared<DynamicLibrary>(lib.c_str());
    fn = dll->sym<void (*)(void **)>(fn_name.c_str());
    std::remove(src.c_str());
    mem_sizes = gen.memory_sizes();
    mem.assign(mem_sizes.size(), nullptr);
    for (size_t i = 0; i < mem_sizes.size(); ++i)
      if (mem_sizes[i] > 0) mem[i] = calloc(mem_sizes[i], sizeof(float));
  }

  ~CCompiled() {
    for (size_t i = 0; i < mem_sizes.size(); ++i)
      if (mem_sizes[i] > 0) free(mem[i]);
  }

  void run(const std::vector<void *> &memory, bool sync) const override {
    for (size_t i = 0; i < memory.size(); ++i) mem[i] = memory[i];
    fn(mem.data());
  }

  std::string dump() const override { return code; }
};

struct CBackend : public Backend {
  CBackend() : Backend("cpp") {}
  std::unique_ptr<Compiled> compile_impl(const LoopTree &lt) const override {
    return std::make_unique<CCompiled>(lt);
  }
  int hardware_requirement() const override { return 1 << 0; }
};

static RegisterBackend register_cpp(std::make_shared<CBackend>());

}  // namespace oracle
}  // namespace loop_tool
