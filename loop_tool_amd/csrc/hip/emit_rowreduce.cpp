// Row-cooperative lowering for nests shaped like a row softmax / row norm:
// the output's innermost var `c` is reduced by one or more nodes whose
// results are broadcast back along `c` (e.g. max, subtract, exp, sum,
// divide over the columns of each row).
//
// One workgroup owns one row.  Lanes own columns c = (k*T + lane)*VEC + j
// (VEC-wide vector loads, so every wave reads contiguous 16-byte chunks;
// plain loads, which measure ~10% faster here than nontemporal ones, and
// nontemporal stores);
// the row's inputs are loaded from HBM once into registers, every column
// node is evaluated in registers, and every row reduction is done as a
// per-lane partial, a wavefront butterfly (cross-lane shuffles), and a
// cross-wave combine through LDS.  The output row is written once.  HBM
// traffic is therefore one read of each input row and one write of the
// output row: the algorithmic minimum.
//
// The tree-shaped reduction order differs from the sequential order of the
// reference C backend, so results agree to rounding (checked at <= 1e-5
// normwise), not bit for bit; max/min reductions are exact.  To keep the
// kernel HBM-bound rather than ALU-bound, exp is a 2-ulp exp2-based form
// (lt_exp) instead of the libm-accurate expf, and a division of a column by
// a row scalar is a multiply by the row's reciprocal (1 ulp).  The exact
// pointwise lowering ("hip:plain") keeps the reference's operations.
#include <functional>
#include <set>
#include <sstream>

#include "emit_internal.h"

namespace loop_tool {
namespace hip {

using symbolic::Expr;
using symbolic::Symbol;

namespace {

bool is_identity(const DFInput &in, const Dataflow &df, const std::vector<IR::VarRef> &vars) {
  const auto &src = df.nodes.at(in.src);
  if (in.coords.size() != src.vars.size()) return false;
  for (size_t d = 0; d < in.coords.size(); ++d) {
    const auto &e = in.coords[d];
    if (e.type() != Expr::Type::symbol) return false;
    if (!(e.symbol() == df.sym(src.vars[d]))) return false;
  }
  return true;
}

const char *combine_op(Operation op) {
  switch (op) {
    case Operation::add: return "add";
    case Operation::multiply: return "mul";
    case Operation::max: return "max";
    case Operation::min: return "min";
    default: return nullptr;
  }
}

}  // namespace

bool try_rowreduce(const Dataflow &df, const Knobs &knobs, Plan &plan) {
  if (df.writes.size() != 1) return false;
  const auto &W = df.nodes.at(df.writes[0]);
  if (W.vars.empty()) return false;
  const IR::VarRef cvar = W.vars.back();
  std::vector<IR::VarRef> rvars(W.vars.begin(), W.vars.end() - 1);
  std::set<IR::VarRef> rowset(rvars.begin(), rvars.end());
  const int64_t C = df.size(cvar);
  const int64_t rows = df.numel(rvars);
  if (C < 128 || rows < 1) return false;

  // classify every compute node reachable from the write
  enum Kind { COLUMN, ROW_SCALAR, ROW_REDUCE };
  std::map<IR::NodeRef, Kind> kind;
  bool any_reduce = false;
  for (auto nr : df.topo) {
    const auto &n = df.nodes.at(nr);
    if (n.op == Operation::write) continue;
    bool has_c = false;
    for (auto v : n.vars) {
      if (v == cvar) has_c = true;
      else if (!rowset.count(v)) return false;
    }
    if (!n.red_vars.empty()) {
      if (n.red_vars.size() != 1 || n.red_vars[0] != cvar || has_c) return false;
      if (!combine_op(n.op)) return false;
      kind[nr] = ROW_REDUCE;
      any_reduce = true;
    } else {
      kind[nr] = has_c ? COLUMN : ROW_SCALAR;
    }
    // compute-node inputs must be read at their own coordinates
    for (const auto &in : n.inputs) {
      const auto &src = df.nodes.at(in.src);
      if (src.op != Operation::read && !is_identity(in, df, src.vars)) return false;
      if (src.op == Operation::read) {
        for (const auto &e : in.coords)
          for (const auto &s : e.symbols()) {
            bool ok = s == df.sym(cvar);
            for (auto v : rvars) ok |= s == df.sym(v);
            if (!ok) return false;
          }
      }
    }
  }
  if (!any_reduce) return false;
  if (!is_identity(W.inputs[0], df, W.vars)) {
    if (df.nodes.at(W.inputs[0].src).op != Operation::read) return false;
  }

  // launch shape: T lanes x VEC elements x NK chunks covers the row
  // the schedule's split of the column var over the write: its innermost
  // extent (1, 2 or 4) sets the values per vector, the loop above it (64 to
  // 1024, a multiple of 64) the lanes per row
  const int csym = df.sym(cvar).id();
  const int64_t s_vec = df.split_extent(W.nr, csym), s_lanes = df.split_extent(W.nr, csym, 1);
  int vec = knobs.vec > 0 ? knobs.vec
            : (s_vec == 1 || s_vec == 2 || s_vec == 4) ? (int)s_vec
                                                       : (C % 4 == 0 ? 4 : 1);
  if (vec != 1 && vec != 2 && vec != 4) vec = 1;
  if (C % vec) vec = 1;
  int T = knobs.block;
  if (!knobs.block_set && s_lanes >= 64 && s_lanes <= 1024 && s_lanes % 64 == 0)
    T = (int)s_lanes;
  else if (!knobs.block_set) {
    // default: two vectors (8 values) per lane, as many lanes as that needs
    // up to 1024 (measured fastest for rows of 1K-8K floats: more rows in
    // flight per CU hide the two reductions' barriers)
    int64_t want = (C + vec * 2 - 1) / (vec * 2);
    T = 64;
    while (T < 1024 && T < want) T *= 2;
  }
  const int64_t NK = (C + (int64_t)T * vec - 1) / ((int64_t)T * vec);
  if (NK * vec > 64) return false;  // too long for registers
  const bool full = (C % ((int64_t)T * vec)) == 0;

  std::stringstream k;
  k << kernel_prelude();
  k << R"(
// e^x as 2^(x log2 e): the product's rounding error (an fma) and the low
// part of log2 e are folded back as a first-order factor around the
// hardware 2^t (v_exp_f32); about 2 ulp, overflow to +inf kept.
__device__ __forceinline__ float lt_exp(float x) {
  const float t = x * 1.44269502f;
  const float err = __builtin_fmaf(x, 1.44269502f, -t) + x * 1.92596303e-8f;
  const float r = __builtin_amdgcn_exp2f(t);
  return r == __builtin_inff() ? r : __builtin_fmaf(r, err * 0.693147181f, r);
}
__device__ __forceinline__ float lt_comb_add(float a, float b) { return a + b; }
__device__ __forceinline__ float lt_comb_mul(float a, float b) { return a * b; }
__device__ __forceinline__ float lt_comb_max(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float lt_comb_min(float a, float b) { return a < b ? a : b; }
)";
  const int waves = T / 64;
  k << "#define LT_T " << T << "\n#define LT_VEC " << vec << "\n#define LT_NK " << NK
    << "\n#define LT_C " << C << "L\n#define LT_WAVES " << waves << "\n";
  // wavefront butterfly + cross-wave combine through LDS
  for (const char *op : {"add", "mul", "max", "min"}) {
    k << "__device__ __forceinline__ float lt_block_" << op
      << "(float v, float* lds) {\n"
         "#pragma unroll\n"
         "  for (int o = 32; o > 0; o >>= 1) v = lt_comb_"
      << op << "(v, __shfl_xor(v, o, 64));\n"
         "  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;\n"
         "  if (LT_WAVES == 1) return v;\n"
         "  if (lane == 0) lds[wave] = v;\n"
         "  __syncthreads();\n"
         "  float r = lds[0];\n"
         "#pragma unroll\n"
         "  for (int w = 1; w < LT_WAVES; ++w) r = lt_comb_"
      << op << "(r, lds[w]);\n"
         "  return r;\n}\n";
  }

  std::stringstream sig;
  for (int i = 0; i < df.num_inputs; ++i) sig << "const float* __restrict__ in" << i << ", ";
  sig << "float* __restrict__ out";

  std::stringstream b;  // kernel body
  b << "  const long row = blockIdx.x;\n  const int lane = threadIdx.x;\n";
  std::map<int, std::string> names;  // symbol id -> C name
  {
    std::string t = "row";
    b << "  long rt = row;\n";
    for (int d = (int)rvars.size() - 1; d >= 0; --d) {
      auto nm = "r" + std::to_string(d);
      names[df.sym(rvars[d]).id()] = nm;
      if (d > 0)
        b << "  const long " << nm << " = rt % " << df.size(rvars[d]) << "L; rt /= "
          << df.size(rvars[d]) << "L;\n";
      else
        b << "  const long " << nm << " = rt;\n";
    }
  }
  names[df.sym(cvar).id()] = "c";
  auto print = [&](const Expr &e) {
    std::unordered_map<Symbol, std::string, symbolic::Hash<Symbol>> repl;
    for (const auto &s : e.symbols()) repl[s] = names.at(s.id());
    return "(" + e.simplify().dump(false, repl) + ")";
  };
  std::map<int, int64_t> extents;
  for (auto v : rvars) extents[df.sym(v).id()] = df.size(v);
  extents[df.sym(cvar).id()] = C;

  int lds_slots = 0;
  std::map<IR::NodeRef, std::string> scalar;  // row scalars
  std::map<IR::NodeRef, std::string> column;  // column arrays v_<n>[NK*VEC]
  int tmp = 0;
  // element loop helper: body sees `c` (column index) and `e` (register slot)
  auto each = [&](const std::string &body) {
    std::stringstream s;
    s << "  #pragma unroll\n  for (int kk = 0; kk < LT_NK; ++kk) {\n"
      << "    #pragma unroll\n    for (int j = 0; j < LT_VEC; ++j) {\n"
      << "      const int e = kk * LT_VEC + j;\n"
      << "      const long c = ((long)kk * LT_T + lane) * LT_VEC + j;\n"
      << (full ? "" : "      if (c < LT_C) {\n") << body << (full ? "" : "      }\n")
      << "    }\n  }\n";
    return s.str();
  };

  // value of a read at column-or-row coordinates, as an expression usable
  // inside `each` (column reads) or at row scope (row reads)
  auto read_expr = [&](const DFInput &in) {
    const auto &src = df.nodes.at(in.src);
    std::vector<std::string> conds;
    Expr idx(0);
    int64_t stride = 1;
    for (int d = (int)src.vars.size() - 1; d >= 0; --d) {
      idx = idx + in.coords[d] * Expr(stride);
      auto r = expr_range(in.coords[d], extents);
      auto sz = df.size(src.vars[d]);
      auto s = print(in.coords[d]);
      if (!r.ok || r.lo < 0) conds.push_back(s + " >= 0");
      if (!r.ok || r.hi > sz - 1) conds.push_back(s + " < " + std::to_string(sz) + "L");
      stride *= sz;
    }
    std::string load = "in" + std::to_string(src.input_index) + "[" + print(idx) + "]";
    if (conds.empty()) return load;
    std::string p;
    for (size_t i = 0; i < conds.size(); ++i) p += (i ? " && " : "") + conds[i];
    return "((" + p + ") ? " + load + " : 0.0f)";
  };
  auto depends_on_c = [&](const DFInput &in) {
    for (const auto &e : in.coords)
      if (e.contains(df.sym(cvar))) return true;
    return false;
  };
  // a read whose innermost coordinate is exactly c and whose other
  // coordinates do not involve c is loaded with vector loads
  auto contiguous_read = [&](const DFInput &in) {
    const auto &src = df.nodes.at(in.src);
    if (src.op != Operation::read || src.vars.empty()) return false;
    const auto &last = in.coords.back();
    if (!(last.type() == Expr::Type::symbol && last.symbol() == df.sym(cvar))) return false;
    if (df.size(src.vars.back()) != C) return false;
    for (size_t d = 0; d + 1 < in.coords.size(); ++d)
      if (in.coords[d].contains(df.sym(cvar))) return false;
    return true;
  };

  std::map<std::string, std::string> loaded;  // read expr key -> array name
  // returns a C expression for input `in` inside `each` (slot e, column c)
  std::function<std::string(const DFInput &)> operand = [&](const DFInput &in) -> std::string {
    const auto &src = df.nodes.at(in.src);
    if (src.op == Operation::read) {
      if (!depends_on_c(in)) {
        std::string key = "row:" + read_expr(in);
        if (!loaded.count(key)) {
          auto nm = "q" + std::to_string(tmp++);
          b << "  const float " << nm << " = " << read_expr(in) << ";\n";
          loaded[key] = nm;
        }
        return loaded[key];
      }
      std::string key = read_expr(in);
      if (!loaded.count(key)) {
        auto nm = "x" + std::to_string(tmp++);
        b << "  float " << nm << "[LT_NK * LT_VEC];\n";
        if (contiguous_read(in) && vec > 1) {
          // row base = index at c = 0
          Expr base(0);
          int64_t stride = 1;
          for (int d = (int)src.vars.size() - 1; d >= 0; --d) {
            if (d != (int)src.vars.size() - 1) base = base + in.coords[d] * Expr(stride);
            stride *= df.size(src.vars[d]);
          }
          std::string vt = vec == 4 ? "lt_f4" : "lt_f2";
          b << "  {\n    const " << vt << "* p = (const " << vt << "*)(in"
            << src.input_index << " + " << print(base) << ");\n"
            << "    #pragma unroll\n    for (int kk = 0; kk < LT_NK; ++kk) {\n"
            << "      const long vi = (long)kk * LT_T + lane;\n";
          if (!full) b << "      if (vi * LT_VEC < LT_C) {\n";
          b << "      const " << vt << " t = p[vi];\n";
          const char *comp[4] = {"x", "y", "z", "w"};
          for (int j = 0; j < vec; ++j)
            b << "      " << nm << "[kk * LT_VEC + " << j << "] = t." << comp[j] << ";\n";
          if (!full) b << "      }\n";
          b << "    }\n  }\n";
        } else {
          b << each("      " + nm + "[e] = " + read_expr(in) + ";\n");
        }
        loaded[key] = nm;
      }
      return loaded[key] + "[e]";
    }
    if (kind.at(in.src) == COLUMN) return column.at(in.src) + "[e]";
    return scalar.at(in.src);
  };
  auto combine = [&](Operation op, const std::vector<std::string> &xs) -> std::string {
    switch (op) {
      case Operation::add:
      case Operation::subtract:
      case Operation::multiply:
      case Operation::divide: {
        const char *sym = op == Operation::add ? " + " : op == Operation::subtract ? " - "
                          : op == Operation::multiply ? " * " : " / ";
        std::string s = "(";
        for (size_t i = 0; i < xs.size(); ++i) s += (i ? sym : "") + xs[i];
        return s + ")";
      }
      case Operation::max:
      case Operation::min: {
        std::function<std::string(size_t)> nest = [&](size_t i) -> std::string {
          if (i + 1 == xs.size()) return xs[i];
          return std::string(op == Operation::max ? "lt_selmax(" : "lt_selmin(") + xs[i] +
                 ", " + nest(i + 1) + ")";
        };
        return nest(0);
      }
      case Operation::exp:
        return "lt_exp(" + xs[0] + ")";
      case Operation::log:
        return "logf(" + xs[0] + ")";
      case Operation::sqrt:
        return "sqrtf(" + xs[0] + ")";
      default:
        return unary_expr(op, xs[0]);
    }
  };

  std::stringstream desc;
  desc << "[hip] row-cooperative lowering: " << rows << " rows x " << C
       << " columns, " << T << " lanes x " << vec << " wide x " << NK << " chunks\n";
  for (auto nr : df.topo) {
    const auto &n = df.nodes.at(nr);
    if (n.op == Operation::write) continue;
    auto kd = kind.at(nr);
    if (kd == ROW_REDUCE) {
      auto part = "p" + std::to_string(tmp++);
      std::string init = n.op == Operation::add ? "0.0f" : n.op == Operation::multiply
                         ? "1.0f" : n.op == Operation::max ? float_literal(-3.40282e+38)
                         : float_literal(3.40282e+38);
      b << "  float " << part << " = " << init << ";\n";
      std::vector<std::string> xs{part};
      std::string body;
      {
        // operands are evaluated inside the element loop
        std::vector<std::string> ops;
        for (const auto &in : n.inputs) ops.push_back(operand(in));
        xs.insert(xs.end(), ops.begin(), ops.end());
      }
      b << each("      " + part + " = " + combine(n.op, xs) + ";\n");
      auto nm = "s" + std::to_string(tmp++);
      b << "  const float " << nm << " = lt_block_" << combine_op(n.op) << "(" << part
        << ", lds + " << lds_slots << ");\n";
      lds_slots += waves;
      scalar[nr] = nm;
      desc << "  row reduction " << df.ir().dump(nr) << "\n";
    } else if (kd == ROW_SCALAR) {
      std::vector<std::string> xs;
      for (const auto &in : n.inputs) xs.push_back(operand(in));
      auto nm = "s" + std::to_string(tmp++);
      b << "  const float " << nm << " = " << combine(n.op, xs) << ";\n";
      scalar[nr] = nm;
    } else {
      std::vector<std::string> xs;
      for (const auto &in : n.inputs) xs.push_back(operand(in));
      Operation op = n.op;
      if (op == Operation::divide && xs.size() == 2 &&
          (xs[1].size() < 3 || xs[1].compare(xs[1].size() - 3, 3, "[e]") != 0)) {
        // column / row scalar: one reciprocal per row, a multiply per element
        auto rn = "r" + std::to_string(tmp++);
        b << "  const float " << rn << " = 1.0f / " << xs[1] << ";\n";
        xs[1] = rn;
        op = Operation::multiply;
      }
      auto nm = "v" + std::to_string(tmp++);
      b << "  float " << nm << "[LT_NK * LT_VEC];\n";
      b << each("      " + nm + "[e] = " + combine(op, xs) + ";\n");
      column[nr] = nm;
    }
  }
  // output row: out[row * C + c]
  std::string val = operand(W.inputs[0]);
  if (vec > 1) {
    std::string vt = vec == 4 ? "lt_f4" : "lt_f2";
    std::string arr = val.substr(0, val.size() - 3);  // strip "[e]"
    const char *comp[4] = {"x", "y", "z", "w"};
    b << "  {\n    " << vt << "* p = (" << vt << "*)(out + row * LT_C);\n"
      << "    #pragma unroll\n    for (int kk = 0; kk < LT_NK; ++kk) {\n"
      << "      const long vi = (long)kk * LT_T + lane;\n";
    if (!full) b << "      if (vi * LT_VEC < LT_C) {\n";
    b << "      " << vt << " t;\n";
    for (int j = 0; j < vec; ++j)
      b << "      t." << comp[j] << " = " << arr << "[kk * LT_VEC + " << j << "];\n";
    b << "      __builtin_nontemporal_store(t, p + vi);\n";
    if (!full) b << "      }\n";
    b << "    }\n  }\n";
  } else {
    b << each("      out[row * LT_C + c] = " + val + ";\n");
  }

  k << "extern \"C\" __global__ void __launch_bounds__(" << T << ") lt_row0(" << sig.str()
    << ") {\n  __shared__ float lds[" << std::max(lds_slots, 1) << "];\n"
    << b.str() << "}\n";

  Stage st;
  st.kernel = "lt_row0";
  st.kind = "rowreduce";
  for (int i = 0; i < df.num_inputs; ++i) st.buffers.push_back(i);
  st.buffers.push_back(df.num_inputs);
  st.writes.push_back(df.num_inputs);
  ASSERT(rows < (1LL << 31)) << "too many rows";
  st.grid[0] = (unsigned)rows;
  st.block[0] = T;
  plan.config["vec"] = vec;
  plan.config["block"] = T;
  int64_t bytes = rows * C * 4;  // output
  for (int i = 0; i < df.num_inputs; ++i)
    bytes += df.numel(df.nodes.at(df.ir().inputs()[i]).vars) * 4;
  st.bytes = bytes;
  plan.stages.push_back(st);
  plan.source = k.str();
  plan.description = desc.str();
  return true;
}

}  // namespace hip
}  // namespace loop_tool
