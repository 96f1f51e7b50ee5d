// Contraction lowering on the MFMA matrix cores.
//
// Applies to nests that really are dense contractions:
//     write <- add-reduce over K <- multiply(X, Y)
// where X and Y are IR inputs read through affine (possibly predicated) view
// indices -- matmul, batched matmul, and convolution through its pad and
// window views (the generalisation of the reference's `is_fma_nest`,
// src/backends/cpu/loop_nest.cpp:41-86, to views).  The output vars split
// into I (only in one operand), J (only in the other) and batch vars (in
// both); every reduced var must index both operands.  The kernel computes
//     D[b][i][j] = sum_k P[b][i][k] * Q[b][k][j]
// as an implicit GEMM: P/Q tiles are gathered from the operands with the
// view's affine index (out-of-range elements read 0, as `cond ? x : 0` in
// src/backends/cpu/cpp.cpp:64-106), converted to the MFMA operand type,
// staged in LDS (double-buffered, one barrier per K tile) and multiplied with
// v_mfma_f32_32x32x16_{bf16,f16} or v_mfma_f32_32x32x2_f32 with float32
// accumulation.  The operand whose vars include the output's contiguous var
// provides the MFMA columns, so each 32-lane half-wave stores 128 contiguous
// bytes of output.
//
// Accumulation order differs from the reference's sequential sum, so FP32
// results match the C oracle to rounding (<= 1e-5 normwise); BF16/FP16
// operands are rounded once (RNE) from FP32 and their error is stated
// against a float64 evaluation.
#include <algorithm>
#include <functional>
#include <set>
#include <sstream>

#include "contract.h"
#include "emit_internal.h"

namespace loop_tool {
namespace hip {

using symbolic::Expr;
using symbolic::Op;
using symbolic::Symbol;

namespace {

// e = sum coef[v] * v + c over symbols; false when not affine
bool affine(const Expr &e0, std::map<int, int64_t> &coef, int64_t &c) {
  std::function<bool(const Expr &, int64_t)> rec = [&](const Expr &x, int64_t s) -> bool {
    if (x.type() == Expr::Type::value) {
      c += s * x.value();
      return true;
    }
    if (x.type() == Expr::Type::symbol) {
      coef[x.symbol().id()] += s;
      return true;
    }
    switch (x.op()) {
      case Op::add:
        return rec(x.arg(0), s) && rec(x.arg(1), s);
      case Op::negate:
        return rec(x.arg(0), -s);
      case Op::multiply:
        if (x.arg(1).type() == Expr::Type::value) return rec(x.arg(0), s * x.arg(1).value());
        if (x.arg(0).type() == Expr::Type::value) return rec(x.arg(1), s * x.arg(0).value());
        return false;
      default:
        return false;
    }
  };
  c = 0;
  coef.clear();
  bool ok = rec(e0.simplify(), 1);
  for (auto it = coef.begin(); it != coef.end();)
    it = it->second == 0 ? coef.erase(it) : std::next(it);
  return ok;
}

}  // namespace

bool try_contract(const Dataflow &df, const Knobs &knobs_in, Plan &plan) {
  Knobs knobs = knobs_in;
  if (df.writes.size() != 1) return false;
  const auto &W = df.nodes.at(df.writes[0]);
  const auto &win = W.inputs.at(0);
  const auto &R = df.nodes.at(win.src);
  if (R.op != Operation::add || R.red_vars.empty() || R.inputs.size() != 1) return false;
  // the write must read R at a permutation of its own vars
  std::map<int, int> wpos;  // R var sym id -> W dim
  for (size_t d = 0; d < win.coords.size(); ++d) {
    const auto &e = win.coords[d];
    if (e.type() != Expr::Type::symbol) return false;
    for (size_t w = 0; w < W.vars.size(); ++w)
      if (e.symbol() == df.sym(W.vars[w])) wpos[df.sym(R.vars[d]).id()] = (int)w;
  }
  if (wpos.size() != R.vars.size() || R.vars.size() != W.vars.size()) return false;
  const auto &Min = R.inputs[0];
  const auto &Mn = df.nodes.at(Min.src);
  if (Mn.op != Operation::multiply || Mn.inputs.size() != 2 || !Mn.red_vars.empty())
    return false;
  for (size_t d = 0; d < Min.coords.size(); ++d) {
    const auto &e = Min.coords[d];
    if (e.type() != Expr::Type::symbol || !(e.symbol() == df.sym(Mn.vars[d]))) return false;
  }
  std::map<int, int64_t> extents;
  for (auto v : Mn.vars) extents[df.sym(v).id()] = df.size(v);

  Operand ops[2];
  std::set<int> used[2];
  for (int o = 0; o < 2; ++o) {
    const auto &in = Mn.inputs[o];
    const auto &src = df.nodes.at(in.src);
    if (src.op != Operation::read) return false;
    auto &op = ops[o];
    op.input_index = src.input_index;
    int64_t st = 1;
    op.sizes.resize(src.vars.size());
    op.strides.resize(src.vars.size());
    for (int d = (int)src.vars.size() - 1; d >= 0; --d) {
      op.sizes[d] = df.size(src.vars[d]);
      op.strides[d] = st;
      st *= op.sizes[d];
    }
    for (size_t d = 0; d < in.coords.size(); ++d) {
      std::map<int, int64_t> cf;
      int64_t c = 0;
      if (!affine(in.coords[d], cf, c)) return false;
      for (const auto &p : cf) used[o].insert(p.first);
      op.coef.push_back(cf);
      op.cst.push_back(c);
      auto r = expr_range(in.coords[d], extents);
      if (!r.ok) return false;
      op.lo.push_back(r.lo);
      op.hi.push_back(r.hi);
      if (r.lo < 0 || r.hi > op.sizes[d] - 1) op.pred_dims.push_back((int)d);
    }
  }
  // classify output vars
  std::vector<IR::VarRef> side[2], batch;
  for (auto v : W.vars) {  // W order: row-major output
    int s = df.sym(v).id();
    bool a = used[0].count(s), b = used[1].count(s);
    if (a && b) batch.push_back(v);
    else if (a) side[0].push_back(v);
    else if (b) side[1].push_back(v);
    else return false;
  }
  for (auto v : R.red_vars) {
    int s = df.sym(v).id();
    if (!used[0].count(s) || !used[1].count(s)) return false;
  }
  if (side[0].empty() || side[1].empty()) return false;
  // the operand owning the output's contiguous var provides the columns (Q)
  int qo = std::find(side[0].begin(), side[0].end(), W.vars.back()) != side[0].end() ? 0 : 1;
  if (std::find(side[qo].begin(), side[qo].end(), W.vars.back()) == side[qo].end())
    return false;  // contiguous var is a batch var
  int po = 1 - qo;

  auto infos = [&](const std::vector<IR::VarRef> &vs, const std::string &prefix) {
    std::vector<VarInfo> out;
    for (size_t i = 0; i < vs.size(); ++i)
      out.push_back({df.sym(vs[i]).id(), df.size(vs[i]), prefix + std::to_string(i)});
    return out;
  };
  std::vector<VarInfo> I = infos(side[po], "vi"), J = infos(side[qo], "vj"),
                       B = infos(batch, "vb");
  // K order: as the P operand's dims list them (its innermost var last)
  std::vector<IR::VarRef> kvars;
  {
    std::set<int> red;
    for (auto v : R.red_vars) red.insert(df.sym(v).id());
    for (const auto &cf : ops[po].coef)
      for (auto v : R.red_vars)
        if (cf.count(df.sym(v).id()) &&
            std::find(kvars.begin(), kvars.end(), v) == kvars.end())
          kvars.push_back(v);
    for (auto v : R.red_vars)
      if (std::find(kvars.begin(), kvars.end(), v) == kvars.end()) kvars.push_back(v);
  }
  std::vector<VarInfo> K = infos(kvars, "vk");
  // tile sides from the schedule: a single row (column) var split into
  // nested loops above the reduction gives the tile's rows (columns) as its
  // innermost extent (64 to 256, a multiple of 64); a tile=MxN annotation
  // takes precedence
  if (knobs.tile_m == 0) {
    auto side = [&](const std::vector<VarInfo> &U) -> int {
      if (U.size() != 1) return 0;
      const int64_t s = df.split_extent(R.nr, U[0].sym);
      return s >= 64 && s <= 256 && s % 64 == 0 ? (int)s : 0;
    };
    const int sm = side(I), sn = side(J);
    if (sm || sn) {
      knobs.tile_m = sm ? sm : 128;
      knobs.tile_n = sn ? sn : 128;
    }
  }
  int64_t MD = 1, ND = 1, KD = 1, BD = 1;
  for (auto &v : I) MD *= v.size;
  for (auto &v : J) ND *= v.size;
  for (auto &v : K) KD *= v.size;
  for (auto &v : B) BD *= v.size;

  // a plain matmul: one var each for rows, columns and K, A[M][K] and
  // B[K][N] read at exactly (i, k) and (k, j), C[M][N] row-major
  {
    auto plain = [&](const Operand &op, int s0, int s1, int64_t n0, int64_t n1) {
      return op.sizes.size() == 2 && op.pred_dims.empty() && op.sizes[0] == n0 &&
             op.sizes[1] == n1 && op.cst[0] == 0 && op.cst[1] == 0 &&
             op.coef[0].size() == 1 && op.coef[0].count(s0) && op.coef[0].at(s0) == 1 &&
             op.coef[1].size() == 1 && op.coef[1].count(s1) && op.coef[1].at(s1) == 1;
    };
    if (B.empty() && I.size() == 1 && J.size() == 1 && K.size() == 1 &&
        W.vars.size() == 2 && df.sym(W.vars[0]).id() == I[0].sym &&
        df.sym(W.vars[1]).id() == J[0].sym &&
        plain(ops[po], I[0].sym, K[0].sym, MD, KD) && plain(ops[qo], K[0].sym, J[0].sym, KD, ND)) {
      plan.matmul.valid = true;
      plan.matmul.a_input = ops[po].input_index;
      plan.matmul.b_input = ops[qo].input_index;
      plan.matmul.M = MD;
      plan.matmul.N = ND;
      plan.matmul.K = KD;
      plan.config["plain_matmul"] = 1;
    }
  }

  // staged operands and the implicit-GEMM tile (emit_igemm.cpp): 16-bit
  // operands, and float operands where the pipelined float tile applies
  if (!knobs.gather) {
    Contraction c;
    c.ops[0] = ops[0];
    c.ops[1] = ops[1];
    c.po = po;
    c.qo = qo;
    c.I = I;
    c.J = J;
    c.B = B;
    c.K = K;
    c.MD = MD;
    c.ND = ND;
    c.KD = KD;
    c.BD = BD;
    c.red_node = R.nr;
    int64_t st = 1;
    for (int d = (int)W.vars.size() - 1; d >= 0; --d) {
      c.ostride[df.sym(W.vars[d]).id()] = st;
      st *= df.size(W.vars[d]);
    }
    if (try_contract_staged(df, c, knobs, plan)) return true;
  }

  // ---- tile configuration ----
  const bool bf16 = knobs.bf16, f16 = knobs.f16 && !knobs.bf16;
  const bool lowp = bf16 || f16;
  int BM = 128, BN = 128;
  if (knobs.tile_m > 0) BM = knobs.tile_m;
  if (knobs.tile_n > 0) BN = knobs.tile_n;
  if (knobs.tile_m == 0 && MD * ND <= 256LL * 128 * 128 && MD <= 512 && ND <= 512) {
    BM = 64;
    BN = 64;
  }
  if (BM % 64 || BN % 64 || BM > 256 || BN > 256) return false;
  const int KI = lowp ? 16 : 2;  // k per MFMA
  // Nests too small to fill the GPU with 64 x 64 tiles (fewer than 64
  // workgroups): 32 x 32 tiles whose K tiles are split over the
  // workgroup's four waves -- each wave multiplies its quarter of every K
  // tile, and the four partial sums are added through LDS in wave order --
  // for four times the workgroups and sixteen times the waves, every wave
  // with a quarter of the MFMAs.  "ksplit=0|1" forces the choice.
  const int BKW = lowp ? 32 : 16;  // K per wave and K tile
  const int64_t wg64 = ((MD + 63) / 64) * ((ND + 63) / 64) * BD;
  const bool ksplit = knobs.ksplit == 1 ||
                      (knobs.ksplit != 0 && knobs.tile_m == 0 && BM == 64 && BN == 64 &&
                       wg64 < 64 && KD >= 4 * BKW);
  const int KS = ksplit ? 4 : 1;
  if (ksplit) BM = BN = 32;
  const int BK = KS * BKW;  // K extent of a K tile (the LDS tile)
  const int WAVES_M = ksplit ? 1 : 2, WAVES_N = ksplit ? 1 : 2, NT = 256;
  const int WM = BM / WAVES_M, WN = BN / WAVES_N;  // per-wave tile
  const int TM = WM / 32, TN = WN / 32;            // 32x32 MFMA tiles per wave
  const int PAD = lowp ? 8 : 1;
  const char *elt = bf16 ? "__bf16" : f16 ? "_Float16" : "float";

  // ---- operand loaders ----
  struct LoadPlan {
    bool kfast;
    int per;   // elements per thread per tile
    int rs;    // stride between a thread's elements (rows or ks)
  };
  auto plan_load = [&](int o, const std::vector<VarInfo> &U, int BU) {
    const auto &op = ops[o];
    // contiguous along k or along u?
    auto dim_coef = [&](int sym) {
      int64_t s = 0;
      for (size_t d = 0; d < op.coef.size(); ++d) {
        auto it = op.coef[d].find(sym);
        if (it != op.coef[d].end()) s += it->second * op.strides[d];
      }
      return s;
    };
    bool kfast = dim_coef(K.back().sym) == 1;
    bool ufast = dim_coef(U.back().sym) == 1;
    LoadPlan lp;
    lp.kfast = !ufast;  // prefer coalescing along the output-side var
    lp.per = BU * BK / NT;
    lp.rs = lp.kfast ? NT / BK : NT / BU;
    return lp;
  };
  LoadPlan lpP = plan_load(po, I, BM), lpQ = plan_load(qo, J, BN);
  if (lpP.per < 1 || lpQ.per < 1) return false;

  std::stringstream k;
  k << kernel_prelude();
  k << "typedef __bf16 lt_bf16x8 __attribute__((ext_vector_type(8)));\n"
       "typedef _Float16 lt_f16x8 __attribute__((ext_vector_type(8)));\n"
       "typedef float lt_f32x16 __attribute__((ext_vector_type(16)));\n";
  k << "extern \"C\" __global__ void __launch_bounds__(" << NT << ") lt_gemm0(";
  for (int i = 0; i < df.num_inputs; ++i) k << "const float* __restrict__ in" << i << ", ";
  k << "float* __restrict__ out) {\n";
  const int64_t NKT = (KD + BK - 1) / BK;
  const int NBUF = NKT > 1 ? 2 : 1;  // LDS buffers: one when a single K tile
  k << "  __shared__ " << elt << " sP[" << NBUF << "][" << BM << "][" << BK + PAD << "];\n";
  k << "  __shared__ " << elt << " sQ[" << NBUF << "][" << BN << "][" << BK + PAD << "];\n";
  k << "  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;\n";
  if (ksplit)  // every wave covers the whole tile, over its quarter of K
    k << "  const int wm = 0, wn = 0, kw = wave * " << BKW << ";\n"
      << "  __shared__ float lt_red[" << KS - 1 << "][16][64];\n";
  else
    k << "  const int wm = wave / " << WAVES_N << ", wn = wave % " << WAVES_N
      << ", kw = 0;\n";
  k << "  const long i0 = (long)blockIdx.y * " << BM << ", j0 = (long)blockIdx.x * " << BN
    << ";\n";
  // batch decomposition
  for (auto &v : B) k << "  long " << v.name << ";\n";
  k << decompose("(long)blockIdx.z", B, "  ");

  // per-operand: hoisted parts and per-tile loader code
  auto emit_loader = [&](int o, const std::vector<VarInfo> &U, const LoadPlan &lp,
                         const std::string &tag, int BU, int64_t UD) {
    const auto &op = ops[o];
    // linear index coefficient of each var
    auto idx_coef = [&](const std::vector<VarInfo> &vs) {
      std::map<int, int64_t> c;
      for (size_t d = 0; d < op.coef.size(); ++d)
        for (const auto &p : op.coef[d]) c[p.first] += p.second * op.strides[d];
      std::map<int, int64_t> out;
      for (auto &v : vs)
        if (c.count(v.sym)) out[v.sym] = c[v.sym];
      return out;
    };
    int64_t cst = 0;
    for (size_t d = 0; d < op.cst.size(); ++d) cst += op.cst[d] * op.strides[d];
    std::stringstream s;
    const int ND_ = (int)op.pred_dims.size();
    // u-side (incl. batch) parts
    auto upart = [&](const std::string &base, const std::string &pu, const std::string &ulin,
                     const std::string &ind) {
      std::stringstream t;
      t << ind << "{\n";
      for (auto &v : U) t << ind << "  long " << v.name << ";\n";
      t << decompose(ulin, U, ind + "  ");
      auto cu = idx_coef(U);
      auto cb = idx_coef(B);
      t << ind << "  " << base << " = " << lin_expr(cu, U) << " + " << lin_expr(cb, B)
        << " + " << cst << "L;\n";
      for (int q = 0; q < ND_; ++q) {
        int d = op.pred_dims[q];
        std::map<int, int64_t> cu_d, cb_d;
        for (auto &v : U)
          if (op.coef[d].count(v.sym)) cu_d[v.sym] = op.coef[d].at(v.sym);
        for (auto &v : B)
          if (op.coef[d].count(v.sym)) cb_d[v.sym] = op.coef[d].at(v.sym);
        t << ind << "  " << pu << "[" << q << "] = (int)(" << lin_expr(cu_d, U) << " + "
          << lin_expr(cb_d, B) << " + " << op.cst[d] << "L);\n";
      }
      t << ind << "}\n";
      return t.str();
    };
    auto kpart = [&](const std::string &off, const std::string &pk, const std::string &klin,
                     const std::string &ind) {
      std::stringstream t;
      t << ind << "{\n";
      for (auto &v : K) t << ind << "  long " << v.name << ";\n";
      t << decompose(klin, K, ind + "  ");
      t << ind << "  " << off << " = " << lin_expr(idx_coef(K), K) << ";\n";
      for (int q = 0; q < ND_; ++q) {
        int d = op.pred_dims[q];
        std::map<int, int64_t> ck;
        for (auto &v : K)
          if (op.coef[d].count(v.sym)) ck[v.sym] = op.coef[d].at(v.sym);
        t << ind << "  " << pk << "[" << q << "] = (int)(" << lin_expr(ck, K) << ");\n";
      }
      t << ind << "}\n";
      return t.str();
    };
    auto valid = [&](const std::string &pu, const std::string &pk) {
      std::string v;
      for (int q = 0; q < ND_; ++q) {
        int d = op.pred_dims[q];
        auto e = "(" + pu + "[" + std::to_string(q) + "] + " + pk + "[" + std::to_string(q) + "])";
        v += " && " + e + " >= 0 && " + e + " < " + std::to_string(op.sizes[d]);
      }
      return v;
    };
    const std::string arr = "in" + std::to_string(op.input_index);
    const int P = lp.per;
    // hoisted u-side state
    std::stringstream h;
    if (lp.kfast) {
      h << "  const int " << tag << "_tk = tid % " << BK << ", " << tag << "_tu = tid / " << BK
        << ";\n";
      h << "  long " << tag << "_base[" << P << "]; int " << tag << "_pu[" << P << "]["
        << std::max(ND_, 1) << "]; bool " << tag << "_uok[" << P << "];\n";
      h << "  #pragma unroll\n  for (int r = 0; r < " << P << "; ++r) {\n"
        << "    const long u = (" << (tag == "p" ? "i0" : "j0") << ") + " << tag << "_tu + r * "
        << lp.rs << ";\n"
        << "    " << tag << "_uok[r] = u < " << UD << "L;\n"
        << "    long b_ = 0; int pu_[" << std::max(ND_, 1) << "];\n"
        << upart("b_", "pu_", "u", "    ")
        << "    " << tag << "_base[r] = b_;\n";
      for (int q = 0; q < ND_; ++q) h << "    " << tag << "_pu[r][" << q << "] = pu_[" << q << "];\n";
      h << "  }\n";
    } else {
      h << "  const int " << tag << "_tu = tid % " << BU << ", " << tag << "_tk = tid / " << BU
        << ";\n";
      h << "  long " << tag << "_base; int " << tag << "_pu[" << std::max(ND_, 1) << "]; bool "
        << tag << "_uok;\n";
      h << "  {\n    const long u = (" << (tag == "p" ? "i0" : "j0") << ") + " << tag
        << "_tu;\n    " << tag << "_uok = u < " << UD << "L;\n"
        << upart(tag + "_base", tag + "_pu", "u", "    ") << "  }\n";
    }
    // per-tile load into registers
    std::stringstream ld;
    ld << "  __device__ __forceinline__ void lt_load_" << tag << "(";
    (void)ld;
    std::stringstream body;
    body << "    {\n";
    if (lp.kfast) {
      body << "      const long k = k0 + " << tag << "_tk;\n"
           << "      long off_ = 0; int pk_[" << std::max(ND_, 1) << "];\n"
           << kpart("off_", "pk_", "k", "      ")
           << "      const bool kok = k < " << KD << "L;\n"
           << "      #pragma unroll\n      for (int r = 0; r < " << P << "; ++r) {\n"
           << "        const bool ok = kok && " << tag << "_uok[r]"
           << valid(tag + "_pu[r]", "pk_") << ";\n"
           << "        " << tag << "_reg[r] = ok ? " << arr << "[" << tag << "_base[r] + off_] : 0.0f;\n"
           << "      }\n";
    } else {
      body << "      #pragma unroll\n      for (int r = 0; r < " << P << "; ++r) {\n"
           << "        const long k = k0 + " << tag << "_tk + r * " << lp.rs << ";\n"
           << "        long off_ = 0; int pk_[" << std::max(ND_, 1) << "];\n"
           << kpart("off_", "pk_", "k", "        ")
           << "        const bool ok = k < " << KD << "L && " << tag << "_uok"
           << valid(tag + "_pu", "pk_") << ";\n"
           << "        " << tag << "_reg[r] = ok ? " << arr << "[" << tag << "_base + off_] : 0.0f;\n"
           << "      }\n";
    }
    body << "    }\n";
    // store registers to LDS buffer `buf`
    std::stringstream st;
    st << "    #pragma unroll\n    for (int r = 0; r < " << P << "; ++r) {\n";
    if (lp.kfast)
      st << "      s" << (tag == "p" ? "P" : "Q") << "[buf][" << tag << "_tu + r * " << lp.rs
         << "][" << tag << "_tk] = (" << elt << ")" << tag << "_reg[r];\n";
    else
      st << "      s" << (tag == "p" ? "P" : "Q") << "[buf][" << tag << "_tu][" << tag
         << "_tk + r * " << lp.rs << "] = (" << elt << ")" << tag << "_reg[r];\n";
    st << "    }\n";
    h << "  float " << tag << "_reg[" << P << "];\n";
    return std::make_tuple(h.str(), body.str(), st.str());
  };
  auto [hP, loadP, storeP] = emit_loader(po, I, lpP, "p", BM, MD);
  auto [hQ, loadQ, storeQ] = emit_loader(qo, J, lpQ, "q", BN, ND);
  k << hP << hQ;
  k << "  lt_f32x16 acc[" << TM << "][" << TN << "];\n";
  k << "  #pragma unroll\n  for (int a = 0; a < " << TM << "; ++a)\n  #pragma unroll\n"
       "  for (int b = 0; b < " << TN << "; ++b) acc[a][b] = (lt_f32x16){};\n";
  k << "  long k0 = 0;\n  int buf = 0;\n";
  k << loadP << loadQ << storeP << storeQ << "  __syncthreads();\n";
  k << "  for (long kt = 0; kt < " << NKT << "; ++kt) {\n";
  k << "    const bool more = kt + 1 < " << NKT << ";\n";
  k << "    if (more) {\n      k0 = (kt + 1) * " << BK << ";\n" << loadP << loadQ << "    }\n";
  // MFMA over the current buffer
  k << "    #pragma unroll\n    for (int kk = 0; kk < " << BKW / KI << "; ++kk) {\n";
  if (lowp) {
    const char *vt = bf16 ? "lt_bf16x8" : "lt_f16x8";
    const char *fn = bf16 ? "__builtin_amdgcn_mfma_f32_32x32x16_bf16"
                          : "__builtin_amdgcn_mfma_f32_32x32x16_f16";
    k << "      " << vt << " av[" << TM << "], bv[" << TN << "];\n";
    k << "      #pragma unroll\n      for (int a = 0; a < " << TM << "; ++a)\n"
      << "        av[a] = *(const " << vt << "*)&sP[buf][wm * " << WM
      << " + a * 32 + (lane & 31)][kw + kk * 16 + (lane >> 5) * 8];\n";
    k << "      #pragma unroll\n      for (int b = 0; b < " << TN << "; ++b)\n"
      << "        bv[b] = *(const " << vt << "*)&sQ[buf][wn * " << WN
      << " + b * 32 + (lane & 31)][kw + kk * 16 + (lane >> 5) * 8];\n";
    k << "      #pragma unroll\n      for (int a = 0; a < " << TM << "; ++a)\n"
      << "      #pragma unroll\n      for (int b = 0; b < " << TN << "; ++b)\n"
      << "        acc[a][b] = " << fn << "(av[a], bv[b], acc[a][b], 0, 0, 0);\n";
  } else {
    k << "      float av[" << TM << "], bv[" << TN << "];\n";
    k << "      #pragma unroll\n      for (int a = 0; a < " << TM << "; ++a)\n"
      << "        av[a] = sP[buf][wm * " << WM << " + a * 32 + (lane & 31)][kw + kk * 2 + (lane >> 5)];\n";
    k << "      #pragma unroll\n      for (int b = 0; b < " << TN << "; ++b)\n"
      << "        bv[b] = sQ[buf][wn * " << WN << " + b * 32 + (lane & 31)][kw + kk * 2 + (lane >> 5)];\n";
    k << "      #pragma unroll\n      for (int a = 0; a < " << TM << "; ++a)\n"
      << "      #pragma unroll\n      for (int b = 0; b < " << TN << "; ++b)\n"
      << "        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[a], bv[b], acc[a][b], 0, 0, 0);\n";
  }
  k << "    }\n";
  k << "    if (more) {\n      buf ^= 1;\n" << storeP << storeQ << "    }\n";
  k << "    __syncthreads();\n  }\n";
  if (ksplit) {
    // the partial sums of waves 1..3 added to wave 0's in wave order
    k << "  if (wave > 0) {\n"
      << "    #pragma unroll\n    for (int r = 0; r < 16; ++r) lt_red[wave - 1][r][lane] = "
         "acc[0][0][r];\n  }\n"
      << "  __syncthreads();\n  if (wave > 0) return;\n"
      << "  #pragma unroll\n  for (int w = 0; w < " << KS - 1 << "; ++w)\n"
      << "  #pragma unroll\n    for (int r = 0; r < 16; ++r) acc[0][0][r] += lt_red[w][r][lane];\n";
  }

  // ---- epilogue: D layout of 32x32 MFMA: col = lane%32,
  // row = (r/4)*8 + (lane/32)*4 + r%4 for accumulator register r
  std::map<int, int64_t> ostride;
  {
    int64_t st = 1;
    for (int d = (int)W.vars.size() - 1; d >= 0; --d) {
      ostride[df.sym(W.vars[d]).id()] = st;
      st *= df.size(W.vars[d]);
    }
  }
  auto part = [&](const std::vector<VarInfo> &vs) {
    std::map<int, int64_t> c;
    for (auto &v : vs) c[v.sym] = ostride.at(v.sym);
    return lin_expr(c, vs);
  };
  k << "  long ob = " << part(B) << ";\n";
  for (auto &v : I) k << "  long " << v.name << ";\n";
  for (auto &v : J) k << "  long " << v.name << ";\n";
  k << "  #pragma unroll\n  for (int b = 0; b < " << TN << "; ++b) {\n"
    << "    const long j = j0 + wn * " << WN << " + b * 32 + (lane & 31);\n"
    << "    if (j >= " << ND << "L) continue;\n"
    << decompose("j", J, "    ") << "    const long oj = " << part(J) << ";\n"
    << "    #pragma unroll\n    for (int a = 0; a < " << TM << "; ++a) {\n"
    << "      #pragma unroll\n      for (int r = 0; r < 16; ++r) {\n"
    << "        const long i = i0 + wm * " << WM
    << " + a * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);\n"
    << "        if (i >= " << MD << "L) continue;\n"
    << decompose("i", I, "        ") << "        out[ob + oj + " << part(I)
    << "] = acc[a][b][r];\n      }\n    }\n  }\n}\n";

  Stage st;
  st.kernel = "lt_gemm0";
  st.kind = lowp ? (bf16 ? "contract_bf16" : "contract_f16") : "contract_f32";
  for (int i = 0; i < df.num_inputs; ++i) st.buffers.push_back(i);
  st.buffers.push_back(df.num_inputs);
  st.writes.push_back(df.num_inputs);
  auto gx = (ND + BN - 1) / BN, gy = (MD + BM - 1) / BM;
  ASSERT(gx < (1LL << 31) && gy < 65536 && BD < 65536) << "contraction grid too large";
  st.grid[0] = (unsigned)gx;
  st.grid[1] = (unsigned)gy;
  st.grid[2] = (unsigned)BD;
  st.block[0] = NT;
  st.flops = 2 * MD * ND * KD * BD;
  int64_t bytes = MD * ND * BD * 4;
  for (int i = 0; i < df.num_inputs; ++i)
    bytes += df.numel(df.nodes.at(df.ir().inputs()[i]).vars) * 4;
  st.bytes = bytes;
  plan.stages.push_back(st);
  plan.config["tile_m"] = BM;
  plan.config["tile_n"] = BN;
  plan.config["bk"] = BK;
  plan.config["ksplit"] = KS;
  plan.source = k.str();
  std::stringstream desc;
  desc << "[hip] contraction on MFMA (" << (lowp ? (bf16 ? "bf16" : "f16") : "f32")
       << " operands, f32 accumulate): D[" << BD << "][" << MD << "][" << ND
       << "] = P x Q over K=" << KD << ", tile " << BM << "x" << BN << "x" << BK
       << (ksplit ? ", K tiles split over 4 waves" : "") << "\n";
  plan.description = desc.str();
  return true;
}

}  // namespace hip
}  // namespace loop_tool
