// Staged-operand contraction lowering for 16-bit MFMA operands.
//
// For a contraction nest (see emit_contract.cpp for the analysis) whose two
// operands share a reduced var k_in that indexes one dim of each operand
// alone (coefficient 1, no offset), the plan is
//   stage 0, 1  one pass per operand: permute the operand so k_in's dim is
//               innermost, pad every dim that a view can index out of range
//               with a zero halo, zero-pad k_in to a multiple of the K
//               tile, round to bf16/fp16 (RNE);
//   stage 2     the implicit-GEMM tile of csrc/kernels/igemm_tile.h: every
//               operand access is now row_base(row) + outer_offset(k_out) +
//               k_in with no bounds checks, loaded as 16-byte vectors.
// For a 3x3 pad-1 convolution X[N][C][H][W] * W[Co][C][3][3] this stages
// X'[N][H+2][W+2][C] and W'[Co][3][3][C]: K = (kh, kw, ci), k_in = ci.  For
// a matmul A[M][K] B[K][N] it rounds A and transposes B to Bt[N][K].
//
// The staged buffers are backend scratch.  Out-of-range view reads still
// read 0 (they land in the halo), as `cond ? x : 0` in the reference C
// backend (src/backends/cpu/cpp.cpp:64-106).
#include <algorithm>
#include <sstream>

#include "contract.h"

namespace loop_tool {
namespace hip {

const char *igemm_tile_source() {
  return
#include "igemm_tile.inc"
      ;
}

namespace {

struct Staged {
  int input_index;
  int inner_dim;                     // source dim indexed by k_in alone
  std::vector<int> order;            // staged dims as source dims, outer to inner
  std::vector<int64_t> padlo, ssize; // per source dim
  std::vector<int64_t> sstride;      // per source dim (element stride in staging)
  std::map<int, int64_t> scoef;      // var sym -> staged element coefficient
  int64_t cst = 0;                   // staged offset of the all-zero index
  int64_t total = 1;                 // staged elements
};

int find_inner_dim(const Operand &op, int sym) {
  int found = -1;
  for (size_t d = 0; d < op.coef.size(); ++d) {
    if (!op.coef[d].count(sym)) continue;
    if (found >= 0) return -1;  // k_in in two dims
    found = (int)d;
  }
  if (found < 0) return -1;
  const auto &cf = op.coef[found];
  if (cf.size() != 1 || cf.at(sym) != 1 || op.cst[found] != 0) return -1;
  return found;
}

Staged stage_layout(const Operand &op, int inner_dim, int64_t kin_padded) {
  Staged s;
  s.input_index = op.input_index;
  s.inner_dim = inner_dim;
  const int nd = (int)op.sizes.size();
  for (int d = 0; d < nd; ++d)
    if (d != inner_dim) s.order.push_back(d);
  s.order.push_back(inner_dim);
  s.padlo.assign(nd, 0);
  s.ssize.assign(nd, 0);
  s.sstride.assign(nd, 0);
  for (int d = 0; d < nd; ++d) {
    int64_t lo = std::max<int64_t>(0, -op.lo[d]);
    int64_t hi = std::max<int64_t>(0, op.hi[d] - (op.sizes[d] - 1));
    s.padlo[d] = lo;
    s.ssize[d] = op.sizes[d] + lo + hi;
  }
  s.ssize[inner_dim] = kin_padded;  // k_in in [0, KI), zeros up to the padded extent
  int64_t st = 1;
  for (int q = nd - 1; q >= 0; --q) {
    int d = s.order[q];
    s.sstride[d] = st;
    st *= s.ssize[d];
  }
  s.total = st;
  for (int d = 0; d < nd; ++d) {
    for (const auto &p : op.coef[d]) s.scoef[p.first] += p.second * s.sstride[d];
    s.cst += (op.cst[d] + s.padlo[d]) * s.sstride[d];
  }
  return s;
}

// The staged dim that is contiguous in the source, when it is not the
// innermost staged dim (-1 otherwise).
int source_contiguous_dim(const Operand &op, const Staged &s) {
  const int nd = (int)op.sizes.size(), in = s.inner_dim;
  int ld = -1;
  int64_t best = op.strides[in];
  for (int d = 0; d < nd; ++d)
    if (d != in && s.ssize[d] > 1 && op.strides[d] < best) {
      best = op.strides[d];
      ld = d;
    }
  return ld;
}

// Staging as a transpose through LDS, for an operand whose source-contiguous
// dim `ld` is not the staged inner dim (a convolution input X[N][C][H][W]
// staged as X'[N][H+2][W+2][C]): a workgroup reads a 64 (ld) x TI (inner)
// tile with lanes along ld (coalesced reads), and writes it back with
// TI/8 lanes covering the 16-byte chunks of one ld position (coalesced
// writes).
std::string staging_transpose_kernel(const std::string &name, const Operand &op,
                                     const Staged &s, int ld, int64_t kin, const char *elt,
                                     unsigned grid[3]) {
  const int nd = (int)op.sizes.size(), in = s.inner_dim;
  const int TL = 64, TI = s.ssize[in] % 64 == 0 ? 64 : 32;
  std::vector<int> other;
  for (int q = 0; q < nd; ++q)
    if (s.order[q] != ld && s.order[q] != in) other.push_back(s.order[q]);
  int64_t nz = 1;
  for (int d : other) nz *= s.ssize[d];
  grid[0] = (unsigned)((s.ssize[ld] + TL - 1) / TL);
  grid[1] = (unsigned)(s.ssize[in] / TI);
  ASSERT(nz < 65536) << "staging grid too large";
  grid[2] = (unsigned)nz;
  auto idx = [&](int d, const std::string &sv) {
    return "(" + sv + " - " + std::to_string(s.padlo[d]) + "L)";
  };
  std::stringstream k;
  k << "extern \"C\" __global__ void __launch_bounds__(256) " << name
    << "(const float* __restrict__ in, " << elt << "* __restrict__ S) {\n"
    << "  typedef " << elt << " T8 __attribute__((ext_vector_type(8)));\n"
    << "  __shared__ float tile[" << TI << "][" << TL + 1 << "];\n"
    << "  const int tid = threadIdx.x;\n"
    << "  const long l0 = (long)blockIdx.x * " << TL << ", c0 = (long)blockIdx.y * " << TI
    << ";\n  long z_ = blockIdx.z;\n";
  for (int q = (int)other.size() - 1; q >= 0; --q) {
    int d = other[q];
    if (q > 0)
      k << "  const long s" << d << " = z_ % " << s.ssize[d] << "L; z_ /= " << s.ssize[d]
        << "L;\n";
    else
      k << "  const long s" << d << " = z_;\n";
  }
  // validity and offsets of the dims fixed for the workgroup
  std::string ok = "true", base = "0L", dbase = "0L";
  for (int d : other) {
    const std::string sv = "s" + std::to_string(d);
    if (s.ssize[d] != op.sizes[d])
      ok += " && " + idx(d, sv) + " >= 0 && " + idx(d, sv) + " < " +
            std::to_string(op.sizes[d]) + "L";
    base += " + " + idx(d, sv) + " * " + std::to_string(op.strides[d]) + "L";
    dbase += " + " + sv + " * " + std::to_string(s.sstride[d]) + "L";
  }
  std::string lok;
  if (s.ssize[ld] != op.sizes[ld])
    lok = " && " + idx(ld, "sl") + " >= 0 && " + idx(ld, "sl") + " < " +
          std::to_string(op.sizes[ld]) + "L";
  k << "  const bool fixed_ok = " << ok << ";\n"
    << "  const long src0 = " << base << ", dst0 = " << dbase << ";\n"
    << "  #pragma unroll\n  for (int q = 0; q < " << TI / 4 << "; ++q) {\n"
    << "    const int r = (tid >> 6) + q * 4, l = tid & 63;\n"
    << "    const long sl = l0 + l, si = c0 + r;\n"
    << "    const bool ok = fixed_ok && sl < " << s.ssize[ld] << "L && si < " << kin << "L"
    << lok << ";\n"
    << "    tile[r][l] = ok ? in[src0 + " << idx(ld, "sl") << " * " << op.strides[ld]
    << "L + si * " << op.strides[in] << "L] : 0.0f;\n  }\n"
    << "  __syncthreads();\n"
    << "  #pragma unroll\n  for (int q = 0; q < " << TL * TI / 8 / 256 << "; ++q) {\n"
    << "    const int e = tid + q * 256, g = e % " << TI / 8 << ", l = e / " << TI / 8
    << ";\n"
    << "    const long sl = l0 + l;\n"
    << "    if (sl >= " << s.ssize[ld] << "L) continue;\n"
    << "    T8 v;\n    #pragma unroll\n    for (int j = 0; j < 8; ++j) v[j] = (" << elt
    << ")tile[g * 8 + j][l];\n"
    << "    *(T8*)(S + dst0 + sl * " << s.sstride[ld] << "L + c0 + g * 8) = v;\n  }\n}\n";
  return k.str();
}

// One thread per 8 staged elements along the innermost staged dim.  Lanes
// walk the dim that is contiguous in the source (when it is not the innermost
// staged dim) so the source reads coalesce.
std::string staging_kernel(const std::string &name, const Operand &op, const Staged &s,
                           int64_t kin, const char *elt, int64_t *threads) {
  const int nd = (int)op.sizes.size();
  const int in = s.inner_dim;
  const int ld = source_contiguous_dim(op, s);
  // thread decomposition, fastest first: (dim, extent)
  std::vector<std::pair<int, int64_t>> tord;
  if (ld >= 0) tord.push_back({ld, s.ssize[ld]});
  tord.push_back({in, s.ssize[in] / 8});
  for (int q = nd - 2; q >= 0; --q)
    if (s.order[q] != ld) tord.push_back({s.order[q], s.ssize[s.order[q]]});
  const int64_t nvec = s.total / 8;
  *threads = nvec;

  std::stringstream k;
  k << "extern \"C\" __global__ void __launch_bounds__(256) " << name
    << "(const float* __restrict__ in, " << elt << "* __restrict__ S) {\n"
    << "  typedef " << elt << " T8 __attribute__((ext_vector_type(8)));\n"
    << "  long t_ = (long)blockIdx.x * 256 + threadIdx.x;\n"
    << "  if (t_ >= " << nvec << "L) return;\n";
  for (int d = 0; d < nd; ++d) k << "  long s" << d << ";\n";
  for (size_t q = 0; q < tord.size(); ++q) {
    int d = tord[q].first;
    if (q + 1 < tord.size())
      k << "  s" << d << " = t_ % " << tord[q].second << "L; t_ /= " << tord[q].second
        << "L;\n";
    else
      k << "  s" << d << " = t_;\n";
  }
  k << "  s" << in << " *= 8;\n";
  std::string ok = "true", src = "0L", dst = "0L";
  for (int d = 0; d < nd; ++d) {
    std::string idx = "(s" + std::to_string(d) + " - " + std::to_string(s.padlo[d]) + "L)";
    if (s.ssize[d] != op.sizes[d])
      ok += " && " + idx + " >= 0 && " + idx + " < " + std::to_string(op.sizes[d]) + "L";
    src += " + " + idx + " * " + std::to_string(op.strides[d]) + "L";
    dst += " + s" + std::to_string(d) + " * " + std::to_string(s.sstride[d]) + "L";
  }
  k << "  const bool ok = " << ok << ";\n"
    << "  T8 v;\n"
    << "  #pragma unroll\n  for (int e = 0; e < 8; ++e)\n"
    << "    v[e] = ok && s" << in << " + e < " << kin << "L ? (" << elt << ")in[" << src
    << " + e * " << op.strides[in] << "L] : (" << elt << ")0.0f;\n"
    << "  *(T8*)(S + " << dst << ") = v;\n}\n";
  return k.str();
}

std::string lin_over(const std::map<int, int64_t> &coef, const std::vector<VarInfo> &vs) {
  std::map<int, int64_t> c;
  for (const auto &v : vs)
    if (coef.count(v.sym)) c[v.sym] = coef.at(v.sym);
  return lin_expr(c, vs);
}

std::string decls(const std::vector<VarInfo> &vs, const std::string &ind) {
  std::string out;
  for (const auto &v : vs) out += ind + "long " + v.name + ";\n";
  return out;
}

}  // namespace

bool try_contract_staged(const Dataflow &df, const Contraction &c, const Knobs &knobs,
                         Plan &plan) {
  const bool bf16 = knobs.bf16;
  const char *elt = bf16 ? "__bf16" : "_Float16";
  // k_in: the largest reduced var that indexes one dim of each operand alone
  int kin = -1, dims[2] = {-1, -1};
  for (size_t q = 0; q < c.K.size(); ++q) {
    int sym = c.K[q].sym;
    int d0 = find_inner_dim(c.ops[0], sym), d1 = find_inner_dim(c.ops[1], sym);
    if (d0 < 0 || d1 < 0) continue;
    if (kin < 0 || c.K[q].size > c.K[kin].size) {
      kin = (int)q;
      dims[0] = d0;
      dims[1] = d1;
    }
  }
  if (kin < 0) return false;
  std::vector<VarInfo> KO;
  int64_t ko_n = 1;
  for (size_t q = 0; q < c.K.size(); ++q)
    if ((int)q != kin) {
      KO.push_back(c.K[q]);
      ko_n *= c.K[q].size;
    }
  const int64_t KI = c.K[kin].size;
  ASSERT(ko_n < (1LL << 31) && KI < (1LL << 31)) << "reduction too large";

  // ---- tile: the largest tile (fewest LDS bytes per flop) that still gives
  // every CU a workgroup; K tile 64 where the LDS of two stages allows ----
  auto tiles = [&](int bm, int bn) {
    return ((c.MD + bm - 1) / bm) * ((c.ND + bn - 1) / bn) * c.BD;
  };
  int BM = 64, BN = 64, BK = 32;
  if (knobs.tile_m > 0) {
    BM = knobs.tile_m;
    BN = knobs.tile_n;
    BK = knobs.bk > 0 ? knobs.bk : 32;
  } else {
    const int cand[][3] = {{256, 256, 64}, {128, 256, 32}, {256, 128, 32}, {128, 128, 64},
                           {128, 64, 32},  {64, 128, 32},  {64, 64, 32}};
    for (auto &t : cand) {
      if (t[0] > ((c.MD + 63) / 64) * 64 || t[1] > ((c.ND + 63) / 64) * 64) continue;
      if (tiles(t[0], t[1]) >= 256) {
        BM = t[0];
        BN = t[1];
        BK = t[2];
        break;
      }
    }
    if (knobs.bk > 0) BK = knobs.bk;
  }
  if (BK != 32 && BK != 64) return false;
  if (BM % 64 || BN % 64 || BM > 256 || BN > 256 || BM * BN > 256 * 256) return false;
  if (2 * (BM + BN) * (BK + 8) * 2 > 160 * 1024) return false;  // LDS of a CU
  const int64_t KIP = (KI + BK - 1) / BK * BK;
  Staged st[2] = {stage_layout(c.ops[0], dims[0], KIP), stage_layout(c.ops[1], dims[1], KIP)};
  const Staged &SP = st[c.po], &SQ = st[c.qo];
  const int lds = 2 * (BM + BN) * (BK + 8) * 2;
  const bool full = c.MD % BM == 0 && c.ND % BN == 0;
  const bool int32 = SP.total < (1LL << 31) && SQ.total < (1LL << 31);

  // ---- buffers: one staged (16-bit) scratch buffer per operand ----
  int sbuf[2];
  int scratch_no = 0;
  for (const auto &b : plan.buffers) scratch_no += b.kind == BufferSpec::SCRATCH;
  for (int o = 0; o < 2; ++o) {
    sbuf[o] = (int)plan.buffers.size();
    int64_t words = ((st[o].total + 1) / 2 + 7) / 8 * 8;
    auto nr = df.ir().inputs().at(c.ops[o].input_index);
    plan.buffers.push_back({BufferSpec::SCRATCH, scratch_no++, words, nr});
  }
  const int out_buf = df.num_inputs;  // the single output

  std::stringstream k;
  k << kernel_prelude() << igemm_tile_source() << "\n";
  for (int o = 0; o < 2; ++o) {
    std::string name = "lt_stage" + std::to_string(o);
    Stage s;
    const int ld = source_contiguous_dim(c.ops[o], st[o]);
    if (ld >= 0 && c.ops[o].strides[ld] == 1 && st[o].ssize[ld] >= 32 &&
        st[o].ssize[st[o].inner_dim] % 32 == 0 &&
        st[o].total / (st[o].ssize[ld] * st[o].ssize[st[o].inner_dim]) < 65536) {
      k << staging_transpose_kernel(name, c.ops[o], st[o], ld, KI, elt, s.grid);
    } else {
      int64_t threads = 0;
      k << staging_kernel(name, c.ops[o], st[o], KI, elt, &threads);
      s.grid[0] = (unsigned)((threads + 255) / 256);
    }
    s.kernel = name;
    s.kind = "stage";
    s.buffers = {c.ops[o].input_index, sbuf[o]};
    s.writes = {sbuf[o]};
    s.block[0] = 256;
    s.bytes = df.numel(df.nodes.at(df.ir().inputs()[c.ops[o].input_index]).vars) * 4 +
              st[o].total * 2;
    plan.stages.push_back(s);
  }
  // geometry of the tile
  k << "struct LtGeo {\n"
    << "  static constexpr int KO = " << ko_n << ", KI = " << KIP << ";\n"
    << "  static constexpr long MD = " << c.MD << "L, ND = " << c.ND << "L;\n"
    << "  long zp, zq, zo;\n"
    << "  __device__ long p_row(long i, bool& ok) const {\n    ok = i < MD;\n"
    << decls(c.I, "    ") << decompose("i", c.I, "    ") << "    return zp + "
    << lin_over(SP.scoef, c.I) << " + " << SP.cst << "L;\n  }\n"
    << "  __device__ long q_row(long j, bool& ok) const {\n    ok = j < ND;\n"
    << decls(c.J, "    ") << decompose("j", c.J, "    ") << "    return zq + "
    << lin_over(SQ.scoef, c.J) << " + " << SQ.cst << "L;\n  }\n"
    << "  __device__ long p_koff(int ko) const {\n"
    << decls(KO, "    ") << decompose("ko", KO, "    ") << "    return "
    << lin_over(SP.scoef, KO) << ";\n  }\n"
    << "  __device__ long q_koff(int ko) const {\n"
    << decls(KO, "    ") << decompose("ko", KO, "    ") << "    return "
    << lin_over(SQ.scoef, KO) << ";\n  }\n"
    << "  __device__ long out_row(long i) const {\n"
    << decls(c.I, "    ") << decompose("i", c.I, "    ") << "    return zo + "
    << lin_over(c.ostride, c.I) << ";\n  }\n"
    << "  __device__ long out_col(long j) const {\n"
    << decls(c.J, "    ") << decompose("j", c.J, "    ") << "    return "
    << lin_over(c.ostride, c.J) << ";\n  }\n"
    << "};\n";
  k << "extern \"C\" __global__ void __launch_bounds__(256) lt_igemm0(const " << elt
    << "* __restrict__ P, const " << elt << "* __restrict__ Q, float* __restrict__ out) {\n"
    << "  __shared__ __attribute__((aligned(16))) unsigned char smem[" << lds << "];\n"
    << "  LtGeo g;\n"
    << decls(c.B, "  ") << decompose("(long)blockIdx.z", c.B, "  ")
    << "  g.zp = " << lin_over(SP.scoef, c.B) << ";\n"
    << "  g.zq = " << lin_over(SQ.scoef, c.B) << ";\n"
    << "  g.zo = " << lin_over(c.ostride, c.B) << ";\n"
    << "  lt_igemm_tile<" << elt << ", " << BM << ", " << BN << ", " << BK << ", "
    << (full ? "true" : "false") << ", " << (int32 ? "int" : "long")
    << ">(g, P, Q, out, (long)blockIdx.y * " << BM
    << ", (long)lt_igemm_swizzle(blockIdx.x, gridDim.x) * " << BN << ", smem);\n}\n";

  Stage g;
  g.kernel = "lt_igemm0";
  g.kind = bf16 ? "contract_bf16" : "contract_f16";
  g.buffers = {sbuf[c.po], sbuf[c.qo], out_buf};
  g.writes = {out_buf};
  auto gx = (c.ND + BN - 1) / BN, gy = (c.MD + BM - 1) / BM;
  ASSERT(gx < (1LL << 31) && gy < 65536 && c.BD < 65536) << "contraction grid too large";
  g.grid[0] = (unsigned)gx;
  g.grid[1] = (unsigned)gy;
  g.grid[2] = (unsigned)c.BD;
  g.block[0] = 256;
  g.flops = 2 * c.MD * c.ND * c.KD * c.BD;
  g.bytes = SP.total * 2 + SQ.total * 2 + c.MD * c.ND * c.BD * 4;
  plan.stages.push_back(g);
  plan.source = k.str();
  std::stringstream desc;
  desc << "[hip] contraction on MFMA (" << (bf16 ? "bf16" : "f16")
       << " staged operands, f32 accumulate): D[" << c.BD << "][" << c.MD << "][" << c.ND
       << "] = P x Q over K=" << ko_n << "x" << KI << (KIP != KI ? " (padded to " + std::to_string(KIP) + ")" : std::string()) << ", tile " << BM << "x" << BN << "x"
       << BK << "; staged P " << SP.total << " and Q " << SQ.total << " elements\n";
  plan.description = desc.str();
  return true;
}

}  // namespace hip
}  // namespace loop_tool
