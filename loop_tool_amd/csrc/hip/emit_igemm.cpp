// Staged-operand contraction lowering: 16-bit MFMA operands (bf16, fp16)
// with fp32 accumulation, and float operands for large nests.
//
// For a contraction nest (see emit_contract.cpp for the analysis) whose two
// operands share a reduced var k_in that indexes one dim of each operand
// alone (coefficient 1, no offset), the plan is
//   stage 0     one launch that stages both operands: zero-pad k_in to a
//               multiple of the K tile BK and cut it into blocks of BK (the
//               block index outermost, the BK elements innermost: k-blocked),
//               put the dims indexed only by reduced vars next, pad every dim
//               that a view can index out of range with a zero halo, and
//               round to bf16/fp16 (RNE) or keep float;
//   stage 1     an implicit-GEMM tile of csrc/kernels/igemm_tile.h: every
//               operand access is row_base(row) + outer_offset(k_out) + k_in
//               with no bounds checks, loaded as 16-byte buffer loads.
// For a 3x3 pad-1 convolution X[N][C][H][W] * W[Co][C][3][3] with BK 32 this
// stages X'[C/32][N][H+2][W+2][32] and W'[C/32][3][3][Co][32]: K = (block,
// kh, kw, ci).  Where the outer-K positions of a block only shift the input
// rows (the window taps), the halo form of the tile loads those rows once
// per block.  For a matmul A[M][K] B[K][N] it stages A'[K/BK][M][BK] and the
// transposed Bt'[K/BK][N][BK].
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
  // k-blocked layout (kblk > 0): k_in = kb * kblk + kr is staged as an
  // outermost dim kb (stride skb, KB blocks) and an innermost dim kr
  int64_t kblk = 0, skb = 0, KB = 1;
};

// staged element offset of inner-dim index `e` (a C expression; multiples of
// 8 never straddle a block)
std::string kin_dst(const Staged &s, const std::string &e) {
  if (!s.kblk) return "(" + e + ")";
  return "((" + e + ") / " + std::to_string(s.kblk) + "L * " + std::to_string(s.skb) +
         "L + (" + e + ") % " + std::to_string(s.kblk) + "L)";
}

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

// kblk > 0: the k-blocked layout [kb][dims indexed only by reduced vars]
// [other dims][kr], so that the rows of one K block are kblk-element runs
// next to each other (a matmul operand A[M][K] becomes A'[K/kblk][M][kblk], a
// convolution input X[N][C][H][W] becomes X'[C/kblk][N][H+2][W+2][kblk]).
Staged stage_layout(const Operand &op, int inner_dim, int64_t kin_padded, int64_t kblk,
                    const std::vector<VarInfo> &K) {
  Staged s;
  s.input_index = op.input_index;
  s.inner_dim = inner_dim;
  const int nd = (int)op.sizes.size();
  auto k_only = [&](int d) {
    if (op.coef[d].empty()) return false;
    for (const auto &p : op.coef[d]) {
      bool in_k = false;
      for (const auto &v : K) in_k = in_k || v.sym == p.first;
      if (!in_k) return false;
    }
    return true;
  };
  const bool blocked = kblk > 0 && kin_padded > kblk;
  if (blocked)
    for (int d = 0; d < nd; ++d)
      if (d != inner_dim && k_only(d)) s.order.push_back(d);
  for (int d = 0; d < nd; ++d)
    if (d != inner_dim && !(blocked && k_only(d))) s.order.push_back(d);
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
    st *= d == inner_dim && blocked ? kblk : s.ssize[d];
  }
  s.total = st;
  if (blocked) {
    s.kblk = kblk;
    s.KB = kin_padded / kblk;
    s.skb = st;
    s.total = st * s.KB;
  }
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

// The fixed-dim prologue shared by the staging bodies: decomposes `z_` over
// the staged dims in `other` (outer to inner) into s<d>, and builds the
// validity test and source / staged offsets of those dims.
void fixed_dims(std::stringstream &k, const Operand &op, const Staged &s,
                const std::vector<int> &other, std::string &ok, std::string &src,
                std::string &dst) {
  for (int q = (int)other.size() - 1; q >= 0; --q) {
    int d = other[q];
    if (q > 0)
      k << "  const long s" << d << " = z_ % " << s.ssize[d] << "L; z_ /= " << s.ssize[d]
        << "L;\n";
    else
      k << "  const long s" << d << " = z_;\n";
  }
  ok = "true";
  src = "0L";
  dst = "0L";
  for (int d : other) {
    const std::string sv = "s" + std::to_string(d);
    const std::string idx = "(" + sv + " - " + std::to_string(s.padlo[d]) + "L)";
    if (s.ssize[d] != op.sizes[d])
      ok += " && " + idx + " >= 0 && " + idx + " < " + std::to_string(op.sizes[d]) + "L";
    src += " + " + idx + " * " + std::to_string(op.strides[d]) + "L";
    dst += " + " + sv + " * " + std::to_string(s.sstride[d]) + "L";
  }
}

// Staging of whole source planes through LDS, for an operand whose two
// fastest source dims are contiguous rows (ld: stride 1; a: stride =
// sizes[ld]) and whose staged inner dim is a third dim (a convolution input
// X[N][C][H][W] staged as X'[N][H+2][W+2][C]).  A workgroup owns RA staged
// rows of `a` and TI values of the inner dim: per inner value it reads the
// RA source rows as one contiguous run (16-byte loads, 8 lanes per 128-byte
// line, 8 inner values per wave instruction), rounds to 16 bits into an LDS
// tile [position][inner] (row pitch TI + 4 elements: the lanes of one store
// instruction hit distinct banks), then writes every staged position of the
// rows, halo included, as 16-byte chunks of TI/8 consecutive inner values.
int64_t staging_plane_body(std::stringstream &k, const std::string &name, const Operand &op,
                           const Staged &s, int ld, int a, int64_t kin, const char *elt,
                           int rows) {
  const int nd = (int)op.sizes.size(), in = s.inner_dim;
  // inner values per workgroup: 64 16-bit or 32 float values (the LDS tile
  // holds NP positions of TI values)
  const bool wide = std::string(elt) == "float";
  const int TI = s.ssize[in] % 64 == 0 && !wide ? 64 : 32, TP = TI + 4;
  const int64_t L = op.sizes[ld];
  // rows per workgroup: 256 positions' worth unless given, within 64 KB of LDS
  int64_t RA = rows > 0 ? rows : std::max<int64_t>(1, 256 / L);
  RA = std::max<int64_t>(1, std::min<int64_t>({RA, s.ssize[a], 65536 / (L * TP * (wide ? 4 : 2))}));
  const int64_t NP = RA * L, NQ = (NP + 3) / 4, NQP = (NQ + 7) / 8 * 8, CP = TI / 2;
  // row blocks start at source rows that are multiples of RA (the staged
  // rows shifted by the halo), so that a block's RA source rows start where
  // a block of rows starts in the source and their runs begin on the same
  // cache-line boundaries (a block may begin with halo rows below 0, which
  // are skipped)
  const int64_t ashift = (RA - s.padlo[a] % RA) % RA;
  const int64_t nA = (s.ssize[a] + ashift + RA - 1) / RA, nC = s.ssize[in] / TI;
  std::vector<int> other;
  for (int q = 0; q < nd; ++q)
    if (s.order[q] != ld && s.order[q] != in && s.order[q] != a) other.push_back(s.order[q]);
  int64_t nz = 1;
  for (int d : other) nz *= s.ssize[d];
  bool vec = L % 4 == 0 && op.strides[in] % 4 == 0;
  for (int d : other) vec = vec && op.strides[d] % 4 == 0;
  const int64_t rd_items = NQP * CP, wr_items = RA * s.ssize[ld] * (TI / 8);
  k << "__device__ __forceinline__ void " << name << "(const float* __restrict__ in, " << elt
    << "* __restrict__ S, unsigned bid) {\n"
    << "  typedef " << elt << " T2 __attribute__((ext_vector_type(2)));\n"
    << "  typedef " << elt << " T4 __attribute__((ext_vector_type(4)));\n"
    << "  typedef " << elt << " T8 __attribute__((ext_vector_type(8)));\n"
    << "  typedef float F4 __attribute__((ext_vector_type(4)));\n"
    << "  __shared__ __attribute__((aligned(16))) " << elt << " tile[" << NP * TP << "];\n"
    << "  const int tid = threadIdx.x;\n"
    << "  const long c0 = (long)(bid % " << nC << "u) * " << TI << ";\n"
    << "  const long a0 = (long)((bid / " << nC << "u) % " << nA << "u) * " << RA << " - "
    << ashift << "L;\n"
    << "  long z_ = bid / " << nC * nA << "u;\n";
  std::string ok, src, dst;
  fixed_dims(k, op, s, other, ok, src, dst);
  const std::string pa = std::to_string(s.padlo[a]);
  k << "  const bool fixed_ok = " << ok << ";\n"
    << "  const long src0 = " << src << " + (a0 - " << pa << "L) * " << L << "L;\n"
    << "  const long dst0 = " << dst << ";\n"
    // positions [pv0, pv1) of the tile hold source rows
    << "  const long pv0 = (a0 < " << pa << "L ? " << pa << "L - a0 : 0L) * " << L << "L;\n"
    << "  long pv1 = " << op.sizes[a] << "L + " << pa << "L - a0;\n"
    << "  pv1 = (pv1 < " << RA << "L ? pv1 : " << RA << "L) * " << L << "L;\n"
    << "  const bool vec = " << (vec ? "true" : "false")
    << " && (((unsigned long)in & 15ul) == 0ul);\n"
    << "  #pragma unroll\n  for (int it = 0; it < " << (rd_items + 255) / 256 << "; ++it) {\n"
    << "    const int e = tid + it * 256;\n";
  if (rd_items % 256) k << "    if (e >= " << rd_items << ") break;\n";
  k << "    const int rest = e >> 6;\n"
    << "    const int q = (rest % " << NQP / 8 << ") * 8 + (e & 7);\n"
    << "    const int cp = (rest / " << NQP / 8 << ") * 8 + ((e >> 3) & 7);\n"
    << "    if (q >= " << NQ << ") continue;\n"
    << "    const long p = (long)q * 4, ch = c0 + 2 * cp;\n"
    << "    const long o0 = src0 + ch * " << op.strides[in] << "L + p, o1 = o0 + "
    << op.strides[in] << "L;\n"
    << "    float x0[4], x1[4];\n"
    << "    if (vec && fixed_ok && p >= pv0 && p + 4 <= pv1 && ch + 1 < " << kin << "L) {\n"
    << "      const F4 u = *(const F4*)(in + o0), w = *(const F4*)(in + o1);\n"
    << "      #pragma unroll\n      for (int i = 0; i < 4; ++i) { x0[i] = u[i]; x1[i] = w[i]; }\n"
    << "    } else {\n"
    << "      #pragma unroll\n      for (int i = 0; i < 4; ++i) {\n"
    << "        const bool pk = fixed_ok && p + i >= pv0 && p + i < pv1;\n"
    << "        x0[i] = pk && ch < " << kin << "L ? in[o0 + i] : 0.0f;\n"
    << "        x1[i] = pk && ch + 1 < " << kin << "L ? in[o1 + i] : 0.0f;\n"
    << "      }\n    }\n"
    << "    #pragma unroll\n    for (int i = 0; i < 4; ++i) {\n"
    << "      if (p + i >= " << NP << "L) break;\n"
    << "      T2 t; t[0] = (" << elt << ")x0[i]; t[1] = (" << elt << ")x1[i];\n"
    << "      *(T2*)(tile + (p + i) * " << TP << " + 2 * cp) = t;\n"
    << "    }\n  }\n"
    << "  __syncthreads();\n"
    << "  #pragma unroll\n  for (int it = 0; it < " << (wr_items + 255) / 256 << "; ++it) {\n"
    << "    const int e = tid + it * 256;\n";
  if (wr_items % 256) k << "    if (e >= " << wr_items << ") break;\n";
  const std::string pl = std::to_string(s.padlo[ld]);
  k << "    const int g = e % " << TI / 8 << ";\n"
    << "    const long sl = (e / " << TI / 8 << ") % " << s.ssize[ld] << ", r = e / "
    << (TI / 8) * s.ssize[ld] << ";\n"
    << "    if (a0 + r >= " << s.ssize[a] << "L) break;\n"
    << "    if (a0 + r < 0L) continue;\n"
    << "    const long w = sl - " << pl << "L, ra = a0 + r - " << pa << "L;\n"
    << "    T8 v = {};\n"
    << "    if (fixed_ok && w >= 0 && w < " << L << "L && ra >= 0 && ra < " << op.sizes[a]
    << "L) {\n"
    << "      const T4* t4 = (const T4*)(tile + (r * " << L << "L + w) * " << TP << " + g * 8);\n"
    << "      const T4 lo = t4[0], hi = t4[1];\n"
    << "      #pragma unroll\n      for (int j = 0; j < 4; ++j) { v[j] = lo[j]; v[j + 4] = hi[j]; }\n"
    << "    }\n"
    << "    *(T8*)(S + dst0 + (a0 + r) * " << s.sstride[a] << "L + sl * " << s.sstride[ld]
    << "L + " << kin_dst(s, "c0 + g * 8") << ") = v;\n  }\n}\n";
  return nz * nA * nC;
}

// Staging as a transpose through LDS, for an operand whose source-contiguous
// dim `ld` is not the staged inner dim (B[K][N] of a matmul staged as
// Bt[N][K]): a workgroup reads a 64 (ld) x TI (inner) tile with lanes along
// ld (coalesced reads), and writes it back with TI/8 lanes covering the
// 16-byte chunks of one ld position (coalesced writes).
int64_t staging_transpose_body(std::stringstream &k, const std::string &name,
                               const Operand &op, const Staged &s, int ld, int64_t kin,
                               const char *elt) {
  const int nd = (int)op.sizes.size(), in = s.inner_dim;
  const int TL = 64, TI = s.ssize[in] % 64 == 0 ? 64 : 32;
  std::vector<int> other;
  for (int q = 0; q < nd; ++q)
    if (s.order[q] != ld && s.order[q] != in) other.push_back(s.order[q]);
  int64_t nz = 1;
  for (int d : other) nz *= s.ssize[d];
  const int64_t gx = (s.ssize[ld] + TL - 1) / TL, gy = s.ssize[in] / TI;
  k << "__device__ __forceinline__ void " << name << "(const float* __restrict__ in, " << elt
    << "* __restrict__ S, unsigned bid) {\n"
    << "  typedef " << elt << " T8 __attribute__((ext_vector_type(8)));\n"
    << "  __shared__ float tile[" << TI << "][" << TL + 1 << "];\n"
    << "  const int tid = threadIdx.x;\n"
    << "  const long l0 = (long)(bid % " << gx << "u) * " << TL << ", c0 = (long)((bid / "
    << gx << "u) % " << gy << "u) * " << TI << ";\n"
    << "  long z_ = bid / " << gx * gy << "u;\n";
  std::string ok, base, dbase;
  fixed_dims(k, op, s, other, ok, base, dbase);
  auto idx = [&](int d, const std::string &sv) {
    return "(" + sv + " - " + std::to_string(s.padlo[d]) + "L)";
  };
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
    << "    *(T8*)(S + dst0 + sl * " << s.sstride[ld] << "L + " << kin_dst(s, "c0 + g * 8")
    << ") = v;\n  }\n}\n";
  return gx * gy * nz;
}

// One thread per 8 staged elements along the innermost staged dim.  Lanes
// walk the dim that is contiguous in the source (when it is not the innermost
// staged dim) so the source reads coalesce.
int64_t staging_generic_body(std::stringstream &k, const std::string &name, const Operand &op,
                             const Staged &s, int64_t kin, const char *elt) {
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

  k << "__device__ __forceinline__ void " << name << "(const float* __restrict__ in, " << elt
    << "* __restrict__ S, unsigned bid) {\n"
    << "  typedef " << elt << " T8 __attribute__((ext_vector_type(8)));\n"
    << "  long t_ = (long)bid * 256 + threadIdx.x;\n"
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
    if (d == in)
      dst += " + " + kin_dst(s, "s" + std::to_string(d));
    else
      dst += " + s" + std::to_string(d) + " * " + std::to_string(s.sstride[d]) + "L";
  }
  k << "  const bool ok = " << ok << ";\n"
    << "  T8 v;\n"
    << "  #pragma unroll\n  for (int e = 0; e < 8; ++e)\n"
    << "    v[e] = ok && s" << in << " + e < " << kin << "L ? (" << elt << ")in[" << src
    << " + e * " << op.strides[in] << "L] : (" << elt << ")0.0f;\n"
    << "  *(T8*)(S + " << dst << ") = v;\n}\n";
  return (nvec + 255) / 256;
}

// Emits the staging of one operand as `__device__ void name(in, S, bid)` and
// returns its workgroup count; picks the plane, transpose or generic form.
int64_t staging_body(std::stringstream &k, const std::string &name, const Operand &op,
                     const Staged &s, int64_t kin, const char *elt, int rows) {
  const int nd = (int)op.sizes.size(), in = s.inner_dim;
  const int ld = source_contiguous_dim(op, s);
  if (ld < 0 || op.strides[ld] != 1 || s.ssize[in] % 32 != 0)
    return staging_generic_body(k, name, op, s, kin, elt);
  for (int a = 0; a < nd; ++a)
    if (a != ld && a != in && op.strides[a] == op.sizes[ld] && s.ssize[a] > 1 &&
        op.sizes[ld] <= 512)
      return staging_plane_body(k, name, op, s, ld, a, kin, elt, rows);
  if (s.ssize[ld] >= 32) return staging_transpose_body(k, name, op, s, ld, kin, elt);
  return staging_generic_body(k, name, op, s, kin, elt);
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

// Halo bound of the Q operand for K tiles of BK over a k-blocked staging
// whose first outer-K var is the block index and the rest are "taps": 0 when
// the halo form does not apply (taps that do not shift Q's rows forward, Q
// rows that are not non-decreasing in j, fewer than two taps), else the
// number of BK-vectors HL such that every tile's rows at every tap lie within
// HL vectors of the tile's first row.  ntap: taps per block.
int64_t halo_rows(const Contraction &c, const Staged &SQ, const std::vector<VarInfo> &KO,
                  int BN, int BK, int &ntap) {
  if (KO.size() < 2 || c.ND > (1LL << 24)) return 0;
  int64_t taps = 1, max_tap = 0;
  for (size_t q = 1; q < KO.size(); ++q) {
    taps *= KO[q].size;
    auto it = SQ.scoef.find(KO[q].sym);
    const int64_t cf = it == SQ.scoef.end() ? 0 : it->second;
    if (cf < 0 || cf % BK) return 0;
    max_tap += cf * (KO[q].size - 1);
  }
  if (taps < 2 || taps > (1 << 20)) return 0;
  // Q row offsets over j (row-major over the J vars, the last fastest)
  std::vector<int64_t> cf(c.J.size());
  for (size_t v = 0; v < c.J.size(); ++v) {
    auto it = SQ.scoef.find(c.J[v].sym);
    cf[v] = it == SQ.scoef.end() ? 0 : it->second;
  }
  std::vector<int64_t> idx(c.J.size(), 0);
  int64_t prev = 0, first = 0, span = 0;
  for (int64_t j = 0; j < c.ND; ++j) {
    int64_t r = 0;
    for (size_t v = 0; v < c.J.size(); ++v) r += cf[v] * idx[v];
    if (j > 0 && r < prev) return 0;
    if (j % BN == 0) first = r;
    if (r % BK) return 0;
    span = std::max(span, r - first);
    prev = r;
    for (int v = (int)c.J.size() - 1; v >= 0; --v) {  // next j
      if (++idx[v] < c.J[v].size) break;
      idx[v] = 0;
    }
  }
  ntap = (int)taps;
  return (span + max_tap) / BK + 1;
}

}  // namespace

bool try_contract_staged(const Dataflow &df, const Contraction &c, const Knobs &knobs,
                         Plan &plan) {
  const bool bf16 = knobs.bf16, f32 = !knobs.bf16 && !knobs.f16;
  const char *elt = bf16 ? "__bf16" : f32 ? "float" : "_Float16";
  const int EB = f32 ? 4 : 2;  // bytes per staged element
  // float operands: staging pays for itself on large nests only (small ones
  // stay on the single-launch gathered tile unless "contract=staged")
  if (f32 && !knobs.staged && c.MD * c.ND * c.KD * c.BD < (1LL << 28)) return false;
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
  // the K tile from the schedule: k_in split into nested loops above the
  // reduction, innermost extent 32 or 64 (a bk=N annotation takes precedence)
  const int64_t s_bk = knobs.bk > 0 ? 0 : df.split_extent(c.red_node, c.K[kin].sym);

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
    // 16-bit: up to 256 x 256; float: up to 128 x 128 (its fragments and
    // loads take twice the registers per k)
    const int cand16[][3] = {{256, 256, 32}, {128, 256, 32}, {256, 128, 32}, {128, 128, 64},
                             {128, 64, 32},  {64, 128, 32},  {64, 64, 32}};
    const int cand32[][3] = {{128, 128, 32}, {128, 64, 32}, {64, 128, 32}, {64, 64, 32},
                             {64, 64, 32},   {64, 64, 32},  {64, 64, 32}};
    for (auto &t : f32 ? cand32 : cand16) {
      if (t[0] > ((c.MD + 63) / 64) * 64 || t[1] > ((c.ND + 63) / 64) * 64) continue;
      if (tiles(t[0], t[1]) >= 256) {
        BM = t[0];
        BN = t[1];
        BK = t[2];
        break;
      }
    }
    if (knobs.bk > 0) BK = knobs.bk;
    // the direct-to-LDS tile (the default for 16-bit operands) runs K tiles
    // of 64: plain contractions (no outer-K positions, so no halo form) whose
    // k_in spans several such tiles
    if (knobs.bk <= 0 && !f32 && ko_n == 1 && (KI % 64 == 0 || KI >= 256) &&
        (knobs.lds_swz == -1 || knobs.lds_swz == 2) && !knobs.igemm_plain)
      BK = 64;
  }
  if (s_bk == 32 || s_bk == 64) BK = (int)s_bk;
  if (BK != 32 && BK != 64) return false;
  if (BM % 64 || BN % 64 || BM > 512 || BN > 512 || BM * BN > 256 * 256) return false;
  if (2 * (BM + BN) * (BK + 16 / EB) * EB > 160 * 1024) return false;  // LDS of a CU
  const int64_t KIP = (KI + BK - 1) / BK * BK;
  // k-blocked staging: K tiles of BK over k_in become separate outer-K
  // blocks, and a K tile's operand rows are adjacent BK-element runs
  // halo form: forced by igemm=halo|halo1, tried by default (one buffer)
  // for 16-bit operands, off with igemm=pipe
  const bool halo_forced = knobs.igemm_halo >= 1;
  const bool want_halo = !f32 && knobs.igemm_halo != 0;
  const bool blocked = (knobs.stage_blocked != 0 || halo_forced) && KIP > BK;
  const bool kb_first = knobs.kb_outer == 1 || (want_halo && knobs.kb_outer != 0);
  Staged st[2] = {stage_layout(c.ops[0], dims[0], KIP, blocked ? BK : 0, c.K),
                  stage_layout(c.ops[1], dims[1], KIP, blocked ? BK : 0, c.K)};
  const Staged &SP = st[c.po], &SQ = st[c.qo];
  int64_t KIT = KIP;  // inner-K extent seen by the tile
  if (blocked) {
    // the block index is one more outer-K var: slowest (kb=outer: all
    // outer-K positions of a block in a row) or fastest
    const int kb_sym = -1000;
    VarInfo kb{kb_sym, KIP / BK, "kb_"};
    if (kb_first) KO.insert(KO.begin(), kb);
    else KO.push_back(kb);
    ko_n *= KIP / BK;
    for (auto &x : st) x.scoef[kb_sym] = x.skb;
    KIT = BK;
  }
  const bool full = c.MD % BM == 0 && c.ND % BN == 0;
  const bool int32 = SP.total < (1LL << 31) && SQ.total < (1LL << 31);
  // the pipelined tiles address each operand with 32-bit byte offsets
  const bool pipe = (!knobs.igemm_plain || f32) && SP.total * EB < (1LL << 31) &&
                    SQ.total * EB < (1LL << 31);
  if (f32 && !pipe) return false;  // float operands: the pipelined tile only
  // pipelined tile: wave grid (4 waves as 2x2, or 8 as 2x4 / 4x2) and LDS rows
  int WGM = 2, WGN = 2;
  bool swz = false;
  if (pipe) {
    auto fits = [&](int wgm, int wgn) {
      const int NT = 64 * wgm * wgn, CPR = BK * EB / 16;
      return BM % (32 * wgm) == 0 && BN % (32 * wgn) == 0 && BM * CPR % NT == 0 &&
             BN * CPR % NT == 0 && BM * CPR >= NT && BN * CPR >= NT;
    };
    // default: 8 waves for a 256 x 256 tile (128 accumulators per lane
    // instead of 256, so the two register sets of the load pipeline fit
    // without spilling), else 4
    const int waves = knobs.waves > 0 ? knobs.waves : (BM * BN == 256 * 256 ? 8 : 4);
    if (waves == 8) {
      if (BN / 4 >= 32) WGN = 4;
      else WGM = 4;
    } else if (waves == 2) {  // two waves of (BM) x (BN/2): fewer LDS reads per MFMA
      WGM = 1;
    }
    if (!fits(WGM, WGN)) WGM = WGN = 2;  // the tile is too small for that wave grid
    if (!fits(WGM, WGN)) return false;
    swz = knobs.lds_swz == 1 && !f32;
  }
  // halo form: Q rows of all taps of a K block loaded once per block
  int64_t HL = 0;
  int ntap = 0;
  const int HB = knobs.igemm_halo == 1 ? 2 : 1;  // halo buffers
  if (pipe && want_halo && blocked && kb_first) HL = halo_rows(c, SQ, KO, BN, BK, ntap);
  if (HL) {
    // eight waves (2 x 4) where the tile allows, else four (2 x 2)
    auto fits = [&](int wgm, int wgn) {
      const int NT = 64 * wgm * wgn;
      return BM % (32 * wgm) == 0 && BN % (32 * wgn) == 0 && BM * (BK / 8) % NT == 0 &&
             BM * (BK / 8) >= NT;
    };
    WGM = 2;
    WGN = knobs.waves != 4 && fits(2, 4) ? 4 : 2;
    if (!fits(WGM, WGN) || (2 * BM + HB * HL) * (BK + 8) * 2 > 160 * 1024) HL = 0;
    if (HL) swz = false;  // the register-staged halo tile pads its LDS rows
  }
  if (halo_forced && !HL) return false;
  // direct-to-LDS tile: 16-bit pipelined tiles outside the halo form, unless
  // lds=swz|pad asks for the register-staged one
  const bool dl = pipe && !f32 && !HL && (knobs.lds_swz == -1 || knobs.lds_swz == 2);
  // halo=direct: the halo tile reads Q's float32 source itself (no staged
  // copy of Q; its rows are rounded to 16 bits as they go to LDS)
  const Operand &QOP = c.ops[c.qo];
  const int64_t q_src_bytes =
      df.numel(df.nodes.at(df.ir().inputs()[QOP.input_index]).vars) * 4;
  // (one value per tap over a block's first 8 taps: windows of 9 or more taps)
  const int64_t qf_lds = 2 * 64 * 4 * (int64_t)(WGM * WGN) *
                         (((HL + 63) / 64 * (BK / 8) + WGM * WGN - 1) / (WGM * WGN));
  const bool qdirect = HL && knobs.halo_direct == 1 && q_src_bytes < (1LL << 31) &&
                       ntap >= 9 &&
                       (2 * BM + HB * HL) * (BK + 8) * 2 + qf_lds <= 160 * 1024;
  if (knobs.halo_direct == 1 && HL && !qdirect) return false;
  // direct-to-LDS halo tile (lds=dl): P stages and halo buffers filled by
  // loads that write LDS; the halo rounded up to whole load rounds
  const int64_t hthreads = 64 * WGM * WGN, hcpr = BK / 8;
  const int64_t HLP = HL * hcpr >= hthreads ? HL : hthreads / hcpr;
  const bool hdl = HL && !qdirect && knobs.lds_swz == 2 &&
                   2 * (BM + HLP) * BK * 2 <= 160 * 1024 &&
                   (HL * hcpr + hthreads - 1) / hthreads <= 15;
  const int lds = hdl   ? (int)(2 * (BM + HLP) * BK * 2)
                  : HL  ? (int)((2 * BM + HB * HL) * (BK + 8) * 2 + (qdirect ? qf_lds : 0))
                  : f32 ? 2 * (BM + BN) * (BK + 4) * 4
                        : 2 * (BM + BN) * (swz || dl ? BK : BK + 8) * 2;
  const int threads = 64 * WGM * WGN;
  const int64_t gx = (c.ND + BN - 1) / BN, gy = (c.MD + BM - 1) / BM;
  ASSERT(gx < (1LL << 31) && gy < 65536 && c.BD < 65536) << "contraction grid too large";
  // stream-K (float tile): when the tiles fill the 2 x 256 workgroup slots
  // unevenly, SK workgroups share all tiles' K tiles evenly instead; every
  // range spans at least a whole tile's K, so a tile has at most two
  // contributors, whose partial sums are added with float atomics to an
  // output that the staging launch zeroes (two commutative adds: the result
  // does not depend on their order)
  const int64_t SK = 512, ntile = gx * gy, KTT = ko_n * (KIT / BK);
  const int64_t sk_rounds = (ntile + SK - 1) / SK;
  const bool streamk = f32 && pipe && c.BD == 1 && knobs.streamk != 0 && ntile >= SK &&
                       (knobs.streamk == 1 || ntile * 10 < sk_rounds * SK * 9);

  // ---- buffers: one staged (16-bit) scratch buffer per operand ----
  int sbuf[2];
  int scratch_no = 0;
  for (const auto &b : plan.buffers) scratch_no += b.kind == BufferSpec::SCRATCH;
  for (int o = 0; o < 2; ++o) {
    sbuf[o] = (int)plan.buffers.size();
    int64_t words = o == c.qo && qdirect ? 8 : ((st[o].total * EB + 3) / 4 + 7) / 8 * 8;
    auto nr = df.ir().inputs().at(c.ops[o].input_index);
    plan.buffers.push_back({BufferSpec::SCRATCH, scratch_no++, words, nr});
  }
  const int out_buf = df.num_inputs;  // the single output
  // stream-K: the tiles that two workgroups share (a range boundary falls
  // inside their K); the staging launch zeroes their output blocks
  std::vector<int64_t> split_tiles;
  if (streamk)
    for (int64_t v = 1; v < SK; ++v) {
      const int64_t b = ntile * KTT * v / SK;
      if (b % KTT && (split_tiles.empty() || split_tiles.back() != b / KTT))
        split_tiles.push_back(b / KTT);
    }
  const int64_t nzero = (int64_t)split_tiles.size();  // one workgroup per tile

  std::stringstream k;
  k << kernel_prelude() << igemm_tile_source() << "\n";
  // geometry of the tile
  k << "struct LtGeo {\n"
    << "  static constexpr int KO = " << ko_n << ", KI = " << KIT << ";\n"
    << "  static constexpr long MD = " << c.MD << "L, ND = " << c.ND << "L;\n"
    << "  static constexpr int NTAP = " << ntap << ";\n"
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
    << lin_over(c.ostride, c.J) << ";\n  }\n";
  if (qdirect) {
    // staged row r of a block -> source offset of its first k_in value;
    // ok false on a zero halo row
    const int in = SQ.inner_dim;
    std::vector<int> rd;
    for (int d : SQ.order)
      if (d != in) rd.push_back(d);
    k << "  static constexpr long QCS = " << QOP.strides[in] << "L, QKIN = " << QOP.sizes[in]
      << "L;\n"
      << "  __device__ long q_src(long r, bool& ok) const {\n    ok = true;\n    long src = 0;\n";
    for (int q = (int)rd.size() - 1; q >= 0; --q) {
      const int d = rd[q];
      const std::string sv = "s" + std::to_string(d);
      if (q > 0)
        k << "    const long " << sv << " = r % " << SQ.ssize[d] << "L - " << SQ.padlo[d]
          << "L; r /= " << SQ.ssize[d] << "L;\n";
      else
        k << "    const long " << sv << " = r - " << SQ.padlo[d] << "L;\n";
      k << "    ok = ok && " << sv << " >= 0 && " << sv << " < " << QOP.sizes[d] << "L;\n"
        << "    src += " << sv << " * " << QOP.strides[d] << "L;\n";
    }
    k << "    return src;\n  }\n";
  }
  k << "};\n";
  // ---- both staging passes in one launch: workgroups [0, nb[0]) stage
  // operand 0, the rest operand 1 ----
  int64_t nb[2];
  for (int o = 0; o < 2; ++o) {
    nb[o] = staging_body(k, "lt_stage_op" + std::to_string(o), c.ops[o], st[o], KI, elt,
                         knobs.stage_rows);
    if (o == c.qo && qdirect) nb[o] = 0;  // the tile reads Q's source
  }
  ASSERT(nb[0] + nb[1] + nzero < (1LL << 31)) << "staging grid too large";
  k << "static constexpr unsigned lt_stage_blocks[2] = {" << nb[0] << "u, " << nb[1] << "u};\n"
    << "extern \"C\" __global__ void __launch_bounds__(256) lt_stage0(const float* __restrict__ "
       "in0, "
    << elt << "* __restrict__ S0, const float* __restrict__ in1, " << elt
    << "* __restrict__ S1" << (streamk ? ", float* __restrict__ zo" : "") << ") {\n"
    << "  if (blockIdx.x < lt_stage_blocks[0]) lt_stage_op0(in0, S0, blockIdx.x);\n";
  if (nzero) {
    k << "  else if (blockIdx.x >= lt_stage_blocks[0] + lt_stage_blocks[1]) {\n"
      << "    // zero a split tile's output block for the stream-K atomic adds\n"
      << "    static constexpr unsigned split_[" << nzero << "] = {";
    for (size_t q = 0; q < split_tiles.size(); ++q) k << (q ? ", " : "") << split_tiles[q] << "u";
    k << "};\n"
      << "    const unsigned t_ = split_[blockIdx.x - lt_stage_blocks[0] - lt_stage_blocks[1]];\n"
      << "    const long i0 = (long)(t_ / " << gx << "u) * " << BM << ", j0 = (long)(t_ % " << gx
      << "u) * " << BN << ";\n"
      << "    LtGeo g;\n    g.zp = g.zq = g.zo = 0;\n"
      << "    for (int e = threadIdx.x; e < " << BM * BN << "; e += 256) {\n"
      << "      const long i = i0 + e / " << BN << ", j = j0 + e % " << BN << ";\n"
      << "      if (i < LtGeo::MD && j < LtGeo::ND) zo[g.out_row(i) + g.out_col(j)] = 0.f;\n"
      << "    }\n  }\n";
  }
  k << "  else lt_stage_op1(in1, S1, blockIdx.x - lt_stage_blocks[0]);\n}\n";
  {
    Stage s;
    s.kernel = "lt_stage0";
    s.kind = "stage";
    s.buffers = {c.ops[0].input_index, sbuf[0], c.ops[1].input_index, sbuf[1]};
    s.writes = {sbuf[0], sbuf[1]};
    if (streamk) {  // (unused by the launch when no tile is shared)
      s.buffers.push_back(out_buf);
      s.writes.push_back(out_buf);
      s.bytes += nzero * BM * BN * 4;
    }
    s.grid[0] = (unsigned)(nb[0] + nb[1] + nzero);
    s.block[0] = 256;
    for (int o = 0; o < 2; ++o)
      if (!(o == c.qo && qdirect))
        s.bytes += df.numel(df.nodes.at(df.ir().inputs()[c.ops[o].input_index]).vars) * 4 +
                   st[o].total * EB;
    plan.stages.push_back(s);
  }
  // eight waves: keep two workgroups per CU (four waves per SIMD, at most
  // 128 VGPRs each) when their LDS fits and a wave's accumulators take at
  // most 64 of those registers
  const bool two_wg = threads == 512 && 2 * lds <= 160 * 1024 &&
                      (BM / WGM) * (BN / WGN) / 64 <= 64;
  k << "extern \"C\" __global__ void __launch_bounds__(" << threads << ")"
    << (two_wg    ? " __attribute__((amdgpu_waves_per_eu(4)))"
        : streamk ? " __attribute__((amdgpu_waves_per_eu(2)))"  // two 4-wave workgroups per CU
                  : "")
    << " lt_igemm0(const " << elt
    << "* __restrict__ P, const " << elt << "* __restrict__ Q, float* __restrict__ out"
    << (qdirect ? ", const float* __restrict__ Qf" : "") << ") {\n"
    << "  __shared__ __attribute__((aligned(16))) unsigned char smem[" << lds << "];\n"
    << "  LtGeo g;\n"
    << decls(c.B, "  ") << decompose("(long)blockIdx.z", c.B, "  ")
    << "  g.zp = " << lin_over(SP.scoef, c.B) << ";\n"
    << "  g.zq = " << lin_over(SQ.scoef, c.B) << ";\n"
    << "  g.zo = " << lin_over(c.ostride, c.B) << ";\n"
    ;
  const std::string i0 = "(long)blockIdx.y * " + std::to_string(BM);
  const std::string j0 =
      "(long)lt_igemm_swizzle(blockIdx.x, gridDim.x) * " + std::to_string(BN);
  const std::string ind = "  ";
  if (streamk)
    k << "  const unsigned long U_ = " << ntile * KTT << "ul;\n"
      << "  const unsigned v_ = lt_igemm_swizzle(blockIdx.x, gridDim.x);\n"
      << "  unsigned long u_ = U_ * v_ / " << SK << "ul;\n"
      << "  const unsigned long ue_ = U_ * (v_ + 1) / " << SK << "ul;\n"
      << "  while (u_ < ue_) {\n"
      << "    const unsigned t_ = (unsigned)(u_ / " << KTT << "ul), k0_ = (unsigned)(u_ % " << KTT
      << "ul);\n"
      << "    const unsigned long r_ = " << KTT << "ul - k0_;\n"
      << "    const unsigned n_ = (unsigned)(r_ < ue_ - u_ ? r_ : ue_ - u_);\n"
      << "    lt_igemm_tile_f32<" << BM << ", " << BN << ", " << BK << ", "
      << (full ? "true" : "false") << ", " << WGM << ", " << WGN << ">(g, P, Q, out, (long)(t_ / "
      << gx << "u) * " << BM << ", (long)(t_ % " << gx << "u) * " << BN << ", smem, "
      << SP.total * 4 << "u, " << SQ.total * 4 << "u, (int)k0_, (int)n_, n_ != " << KTT
      << "u);\n"
      << "    __syncthreads();  // the next tile rewrites this tile's LDS\n"
      << "    u_ += n_;\n  }\n";
  else if (f32)
    k << ind << "lt_igemm_tile_f32<" << BM << ", " << BN << ", " << BK << ", "
      << (full ? "true" : "false") << ", " << WGM << ", " << WGN << ">(g, P, Q, out, " << i0
      << ", " << j0 << ", smem, " << SP.total * 4 << "u, " << SQ.total * 4 << "u);\n";
  else if (hdl)
    k << ind << "lt_igemm_tile_halo_dl<" << elt << ", " << BM << ", " << BN << ", " << BK << ", "
      << (full ? "true" : "false") << ", " << WGM << ", " << WGN << ", " << HL
      << ">(g, P, Q, out, " << i0 << ", " << j0 << ", smem, " << SP.total * 2 << "u, "
      << SQ.total * 2 << "u);\n";
  else if (HL && qdirect)
    k << ind << "lt_igemm_tile_halo<" << elt << ", " << BM << ", " << BN << ", " << BK << ", "
      << (full ? "true" : "false") << ", " << WGM << ", " << WGN << ", " << HL << ", " << HB
      << ", LtGeo, true>(g, P, Q, out, " << i0 << ", " << j0 << ", smem, " << SP.total * 2
      << "u, 0u, Qf, " << q_src_bytes << "u);\n";
  else if (HL)
    k << ind << "lt_igemm_tile_halo<" << elt << ", " << BM << ", " << BN << ", " << BK << ", "
      << (full ? "true" : "false") << ", " << WGM << ", " << WGN << ", " << HL << ", " << HB
      << ">(g, P, Q, out, " << i0 << ", " << j0 << ", smem, " << SP.total * 2 << "u, "
      << SQ.total * 2 << "u);\n";
  else if (dl)
    k << ind << "lt_igemm_tile_dl<" << elt << ", " << BM << ", " << BN << ", " << BK << ", "
      << (full ? "true" : "false") << ", " << WGM << ", " << WGN << ">(g, P, Q, out, " << i0
      << ", " << j0 << ", smem, " << SP.total * 2 << "u, " << SQ.total * 2 << "u);\n";
  else if (pipe)
    k << ind << "lt_igemm_tile_pipe<" << elt << ", " << BM << ", " << BN << ", " << BK << ", "
      << (full ? "true" : "false") << ", " << WGM << ", " << WGN << ", "
      << (swz ? "true" : "false") << ">(g, P, Q, out, " << i0 << ", " << j0 << ", smem, "
      << SP.total * 2 << "u, " << SQ.total * 2 << "u);\n";
  else
    k << ind << "lt_igemm_tile<" << elt << ", " << BM << ", " << BN << ", " << BK << ", "
      << (full ? "true" : "false") << ", " << (int32 ? "int" : "long") << ">(g, P, Q, out, "
      << i0 << ", " << j0 << ", smem);\n";
  k << "}\n";

  Stage g;
  g.kernel = "lt_igemm0";
  g.kind = bf16 ? "contract_bf16" : f32 ? "contract_f32" : "contract_f16";
  g.buffers = {sbuf[c.po], sbuf[c.qo], out_buf};
  if (qdirect) g.buffers.push_back(QOP.input_index);
  g.writes = {out_buf};
  g.grid[0] = (unsigned)(streamk ? SK : gx);
  g.grid[1] = (unsigned)(streamk ? 1 : gy);
  g.grid[2] = (unsigned)c.BD;
  g.block[0] = (unsigned)threads;
  g.flops = 2 * c.MD * c.ND * c.KD * c.BD;
  g.bytes = (qdirect ? SP.total * EB + q_src_bytes : (SP.total + SQ.total) * EB) +
            c.MD * c.ND * c.BD * 4;
  plan.stages.push_back(g);
  plan.config["tile_m"] = BM;
  plan.config["tile_n"] = BN;
  plan.config["bk"] = BK;
  plan.config["waves"] = pipe ? WGM * WGN : 4;
  plan.config["dl"] = dl ? 1 : 0;
  plan.source = k.str();
  std::stringstream desc;
  desc << "[hip] contraction on MFMA (" << (bf16 ? "bf16" : f32 ? "f32" : "f16")
       << " staged operands, f32 accumulate): D[" << c.BD << "][" << c.MD << "][" << c.ND
       << "] = P x Q over K=" << ko_n << "x" << (blocked ? KIT : KI) << (KIP != KI && !blocked ? " (padded to " + std::to_string(KIP) + ")" : std::string()) << ", tile " << BM << "x" << BN << "x"
       << BK << (pipe ? " pipelined, " + std::to_string(WGM * WGN) + " waves" +
                            (dl ? ", direct-to-LDS loads" : swz ? ", swizzled LDS" : "")
                      : std::string())
       << (blocked ? std::string(", k-blocked staging (kb ") + (kb_first ? "outer)" : "inner)")
                   : std::string())
       << (streamk ? ", stream-K over " + std::to_string(SK) + " workgroups" : std::string())
       << (HL ? ", Q halo of " + std::to_string(HL) + " rows per K block" +
                    (HB == 1 ? " (one buffer)" : "") +
                    (qdirect ? " read from the float source" : "")
              : std::string())
       << "; staged P " << SP.total << " and Q " << SQ.total
       << " elements\n";
  plan.description = desc.str();
  return true;
}

}  // namespace hip
}  // namespace loop_tool
