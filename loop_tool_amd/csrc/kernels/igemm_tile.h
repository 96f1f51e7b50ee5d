// Implicit-GEMM tile on the gfx950 matrix cores for 16-bit operands with
// float32 accumulation:
//     out[i, j] = sum_k P[i, k] * Q[j, k]
// where P and Q are *staged* operands: 16-bit tensors in HBM laid out so that
// the address of (row, k) separates as
//     row_base(row) + outer_offset(k / KI) + k % KI
// i.e. the innermost reduction var (extent KI) is the contiguous dim and
// every other var enters through an affine row or outer-K offset.  Staging
// pads the inner dim with zeros to a multiple of BK, so KI % BK == 0 and no
// K bounds checks remain (the check below is kept for callers that do not).  A matmul operand A[M][K] is already staged (KO = 1); a convolution
// input X[N][C][H][W] staged as the zero-haloed X'[N][H+2][W+2][C] gives
// row_base(n, ho, wo) + outer_offset(kh, kw) + ci with no bounds checks.
//
// The geometry policy G supplies the index maps (constants when generated per
// nest and compiled by hipRTC, run-time fields in the prebuilt library):
//   g.KO, g.KI                        outer-K count and inner-K extent
//   g.p_row(i, ok), g.q_row(j, ok)    staged row offsets (elements), row valid
//   g.p_koff(ko), g.q_koff(ko)        outer-K offsets
//   g.out_row(i), g.out_col(j)        output offsets (elements); MD, ND bounds
//
// A workgroup of 256 lanes (4 waves as 2 x 2) computes a BM x BN tile; each
// wave owns (BM/2) x (BN/2) as 32x32 blocks of v_mfma_f32_32x32x16_{bf16,f16}.
// K advances BK (32 or 64) at a time through double-buffered LDS: one barrier
// per K tile, next tile's 16-byte global loads in flight during the MFMAs.
// Rows of LDS are padded by 16 bytes so the 16-byte MFMA operand reads of 32
// consecutive rows spread over the banks.
//
// lt_igemm_swizzle() maps workgroup ids to tiles so that each of the chip's
// 8 XCDs (which receive workgroups round-robin) works on a contiguous run of
// tiles: neighbouring tiles share operand rows (a convolution's halo rows),
// and each XCD has its own L2.
#ifndef LT_IGEMM_TILE_H_
#define LT_IGEMM_TILE_H_

typedef __bf16 lt_ig_bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 lt_ig_f16x8 __attribute__((ext_vector_type(8)));
typedef float lt_ig_f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short lt_ig_u16x8 __attribute__((ext_vector_type(8)));

template <typename T> struct lt_ig_vec8;
template <> struct lt_ig_vec8<__bf16> { typedef lt_ig_bf16x8 type; };
template <> struct lt_ig_vec8<_Float16> { typedef lt_ig_f16x8 type; };

__device__ __forceinline__ lt_ig_f32x16 lt_ig_mfma(lt_ig_bf16x8 a, lt_ig_bf16x8 b,
                                                   lt_ig_f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ lt_ig_f32x16 lt_ig_mfma(lt_ig_f16x8 a, lt_ig_f16x8 b,
                                                   lt_ig_f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// workgroup id -> tile id, contiguous per XCD (bijective for any count)
__device__ __forceinline__ unsigned lt_igemm_swizzle(unsigned b, unsigned n) {
  const unsigned q = n / 8, r = n % 8, x = b % 8;
  return x * q + (x < r ? x : r) + b / 8;
}

// LDS bytes of one tile configuration (double-buffered P and Q tiles)
template <int BM, int BN, int BK>
__host__ __device__ constexpr int lt_igemm_lds_bytes() {
  return 2 * (BM + BN) * (BK + 8) * 2;
}

// Row offsets are computed once per tile in IT (int when every staged
// operand has fewer than 2^31 elements); FULL drops the row bounds checks
// when the tile grid divides the problem exactly.  The next k16 step's MFMA
// fragments are read from LDS while the current step's MFMAs run.
template <typename T, int BM, int BN, int BK, bool FULL = false, typename IT = long, class G>
__device__ __forceinline__ void lt_igemm_tile(const G& g, const T* __restrict__ P,
                                              const T* __restrict__ Q,
                                              float* __restrict__ out, long i0, long j0,
                                              unsigned char* smem) {
  typedef typename lt_ig_vec8<T>::type V;
  constexpr int NT = 256, LDR = BK + 8;  // LDS row (16-bit elements)
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 32, TN = WN / 32;
  constexpr int CPR = BK / 8, KK = BK / 16;   // 16-byte chunks per row, k16 steps
  constexpr int VP = BM * CPR / NT, VQ = BN * CPR / NT;
  static_assert(VP >= 1 && VQ >= 1 && BM * CPR % NT == 0 && BN * CPR % NT == 0,
                "tile too small for 256 lanes");
  const int KIT = (g.KI + BK - 1) / BK;  // K tiles per outer-K index
  const int KT = g.KO * KIT;
  T* sP = (T*)smem;              // [2][BM][LDR]
  T* sQ = sP + 2 * BM * LDR;     // [2][BN][LDR]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int cc = (tid % CPR) * 8;  // this lane's k chunk within a K tile

  // row offsets of the rows this lane loads (fixed for the whole K loop)
  IT pb[VP], qb[VQ];
  bool pok[VP], qok[VQ];
#pragma unroll
  for (int v = 0; v < VP; ++v) pb[v] = (IT)g.p_row(i0 + (tid + v * NT) / CPR, pok[v]);
#pragma unroll
  for (int v = 0; v < VQ; ++v) qb[v] = (IT)g.q_row(j0 + (tid + v * NT) / CPR, qok[v]);

  lt_ig_u16x8 rp[VP], rq[VQ];
  auto load = [&](int kt) {
    const int ko = kt / KIT, ki = (kt % KIT) * BK + cc;
    const bool kok = (g.KI % BK == 0) || ki < g.KI;
    const IT po = (IT)g.p_koff(ko) + ki, qo = (IT)g.q_koff(ko) + ki;
#pragma unroll
    for (int v = 0; v < VP; ++v)
      rp[v] = ((FULL || pok[v]) && kok) ? *(const lt_ig_u16x8*)(P + (pb[v] + po))
                                        : (lt_ig_u16x8){};
#pragma unroll
    for (int v = 0; v < VQ; ++v)
      rq[v] = ((FULL || qok[v]) && kok) ? *(const lt_ig_u16x8*)(Q + (qb[v] + qo))
                                        : (lt_ig_u16x8){};
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int v = 0; v < VP; ++v)
      *(lt_ig_u16x8*)(sP + (buf * BM + (tid + v * NT) / CPR) * LDR + cc) = rp[v];
#pragma unroll
    for (int v = 0; v < VQ; ++v)
      *(lt_ig_u16x8*)(sQ + (buf * BN + (tid + v * NT) / CPR) * LDR + cc) = rq[v];
  };
  V fa[2][TM], fb[2][TN];
  auto frags = [&](int s, int buf, int kk) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
      fa[s][a] = *(const V*)(sP + (buf * BM + wm * WM + a * 32 + (lane & 31)) * LDR +
                             kk * 16 + (lane >> 5) * 8);
#pragma unroll
    for (int b = 0; b < TN; ++b)
      fb[s][b] = *(const V*)(sQ + (buf * BN + wn * WN + b * 32 + (lane & 31)) * LDR +
                             kk * 16 + (lane >> 5) * 8);
  };

  lt_ig_f32x16 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = (lt_ig_f32x16){};

  int buf = 0;
  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < KT; ++kt) {
    const bool more = kt + 1 < KT;
    if (more) load(kt + 1);
    frags(0, buf, 0);
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
      if (kk + 1 < KK) frags((kk + 1) & 1, buf, kk + 1);
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
          acc[a][b] = lt_ig_mfma(fa[kk & 1][a], fb[kk & 1][b], acc[a][b]);
    }
    if (more) {
      buf ^= 1;
      store(buf);
    }
    __syncthreads();
  }
  // 32x32 accumulator layout: col = lane % 32, row = (r/4)*8 + (lane/32)*4 + r%4
#pragma unroll
  for (int b = 0; b < TN; ++b) {
    const long j = j0 + wn * WN + b * 32 + (lane & 31);
    if (!FULL && j >= g.ND) continue;
    const long oj = g.out_col(j);
#pragma unroll
    for (int a = 0; a < TM; ++a) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const long i = i0 + wm * WM + a * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (FULL || i < g.MD)
          __builtin_nontemporal_store(acc[a][b][r], out + g.out_row(i) + oj);
      }
    }
  }
}

#endif  // LT_IGEMM_TILE_H_
