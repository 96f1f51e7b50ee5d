// Implicit-GEMM tile on the gfx950 matrix cores for 16-bit operands with
// float32 accumulation:
//     out[i, j] = sum_k P[i, k] * Q[j, k]
// where P and Q are *staged* operands: 16-bit tensors in HBM laid out so that
// the address of (row, k) separates as
//     row_base(row) + outer_offset(k / KI) + k % KI
// i.e. the innermost reduction var (extent KI) is the contiguous dim and
// every other var enters through an affine row or outer-K offset.  Staging
// pads the inner dim with zeros to a multiple of BK, so KI % BK == 0 and no
// K bounds checks remain (lt_igemm_tile keeps the check for callers that do
// not).  A matmul operand A[M][K] is already staged (KO = 1); a convolution
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
// per K tile, next tile's 16-byte global loads in flight during the MFMAs
// (lt_igemm_tile_pipe: two tiles ahead, with a pinned instruction interleave).
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
typedef unsigned int lt_ig_u32x4 __attribute__((ext_vector_type(4)));

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

// Writes a wave's TM x TN accumulator blocks (rows from i0, columns from j0).
// 32x32 accumulator layout: col = lane % 32, row = (r/4)*8 + (lane/32)*4 + r%4.
// With `add`, the blocks are added to `out` with float atomics instead (a
// tile whose K range two workgroups share; `out` starts at zero).
template <int TM, int TN, bool FULL, class G>
__device__ __forceinline__ void lt_igemm_store(const G& g, const lt_ig_f32x16 (&acc)[TM][TN],
                                               float* __restrict__ out, long i0, long j0,
                                               int lane, bool add = false) {
#pragma unroll
  for (int b = 0; b < TN; ++b) {
    const long j = j0 + b * 32 + (lane & 31);
    if (!FULL && j >= g.ND) continue;
    const long oj = g.out_col(j);
#pragma unroll
    for (int a = 0; a < TM; ++a) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const long i = i0 + a * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (FULL || i < g.MD) {
          if (add)
            unsafeAtomicAdd(out + g.out_row(i) + oj, acc[a][b][r]);
          else
            __builtin_nontemporal_store(acc[a][b][r], out + g.out_row(i) + oj);
        }
      }
    }
  }
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
  lt_igemm_store<TM, TN, FULL>(g, acc, out, i0 + wm * WM, j0 + wn * WN, lane);
}

// Software-pipelined form of the same tile, for staged operands of fewer than
// 2^31 bytes (pbytes, qbytes: operand sizes in bytes) with KI % BK == 0, or,
// with KTAIL, KI % 8 == 0: the last K tile then zeroes the 16-byte chunks
// past KI (they hold the next row's values, or the buffer range check's 0).
//  * WGM x WGN waves (2, 4 or 8) split the tile; each wave owns
//    (BM/WGM) x (BN/WGN) as 32x32 MFMA blocks.  Eight waves put two waves on
//    every SIMD, so one wave's MFMAs cover the other's LDS and load waits;
//    two waves of larger blocks read fewer LDS bytes per MFMA.
//  * Global loads are buffer loads: a 32-bit byte offset per row in a VGPR
//    and the K tile's offset in an SGPR; rows outside the problem get an
//    offset beyond the buffer, which the hardware reads as 0.
//  * Loads run two K tiles ahead into two register sets, so a tile's loads
//    have a whole K tile of MFMAs to land before they are stored to LDS.
//  * Within a K tile, the next k16 step's fragments are read during the
//    current step's MFMAs, and the next tile's LDS stores are interleaved
//    with the last step's MFMAs; one barrier per K tile.
//  * sched_group_barrier pins the interleave in the steady state: one MFMA,
//    then one fragment read and one global load (or one LDS store).
//  * LDS rows are either padded by 16 bytes (SWZ false) or unpadded with the
//    16-byte chunks of each row XOR-permuted (SWZ true): chunk c of row r is
//    stored at chunk c ^ ((r / RP) % CPR), RP = rows per 256 bytes (the 64
//    four-byte banks of the gfx950 LDS).  Any 16 consecutive lanes of a
//    fragment read (16 rows, one logical chunk) and of a tile store (whole
//    rows) then touch 16 distinct 16-byte bank groups.
// Each lane's working set: TM*TN*16 accumulators, 2*(TM+TN) fragments of 4
// VGPRs and 2*(VP+VQ) load registers of 4.  LDS: 2 * (BM + BN) rows of BK
// (SWZ) or BK + 8 16-bit elements.

template <int BK, bool SWZ>
__device__ __forceinline__ int lt_ig_lds_off(int row, int chunk) {
  constexpr int CPR = BK / 8;
  if (SWZ) {
    constexpr int RP = 128 / BK;  // rows per 256 bytes: 4 (BK 32) or 2 (BK 64)
    return row * BK + ((chunk ^ ((row / RP) % CPR)) << 3);
  }
  return row * (BK + 8) + (chunk << 3);
}

template <typename T, int BM, int BN, int BK, bool FULL, int WGM, int WGN, bool SWZ,
          bool KTAIL = false, class G>
__device__ __forceinline__ void lt_igemm_tile_pipe(const G& g, const T* __restrict__ P,
                                                   const T* __restrict__ Q,
                                                   float* __restrict__ out, long i0, long j0,
                                                   unsigned char* smem, unsigned pbytes,
                                                   unsigned qbytes) {
  typedef typename lt_ig_vec8<T>::type V;
  constexpr int NT = 64 * WGM * WGN, LDR = SWZ ? BK : BK + 8;
  constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / 32, TN = WN / 32;
  constexpr int CPR = BK / 8, KK = BK / 16;
  constexpr int VP = BM * CPR / NT, VQ = BN * CPR / NT;
  static_assert(TM >= 1 && TN >= 1 && WM % 32 == 0 && WN % 32 == 0, "wave tile");
  static_assert(VP >= 1 && VQ >= 1 && BM * CPR % NT == 0 && BN * CPR % NT == 0,
                "tile too small for the workgroup");
  static_assert(KK % 2 == 0, "BK must be a multiple of 32");
  const int KIT = KTAIL ? (g.KI + BK - 1) / BK : g.KI / BK;
  const int KT = g.KO * KIT;
  T* sP = (T*)smem;              // [2][BM][LDR]
  T* sQ = sP + 2 * BM * LDR;     // [2][BN][LDR]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int ck = tid % CPR, cc = ck * 8;
  const auto rsP = __builtin_amdgcn_make_buffer_rsrc((void*)P, (short)0, (int)pbytes, 0x00020000);
  const auto rsQ = __builtin_amdgcn_make_buffer_rsrc((void*)Q, (short)0, (int)qbytes, 0x00020000);
  unsigned pv[VP], qv[VQ];
  int sp[VP], sq[VQ];  // LDS element offsets of this lane's tile stores
#pragma unroll
  for (int v = 0; v < VP; ++v) {
    bool ok;
    const int r = (tid + v * NT) / CPR;
    const long o = g.p_row(i0 + r, ok);
    pv[v] = (FULL || ok) ? (unsigned)((o + cc) * 2) : 0x80000000u;
    sp[v] = lt_ig_lds_off<BK, SWZ>(r, ck);
  }
#pragma unroll
  for (int v = 0; v < VQ; ++v) {
    bool ok;
    const int r = (tid + v * NT) / CPR;
    const long o = g.q_row(j0 + r, ok);
    qv[v] = (FULL || ok) ? (unsigned)((o + cc) * 2) : 0x80000000u;
    sq[v] = lt_ig_lds_off<BK, SWZ>(r, ck);
  }
  // fragment rows of this lane; the chunk of k16 step kk is 2 * kk + lane / 32
  const int fr = lane & 31, fh = lane >> 5;
  lt_ig_u32x4 r0p[VP], r0q[VQ], r1p[VP], r1q[VQ];
  auto load = [&](int kt, lt_ig_u32x4* rp, lt_ig_u32x4* rq) {
    const int ko = kt / KIT, kb = (kt % KIT) * BK;
    const int po = (int)((g.p_koff(ko) + kb) * 2), qo = (int)((g.q_koff(ko) + kb) * 2);
#pragma unroll
    for (int v = 0; v < VP; ++v)
      rp[v] = __builtin_amdgcn_raw_buffer_load_b128(rsP, (int)pv[v], po, 0);
#pragma unroll
    for (int v = 0; v < VQ; ++v)
      rq[v] = __builtin_amdgcn_raw_buffer_load_b128(rsQ, (int)qv[v], qo, 0);
    if (KTAIL && kb + cc >= g.KI) {  // this lane's chunk lies past the K extent
#pragma unroll
      for (int v = 0; v < VP; ++v) rp[v] = (lt_ig_u32x4){};
#pragma unroll
      for (int v = 0; v < VQ; ++v) rq[v] = (lt_ig_u32x4){};
    }
  };
  auto store = [&](int buf, const lt_ig_u32x4* rp, const lt_ig_u32x4* rq) {
#pragma unroll
    for (int v = 0; v < VP; ++v) *(lt_ig_u32x4*)(sP + buf * BM * LDR + sp[v]) = rp[v];
#pragma unroll
    for (int v = 0; v < VQ; ++v) *(lt_ig_u32x4*)(sQ + buf * BN * LDR + sq[v]) = rq[v];
  };
  V fa[2][TM], fb[2][TN];
  auto frags = [&](int buf, int kk, int s) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
      fa[s][a] = *(const V*)(sP + buf * BM * LDR +
                             lt_ig_lds_off<BK, SWZ>(wm * WM + a * 32 + fr, 2 * kk + fh));
#pragma unroll
    for (int b = 0; b < TN; ++b)
      fb[s][b] = *(const V*)(sQ + buf * BN * LDR +
                             lt_ig_lds_off<BK, SWZ>(wn * WN + b * 32 + fr, 2 * kk + fh));
  };
  lt_ig_f32x16 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = (lt_ig_f32x16){};
  // One K tile: tile kt is in LDS buffer B and its k16-step-0 fragments in
  // fa/fb[0]; RN holds tile kt+1 (loaded a tile ago); RF receives tile kt+2
  // when ld.  With st, tile kt+1 goes to buffer B^1 and its first fragments
  // are read after the barrier.
  auto step = [&](int kt, int B, const lt_ig_u32x4* RNp, const lt_ig_u32x4* RNq,
                  lt_ig_u32x4* RFp, lt_ig_u32x4* RFq, bool ld, bool st) {
    if (ld) load(kt + 2, RFp, RFq);
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
      const int s = kk & 1;
      if (kk + 1 < KK) frags(B, kk + 1, s ^ 1);
      if (kk == KK - 1 && st) store(B ^ 1, RNp, RNq);
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = lt_ig_mfma(fa[s][a], fb[s][b], acc[a][b]);
    }
    if (ld && st) {
      constexpr int NM = TM * TN, NR = TM + TN, NL = VP + VQ;
#pragma unroll
      for (int kk = 0; kk < KK - 1; ++kk)
#pragma unroll
        for (int i = 0; i < NM; ++i) {
          __builtin_amdgcn_sched_group_barrier(0x0008, 1, 0);                 // MFMA
          if (i < NR) __builtin_amdgcn_sched_group_barrier(0x0100, 1, 0);    // LDS read
          if (kk == 0 && i < NL) __builtin_amdgcn_sched_group_barrier(0x0020, 1, 0);  // load
        }
#pragma unroll
      for (int i = 0; i < NM; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x0008, 1, 0);                   // MFMA
        if (i < NL) __builtin_amdgcn_sched_group_barrier(0x0200, 1, 0);      // LDS write
      }
    }
    if (st) {
      __syncthreads();
      frags(B ^ 1, 0, 0);
    }
  };
  load(0, r0p, r0q);
  if (KT > 1) load(1, r1p, r1q);
  store(0, r0p, r0q);
  __syncthreads();
  frags(0, 0, 0);
  // tile t lives in register set t % 2 and LDS buffer t % 2
  int kt = 0;
  for (; kt + 3 < KT; kt += 2) {
    step(kt, 0, r1p, r1q, r0p, r0q, true, true);
    step(kt + 1, 1, r0p, r0q, r1p, r1q, true, true);
  }
  const int R = KT - kt;  // 1, 2 or 3 tiles left; kt is even
  if (R == 3) {
    step(kt, 0, r1p, r1q, r0p, r0q, true, true);
    step(kt + 1, 1, r0p, r0q, r1p, r1q, false, true);
    step(kt + 2, 0, r1p, r1q, r0p, r0q, false, false);
  } else if (R == 2) {
    step(kt, 0, r1p, r1q, r0p, r0q, false, true);
    step(kt + 1, 1, r0p, r0q, r1p, r1q, false, false);
  } else {
    step(kt, 0, r1p, r1q, r0p, r0q, false, false);
  }
  lt_igemm_store<TM, TN, FULL>(g, acc, out, i0 + wm * WM, j0 + wn * WN, lane);
}

// Direct-to-LDS form of the pipelined tile: the same operands, contract and
// bounds handling as lt_igemm_tile_pipe (SWZ layout), but the global loads
// write LDS themselves (buffer_load_dwordx4 ... lds: lane l of a wave's load
// lands at the wave's LDS base + 16 l), so no registers hold tiles in flight
// and no LDS stores are issued.  Lane l of round v therefore fetches the
// 16-byte chunk that belongs at slot s = v * NT + tid of the XOR-permuted
// tile: row s / CPR, logical chunk (s % CPR) ^ ((row / RP) % CPR).
//  * Two LDS stages.  Tile kt + 1 is fetched into the other stage while tile
//    kt is multiplied: at the top of each K tile a wave waits for its own
//    loads (vmcnt 0), and one barrier makes everyone's visible and frees the
//    stage that the next loads overwrite.
//  * The freed registers pay for BK = 64: half the barriers of BK = 32 per K.
//  * sched_group_barrier interleaves, per k16 step, each MFMA with one
//    fragment read of the next step, and the next tile's loads with the
//    first step's MFMAs.
// LDS: 2 * (BM + BN) * BK 16-bit elements.  KTAIL as in lt_igemm_tile_pipe:
// chunks past KI in the last K tile read the buffer's out-of-range zero.
template <int BM, int BN, int BK>
__host__ __device__ constexpr int lt_igemm_dl_lds_bytes() {
  return 2 * (BM + BN) * BK * 2;
}

typedef __attribute__((address_space(3))) void lt_ig_lds_void;

// s_waitcnt operand that waits for vmcnt <= n (n < 64) and leaves lgkmcnt and
// expcnt alone (gfx9 encoding: vmcnt bits [3:0] and [15:14])
__host__ __device__ constexpr int lt_ig_vmcnt(int n) {
  return (n & 15) | (((n >> 4) & 3) << 14) | (0x7 << 4) | (0xF << 8);
}

// Workgroup barrier of the tiles whose LDS stages are filled by
// direct-to-LDS loads.  Before it, this wave's LDS reads have completed
// (lgkmcnt 0) and its loads up to WAIT (an lt_ig_vmcnt operand) have landed.
// The compiler fences on both sides keep LDS reads on their side: the
// s_barrier builtin alone is no memory barrier to the compiler, which may
// hoist the next K tile's fragment reads above it (reading a stage before
// another wave's load into it has landed) or leave reads in flight past it
// (while another wave's next load rewrites their stage); either gives wrong
// fragments when the loads are slow, e.g. with other kernels on the CU.
template <int WAIT>
__device__ __forceinline__ void lt_ig_lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(WAIT & ~(0xF << 8));  // and lgkmcnt(0)
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

template <typename T, int BM, int BN, int BK, bool FULL, int WGM, int WGN, bool KTAIL = false,
          class G>
__device__ __forceinline__ void lt_igemm_tile_dl(const G& g, const T* __restrict__ P,
                                                 const T* __restrict__ Q,
                                                 float* __restrict__ out, long i0, long j0,
                                                 unsigned char* smem, unsigned pbytes,
                                                 unsigned qbytes) {
  typedef typename lt_ig_vec8<T>::type V;
  constexpr int NT = 64 * WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / 32, TN = WN / 32;
  constexpr int CPR = BK / 8, KK = BK / 16, RP = 128 / BK;
  constexpr int VP = BM * CPR / NT, VQ = BN * CPR / NT;
  constexpr int SP = BM * BK, SQ = BN * BK;  // elements of one stage
  static_assert(TM >= 1 && TN >= 1 && WM % 32 == 0 && WN % 32 == 0, "wave tile");
  static_assert(VP >= 1 && VQ >= 1 && BM * CPR % NT == 0 && BN * CPR % NT == 0,
                "tile too small for the workgroup");
  static_assert(KK % 2 == 0, "BK must be a multiple of 32");
  const int KIT = KTAIL ? (g.KI + BK - 1) / BK : g.KI / BK;
  const int KT = g.KO * KIT;
  T* sP = (T*)smem;         // [2][BM * BK]
  T* sQ = sP + 2 * SP;      // [2][BN * BK]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const auto rsP = __builtin_amdgcn_make_buffer_rsrc((void*)P, (short)0, (int)pbytes, 0x00020000);
  const auto rsQ = __builtin_amdgcn_make_buffer_rsrc((void*)Q, (short)0, (int)qbytes, 0x00020000);
  // byte offset of this lane's chunk in each load round (0x80000000: a row
  // outside the problem, read as zeros) and the chunk's first k
  unsigned pv[VP], qv[VQ];
  int pk[VP], qk[VQ];
#pragma unroll
  for (int v = 0; v < VP; ++v) {
    const int s = v * NT + tid, r = s / CPR, c = (s % CPR) ^ ((r / RP) % CPR);
    bool ok;
    const long o = g.p_row(i0 + r, ok);
    pv[v] = (FULL || ok) ? (unsigned)((o + c * 8) * 2) : 0x80000000u;
    pk[v] = c * 8;
  }
#pragma unroll
  for (int v = 0; v < VQ; ++v) {
    const int s = v * NT + tid, r = s / CPR, c = (s % CPR) ^ ((r / RP) % CPR);
    bool ok;
    const long o = g.q_row(j0 + r, ok);
    qv[v] = (FULL || ok) ? (unsigned)((o + c * 8) * 2) : 0x80000000u;
    qk[v] = c * 8;
  }
  auto load = [&](int kt, int buf) {
    const int ko = kt / KIT, kb = (kt % KIT) * BK;
    const int po = (int)((g.p_koff(ko) + kb) * 2), qo = (int)((g.q_koff(ko) + kb) * 2);
#pragma unroll
    for (int v = 0; v < VP; ++v)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rsP, (lt_ig_lds_void*)(sP + buf * SP + (v * NT + wave * 64) * 8), 16,
          (int)(KTAIL && kb + pk[v] >= g.KI ? 0x80000000u : pv[v]), po, 0, 0);
#pragma unroll
    for (int v = 0; v < VQ; ++v)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rsQ, (lt_ig_lds_void*)(sQ + buf * SQ + (v * NT + wave * 64) * 8), 16,
          (int)(KTAIL && kb + qk[v] >= g.KI ? 0x80000000u : qv[v]), qo, 0, 0);
  };
  const int fr = lane & 31, fh = lane >> 5;
  V fa[2][TM], fb[2][TN];
  auto frags = [&](int buf, int kk, int s) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
      fa[s][a] = *(const V*)(sP + buf * SP +
                             lt_ig_lds_off<BK, true>(wm * WM + a * 32 + fr, 2 * kk + fh));
#pragma unroll
    for (int b = 0; b < TN; ++b)
      fb[s][b] = *(const V*)(sQ + buf * SQ +
                             lt_ig_lds_off<BK, true>(wn * WN + b * 32 + fr, 2 * kk + fh));
  };
  lt_ig_f32x16 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = (lt_ig_f32x16){};
  auto step = [&](int kt, int B) {
    // this wave's loads of tile kt have landed and its reads of stage B^1
    // are done; then everyone's
    lt_ig_lds_barrier<lt_ig_vmcnt(0)>();
    const bool more = kt + 1 < KT;
    frags(B, 0, 0);
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
      const int s = kk & 1;
      if (kk == 0 && more) load(kt + 1, B ^ 1);
      if (kk + 1 < KK) frags(B, kk + 1, s ^ 1);
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = lt_ig_mfma(fa[s][a], fb[s][b], acc[a][b]);
      constexpr int NM = TM * TN, NR = TM + TN, NL = VP + VQ;
#pragma unroll
      for (int i = 0; i < NM; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x0008, 1, 0);  // MFMA
        if (kk + 1 < KK && i < NR) __builtin_amdgcn_sched_group_barrier(0x0100, 1, 0);  // LDS read
        if (kk == 0 && i < NL) __builtin_amdgcn_sched_group_barrier(0x0020, 1, 0);  // load
      }
    }
  };
  load(0, 0);
  int kt = 0;
#pragma unroll 1  // a K loop unrolled whole would hold every K tile's addresses
  for (; kt + 1 < KT; kt += 2) {
    step(kt, 0);
    step(kt + 1, 1);
  }
  if (kt < KT) step(kt, 0);
  lt_igemm_store<TM, TN, FULL>(g, acc, out, i0 + wm * WM, j0 + wn * WN, lane);
}

// Halo form of the pipelined tile, for a k-blocked staged Q whose outer-K
// positions within a K block ("taps", g.NTAP of them) only shift the row
// offset: Q(j, kb, tap, kr) = q_row(j) + q_koff(kb * NTAP + tap) + kr with
// q_row non-decreasing in j and q_koff(kb * NTAP + tap) - q_koff(kb * NTAP)
// a non-negative multiple of BK (a convolution's window taps over a
// position-major staged input X'[C/BK][N][H+2][W+2][BK]).  K tiles run block
// by block, taps inner.  The Q rows all taps of one block need are the HL
// consecutive BK-vectors starting at the tile's first row: they are loaded
// once per block into an LDS halo buffer (HB = 2: double-buffered, the next
// block's halo loaded during the first tap and stored to LDS after the last;
// HB = 1: one buffer, rewritten between two barriers after the last tap, for
// half the LDS and two workgroups per CU) and
// every tap reads its fragments from it at a uniform shift, instead of
// reloading the tile's rows from L2 for every tap.  P goes through the same
// two-tiles-ahead register pipeline and double-buffered LDS as in
// lt_igemm_tile_pipe.  HL (a host-computed bound over all tiles) must cover
// (q_row(j_last) - q_row(j0) + max tap shift) / BK + 1.
//
// QF: the halo is read from Q's float32 source instead of a staged copy.
// Staged row r of a block (the position of one BK-vector) comes from
// g.q_src(r, ok) -- the source offset of its first k_in, ok false on a zero
// halo row -- and the BK values of block kb at k_in = kb * BK + kr, kr < BK,
// G::QCS elements apart (zero past G::QKIN).  The halo is cut into wave
// units of 64 consecutive rows by one 8-value part (unit u = wave + NW * h,
// part u % CPR): a lane owns one row of each of its units, so every dword
// load of a wave reads 64 consecutive source positions of one k_in.  The next
// block's values come one k_in per step (K tile), as dword loads that write a
// wave-private LDS landing slot (two slots, alternating): value c of block b
// at step NTAP * (b - 1) - 1 + c, from the last tap of block b - 2 on.  Two
// steps after its load the lane reads its value back, rounds it to T (as the
// staging launch does) and shifts it into its 16-byte chunk, so no more than
// the chunk itself stays in registers; the chunk is complete at block b - 1's
// last tap, where it is stored.  The first block is loaded into registers
// before the K loop.  The same operand
// values in the same order as the staged form; needs NTAP >= 9.  Qf and
// qfbytes describe the source (under 2^31 bytes); Q and qbytes are unused.
// LDS beyond the staged form: 2 * NW * HQ * 64 floats (lt_igemm_halo_qf_lds).
// LDS bytes of the QF landing slots of lt_igemm_tile_halo
template <int BK, int HL, int NT>
__host__ __device__ constexpr int lt_igemm_halo_qf_lds() {
  return 2 * (NT / 64) * (((HL + 63) / 64 * (BK / 8) + NT / 64 - 1) / (NT / 64)) * 64 * 4;
}

template <typename T, int BM, int BN, int BK, bool FULL, int WGM, int WGN, int HL, int HB,
          class G, bool QF = false>
__device__ __forceinline__ void lt_igemm_tile_halo(const G& g, const T* __restrict__ P,
                                                   const T* __restrict__ Q,
                                                   float* __restrict__ out, long i0, long j0,
                                                   unsigned char* smem, unsigned pbytes,
                                                   unsigned qbytes,
                                                   const float* __restrict__ Qf = nullptr,
                                                   unsigned qfbytes = 0) {
  typedef typename lt_ig_vec8<T>::type V;
  constexpr int NT = 64 * WGM * WGN, LDR = BK + 8;
  constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / 32, TN = WN / 32;
  constexpr int CPR = BK / 8, KK = BK / 16;
  constexpr int VP = BM * CPR / NT;
  constexpr int HC = HL * CPR, HV = (HC + NT - 1) / NT;  // halo 16-byte chunks
  // QF: wave units of 64 halo rows x one part, HQ per wave
  constexpr int NW = NT / 64, HU = (HL + 63) / 64 * CPR, HQ = (HU + NW - 1) / NW;
  constexpr int NHV = QF ? HQ : HV;
  static_assert(TM >= 1 && TN >= 1 && WM % 32 == 0 && WN % 32 == 0, "wave tile");
  static_assert(VP >= 1 && BM * CPR % NT == 0, "tile too small for the workgroup");
  static_assert(KK % 2 == 0, "BK must be a multiple of 32");
  constexpr int NTAP = G::NTAP;
  static_assert(NTAP >= 2, "the halo form needs at least two taps per K block");
  static_assert(!QF || NTAP >= 9, "the float-source halo loads one value per tap over 8 taps");
  const int KT = g.KO, KB = KT / NTAP;
  T* sP = (T*)smem;              // [2][BM][LDR]
  T* sH = sP + 2 * BM * LDR;     // [HB][HL][LDR]
  float* sF = (float*)(sH + HB * HL * LDR);  // QF landing slots [2][NW][HQ][64]
  static_assert(HB == 1 || HB == 2, "one or two halo buffers");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int ck = tid % CPR, cc = ck * 8;
  const auto rsP = __builtin_amdgcn_make_buffer_rsrc((void*)P, (short)0, (int)pbytes, 0x00020000);
  const auto rsQ = __builtin_amdgcn_make_buffer_rsrc((void*)Q, (short)0, (int)qbytes, 0x00020000);
  unsigned pv[VP];
  int sp[VP];
#pragma unroll
  for (int v = 0; v < VP; ++v) {
    bool ok;
    const int r = (tid + v * NT) / CPR;
    const long o = g.p_row(i0 + r, ok);
    pv[v] = (FULL || ok) ? (unsigned)((o + cc) * 2) : 0x80000000u;
    sp[v] = r * LDR + cc;
  }
  // the tile's first Q row and each fragment row's vector offset from it
  // (rows past the problem reuse the last row: their results are not stored)
  bool ok0;
  const long qb0 = g.q_row(j0, ok0);
  const int fr = lane & 31, fh = lane >> 5;
  int lq[TN];
#pragma unroll
  for (int b = 0; b < TN; ++b) {
    long j = j0 + wn * WN + b * 32 + fr;
    if (!FULL && j >= g.ND) j = g.ND - 1;
    bool ok;
    lq[b] = (int)((g.q_row(j, ok) - qb0) / BK) * LDR + fh * 8;
  }
  // halo chunks of this lane.  Staged: chunk e = tid + h * NT is vector
  // e / CPR, part e % CPR.  QF: unit u = wave + NW * h, row (u / CPR) * 64 +
  // lane, part u % CPR; hv is the source offset (bytes) of the row's first
  // value of the part in block 0.
  unsigned hv[NHV];
#pragma unroll
  for (int h = 0; h < NHV; ++h) {
    if constexpr (QF) {
      const int u = wave + NW * h, row = (u / CPR) * 64 + lane;
      bool ok = false;
      const long src = u < HU && row < HL ? g.q_src(qb0 / BK + row, ok) : 0;
      hv[h] = ok ? (unsigned)((src + (long)((u % CPR) * 8) * G::QCS) * 4) : 0x80000000u;
    } else {
      const int e = tid + h * NT;
      hv[h] = e < HC ? (unsigned)((qb0 + (long)(e / CPR) * BK + (e % CPR) * 8) * 2)
                     : 0x80000000u;
    }
  }
  lt_ig_u32x4 hr[NHV];  // the next block's chunks
  const auto rsQf = __builtin_amdgcn_make_buffer_rsrc((void*)Qf, (short)0, (int)qfbytes,
                                                      0x00020000);
  // the block offset goes into the per-lane offset, which the buffer range
  // check covers: a halo tail past the end of Q (or a zero halo row of the
  // source, offset 2^31) reads zeros
  auto qf_off = [&](int h, int kb, int v) -> unsigned {
    if constexpr (QF) {
      const int c = kb * BK + ((wave + NW * h) % CPR) * 8 + v;
      return (G::QKIN % BK == 0 || c < G::QKIN)
                 ? hv[h] + (unsigned)((long)(kb * BK + v) * G::QCS * 4)
                 : 0x80000000u;
    } else {
      return 0x80000000u;
    }
  };
  auto hload = [&](int kb) {
    if constexpr (QF) {
      // the whole block into registers (the first block, before the K loop)
#pragma unroll
      for (int h = 0; h < NHV; ++h) {
        V x;
#pragma unroll
        for (int v = 0; v < 8; ++v)
          x[v] = (T)__builtin_bit_cast(
              float, __builtin_amdgcn_raw_buffer_load_b32(rsQf, (int)qf_off(h, kb, v), 0, 0));
        hr[h] = __builtin_bit_cast(lt_ig_u32x4, x);
      }
    } else {
      const unsigned bo = (unsigned)(g.q_koff(kb * NTAP) * 2);
#pragma unroll
      for (int h = 0; h < NHV; ++h)
        hr[h] = __builtin_amdgcn_raw_buffer_load_b128(rsQ, (int)(hv[h] + bo), 0, 0);
    }
  };
  // QF: value v of every unit of block kb into landing slot v & 1 (each
  // lane's dword lands at the slot's base + 4 * lane)
  const int wave_s = __builtin_amdgcn_readfirstlane(wave);  // the slot base is wave-uniform
  auto hload_tap = [&](int kb, int v, int slot) {
#pragma unroll
    for (int h = 0; h < NHV; ++h)
      if (HU % NW == 0 || wave_s + NW * h < HU)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rsQf, (lt_ig_lds_void*)(sF + ((slot * NW + wave_s) * NHV + h) * 64), 4,
            (int)qf_off(h, kb, v), 0, 0, 0);
  };
  // QF: the value (0..7) of the next blocks' stream loaded at step kt, -1
  // when none is: step kt loads value (kt + 1) % NTAP of block
  // (kt + 1) / NTAP + 1, when that value is below 8 and the block exists
  auto qf_val = [&](int kt) {
    const int c = (kt + 1) % NTAP, b = (kt + 1) / NTAP + 1;
    return kt >= -1 && c < 8 && b < KB ? c : -1;
  };
  // QF: the value of slot (v & 1), rounded to T, shifted in at the top of the
  // chunk (after the 8 values 0..7 of a block, value v is element v)
  auto hconv = [&](int slot) {
    float x[NHV];
#pragma unroll
    for (int h = 0; h < NHV; ++h)  // all reads first, one wait
      x[h] = HU % NW == 0 || wave + NW * h < HU
                 ? sF[((slot * NW + wave) * NHV + h) * 64 + lane]
                 : 0.0f;
#pragma unroll
    for (int h = 0; h < NHV; ++h) {
      const unsigned b = (unsigned)__builtin_bit_cast(unsigned short, (T)x[h]);
      hr[h][0] = (hr[h][0] >> 16) | (hr[h][1] << 16);
      hr[h][1] = (hr[h][1] >> 16) | (hr[h][2] << 16);
      hr[h][2] = (hr[h][2] >> 16) | (hr[h][3] << 16);
      hr[h][3] = (hr[h][3] >> 16) | (b << 16);
    }
    // the chunks as registers here: the compiler may not keep the 8 float
    // values of a chunk live instead and convert them at the store
#pragma unroll
    for (int h = 0; h < NHV; ++h) asm volatile("" : "+v"(hr[h])::"memory");
  };
  auto hstore = [&](int buf) {
#pragma unroll
    for (int h = 0; h < NHV; ++h) {
      if constexpr (QF) {
        const int u = wave + NW * h, row = (u / CPR) * 64 + lane;
        if ((HU % NW == 0 || u < HU) && row < HL)
          *(lt_ig_u32x4*)(sH + buf * HL * LDR + row * LDR + (u % CPR) * 8) = hr[h];
      } else {
        const int e = tid + h * NT;
        if (HC % NT == 0 || e < HC)
          *(lt_ig_u32x4*)(sH + buf * HL * LDR + (e / CPR) * LDR + (e % CPR) * 8) = hr[h];
      }
    }
  };
  lt_ig_u32x4 r0p[VP], r1p[VP];
  auto load = [&](int kt, lt_ig_u32x4* rp) {
    const int po = (int)(g.p_koff(kt) * 2);
#pragma unroll
    for (int v = 0; v < VP; ++v)
      rp[v] = __builtin_amdgcn_raw_buffer_load_b128(rsP, (int)pv[v], po, 0);
  };
  auto store = [&](int buf, const lt_ig_u32x4* rp) {
#pragma unroll
    for (int v = 0; v < VP; ++v) *(lt_ig_u32x4*)(sP + buf * BM * LDR + sp[v]) = rp[v];
  };
  // LDS element offset of tile kt's tap shift within its block's halo buffer
  auto hoff = [&](int kt) {
    const int kb = kt / NTAP;
    return (HB == 2 ? (kb & 1) * HL * LDR : 0) +
           (int)((g.q_koff(kt) - g.q_koff(kb * NTAP)) / BK) * LDR;
  };
  V fa[2][TM], fb[2][TN];
  auto frags = [&](int buf, int ho, int kk, int s) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
      fa[s][a] = *(const V*)(sP + buf * BM * LDR + (wm * WM + a * 32 + fr) * LDR + kk * 16 +
                             fh * 8);
#pragma unroll
    for (int b = 0; b < TN; ++b) fb[s][b] = *(const V*)(sH + ho + lq[b] + kk * 16);
  };
  lt_ig_f32x16 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = (lt_ig_f32x16){};
  // The workgroup barrier between K tiles.  QF: this wave's LDS accesses
  // complete (lgkmcnt 0), then a bare s_barrier.  __syncthreads would also
  // wait for every load in flight (vmcnt 0), the direct-to-LDS halo loads
  // included, which land in wave-private slots that no other wave reads.
  auto bar = [&]() {
    if constexpr (QF) {
      asm volatile("" ::: "memory");
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
      __builtin_amdgcn_s_barrier();
      asm volatile("" ::: "memory");
    } else {
      __syncthreads();
    }
  };
  // One K tile kt: its P tile is in LDS buffer B, its block's halo in halo
  // buffer (kt / NTAP) % 2, its k16-step-0 fragments in fa/fb[0]; RN holds P
  // of tile kt+1, RF receives P of tile kt+2 when ld.  With st, tile kt+1's P
  // goes to buffer B^1 and its first fragments are read after the barrier.
  auto step = [&](int kt, int B, const lt_ig_u32x4* RN, lt_ig_u32x4* RF, bool ld, bool st) {
    const int kb = kt / NTAP, tap = kt - kb * NTAP;
    const int ho = hoff(kt);
    const bool next_halo = kb + 1 < KB;
    if constexpr (QF) {
      // the float halo stream: value c of block b's chunks is loaded at step
      // NTAP * (b - 1) - 1 + c (c < 8: from the last tap of block b - 2 on)
      // into slot (step & 1), and read back two steps later
      if (kt >= 1 && qf_val(kt - 2) >= 0) {
        // wait until this wave's halo loads of step kt - 2 have landed: the
        // loads issued after them (step kt - 1: P, then halo) may stay out
        const bool h1 = qf_val(kt - 1) >= 0, p1 = kt + 1 < KT;
        if (h1 && p1) __builtin_amdgcn_s_waitcnt(lt_ig_vmcnt(NHV + VP));
        else if (h1) __builtin_amdgcn_s_waitcnt(lt_ig_vmcnt(NHV));
        else if (p1) __builtin_amdgcn_s_waitcnt(lt_ig_vmcnt(VP));
        else __builtin_amdgcn_s_waitcnt(lt_ig_vmcnt(0));
        asm volatile("" ::: "memory");
        hconv(kt & 1);
      }
      if (ld) load(kt + 2, RF);
      // (the slot was read back above: hconv's results are pinned first)
      if (qf_val(kt) >= 0) hload_tap((kt + 1) / NTAP + 1, qf_val(kt), kt & 1);
    } else {
      if (ld) load(kt + 2, RF);
      if (tap == 0 && next_halo) hload(kb + 1);
    }
    // raised wave priority while this wave issues its MFMAs: the sequencer
    // favours it over waves that are loading or storing (measured 3-4 %
    // faster on the 3x3 conv tile)
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
      const int s = kk & 1;
      if (kk + 1 < KK) frags(B, ho, kk + 1, s ^ 1);
      if (kk == KK - 1 && st) store(B ^ 1, RN);
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = lt_ig_mfma(fa[s][a], fb[s][b], acc[a][b]);
    }
    __builtin_amdgcn_s_setprio(0);
    if (HB == 2 && tap == NTAP - 1 && next_halo) hstore((kb + 1) & 1);
    if (st) {
      bar();
      if (HB == 1 && tap == NTAP - 1 && next_halo) {
        // one halo buffer: every wave is past block kb's last reads
        hstore(0);
        bar();
      }
      frags(B ^ 1, hoff(kt + 1), 0, 0);
    }
  };
  hload(0);
  load(0, r0p);
  if (KT > 1) load(1, r1p);
  if constexpr (QF)
    if (qf_val(-1) >= 0) hload_tap(1, 0, 1);  // the stream's first value (step -1)
  hstore(0);
  store(0, r0p);
  bar();
  frags(0, hoff(0), 0, 0);
  // tile t's P lives in register set t % 2 and LDS buffer t % 2
  int kt = 0;
  for (; kt + 3 < KT; kt += 2) {
    step(kt, 0, r1p, r0p, true, true);
    step(kt + 1, 1, r0p, r1p, true, true);
  }
  const int R = KT - kt;  // 1, 2 or 3 tiles left; kt is even
  if (R == 3) {
    step(kt, 0, r1p, r0p, true, true);
    step(kt + 1, 1, r0p, r1p, false, true);
    step(kt + 2, 0, r1p, r0p, false, false);
  } else if (R == 2) {
    step(kt, 0, r1p, r0p, false, true);
    step(kt + 1, 1, r0p, r1p, false, false);
  } else {
    step(kt, 0, r1p, r0p, false, false);
  }
  lt_igemm_store<TM, TN, FULL>(g, acc, out, i0 + wm * WM, j0 + wn * WN, lane);
}

// Direct-to-LDS form of the halo tile (same operands, geometry and HL bound
// as lt_igemm_tile_halo): P tiles go through two XOR-permuted LDS stages as in
// lt_igemm_tile_dl, and the halo of each K block through two XOR-permuted
// halo buffers, both filled by buffer loads that write LDS.  Halo rows are
// rounded up to whole load rounds (HV rounds of NT chunks; the extra chunks
// read the out-of-range zero), so every wave issues the same loads and the
// waits can count them.  Order of loads: P of tile kt + 1 at the top of tile
// kt, then, at a block's first tap, the next block's halo; the top of the
// second tap waits for everything but that halo (vmcnt HV), the other taps
// for everything.  A tap's Q fragments are halo rows shifted by the tap's
// row offset; their chunk permutation follows the shifted row.  A halo of
// at least NT chunks is loaded in HV rounds of NT chunks whose last round
// ends at the halo's last chunk (it reloads a few chunks of the round before:
// two waves write the same bytes, before the barrier that precedes any read),
// so the buffers hold exactly HL rows; a smaller halo takes one round, its
// chunks past HL reading the out-of-range zero into padding rows.
// LDS: 2 * (BM + HLP) * BK 16-bit elements, HLP = max(HL, NT / CPR).
template <int BM, int BK, int HL, int NT>
__host__ __device__ constexpr int lt_igemm_halo_dl_lds_bytes() {
  return 2 * (BM + (HL * (BK / 8) >= NT ? HL : NT / (BK / 8))) * BK * 2;
}

template <typename T, int BM, int BN, int BK, bool FULL, int WGM, int WGN, int HL, class G>
__device__ __forceinline__ void lt_igemm_tile_halo_dl(const G& g, const T* __restrict__ P,
                                                      const T* __restrict__ Q,
                                                      float* __restrict__ out, long i0, long j0,
                                                      unsigned char* smem, unsigned pbytes,
                                                      unsigned qbytes) {
  typedef typename lt_ig_vec8<T>::type V;
  constexpr int NT = 64 * WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / 32, TN = WN / 32;
  constexpr int CPR = BK / 8, KK = BK / 16, RP = 128 / BK;
  constexpr int VP = BM * CPR / NT;
  constexpr int HC = HL * CPR, HV = (HC + NT - 1) / NT, HLP = HC >= NT ? HL : NT / CPR;
  constexpr int SP = BM * BK, SH = HLP * BK;  // elements of a P stage, a halo buffer
  static_assert(TM >= 1 && TN >= 1 && WM % 32 == 0 && WN % 32 == 0, "wave tile");
  static_assert(VP >= 1 && BM * CPR % NT == 0, "tile too small for the workgroup");
  static_assert(KK % 2 == 0, "BK must be a multiple of 32");
  static_assert(HV <= 15, "halo loads beyond the vmcnt range");
  constexpr int NTAP = G::NTAP;
  static_assert(NTAP >= 2, "the halo form needs at least two taps per K block");
  const int KT = g.KO, KB = KT / NTAP;
  T* sP = (T*)smem;     // [2][BM * BK]
  T* sH = sP + 2 * SP;  // [2][HLP * BK]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const auto rsP = __builtin_amdgcn_make_buffer_rsrc((void*)P, (short)0, (int)pbytes, 0x00020000);
  const auto rsQ = __builtin_amdgcn_make_buffer_rsrc((void*)Q, (short)0, (int)qbytes, 0x00020000);
  unsigned pv[VP];
#pragma unroll
  for (int v = 0; v < VP; ++v) {
    const int s = v * NT + tid, r = s / CPR, c = (s % CPR) ^ ((r / RP) % CPR);
    bool ok;
    const long o = g.p_row(i0 + r, ok);
    pv[v] = (FULL || ok) ? (unsigned)((o + c * 8) * 2) : 0x80000000u;
  }
  // the tile's first Q row; each fragment row's halo row
  bool ok0;
  const long qb0 = g.q_row(j0, ok0);
  const int fr = lane & 31, fh = lane >> 5;
  int lq[TN];
#pragma unroll
  for (int b = 0; b < TN; ++b) {
    long j = j0 + wn * WN + b * 32 + fr;
    if (!FULL && j >= g.ND) j = g.ND - 1;  // rows past the problem: results not stored
    bool ok;
    lq[b] = (int)((g.q_row(j, ok) - qb0) / BK);
  }
  // halo chunks of this lane: slot s = hbase(h) + tid is halo row s / CPR
  auto hbase = [](int h) { return HC >= NT ? (h * NT < HC - NT ? h * NT : HC - NT) : 0; };
  unsigned hv[HV];
#pragma unroll
  for (int h = 0; h < HV; ++h) {
    const int s = hbase(h) + tid, r = s / CPR, c = (s % CPR) ^ ((r / RP) % CPR);
    hv[h] = s < HC ? (unsigned)((qb0 + (long)r * BK + c * 8) * 2) : 0x80000000u;
  }
  // the block offset goes into the per-lane offset, which the buffer range
  // check covers: a halo tail past the end of Q reads zeros
  auto hload = [&](int kb, int buf) {
    const unsigned bo = (unsigned)(g.q_koff(kb * NTAP) * 2);
#pragma unroll
    for (int h = 0; h < HV; ++h)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rsQ, (lt_ig_lds_void*)(sH + buf * SH + (hbase(h) + wave * 64) * 8), 16,
          (int)(hv[h] + bo), 0, 0, 0);
  };
  auto load = [&](int kt, int buf) {
    const int po = (int)(g.p_koff(kt) * 2);
#pragma unroll
    for (int v = 0; v < VP; ++v)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rsP, (lt_ig_lds_void*)(sP + buf * SP + (v * NT + wave * 64) * 8), 16, (int)pv[v], po,
          0, 0);
  };
  // halo row shift of tile kt's tap
  auto shift = [&](int kt) {
    const int kb = kt / NTAP;
    return (int)((g.q_koff(kt) - g.q_koff(kb * NTAP)) / BK);
  };
  V fa[2][TM], fb[2][TN];
  auto frags = [&](int B, int hb, int sh, int kk, int s) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
      fa[s][a] = *(const V*)(sP + B * SP +
                             lt_ig_lds_off<BK, true>(wm * WM + a * 32 + fr, 2 * kk + fh));
#pragma unroll
    for (int b = 0; b < TN; ++b)
      fb[s][b] = *(const V*)(sH + hb * SH + lt_ig_lds_off<BK, true>(lq[b] + sh, 2 * kk + fh));
  };
  lt_ig_f32x16 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = (lt_ig_f32x16){};
  constexpr int kWaitVm0 = (0xF << 8) | (0x7 << 4);  // vmcnt(0) (gfx9 s_waitcnt encoding)
  constexpr int kWaitVmHV = kWaitVm0 | (HV & 15);    // vmcnt(HV)
  auto step = [&](int kt, int B) {
    const int kb = kt / NTAP, tap = kt - kb * NTAP;
    const bool next_halo = kb + 1 < KB;
    // tile kt's P and block kb's halo are in LDS, stage B^1 is free; the top
    // of a block's second tap leaves the next block's halo in flight
    if (tap == 1 && next_halo) lt_ig_lds_barrier<kWaitVmHV>();
    else lt_ig_lds_barrier<kWaitVm0>();
    const bool more = kt + 1 < KT;
    const int hb = kb & 1, sh = shift(kt);
    frags(B, hb, sh, 0, 0);
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
      const int s = kk & 1;
      if (kk == 0) {
        if (more) load(kt + 1, B ^ 1);
        if (tap == 0 && next_halo) hload(kb + 1, hb ^ 1);
      }
      if (kk + 1 < KK) frags(B, hb, sh, kk + 1, s ^ 1);
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = lt_ig_mfma(fa[s][a], fb[s][b], acc[a][b]);
      constexpr int NM = TM * TN, NR = TM + TN;
#pragma unroll
      for (int i = 0; i < NM; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x0008, 1, 0);  // MFMA
        if (kk + 1 < KK && i < NR) __builtin_amdgcn_sched_group_barrier(0x0100, 1, 0);  // LDS read
      }
    }
  };
  hload(0, 0);
  load(0, 0);
  int kt = 0;
#pragma unroll 1  // a K loop unrolled whole would hold every K tile's addresses
  for (; kt + 1 < KT; kt += 2) {
    step(kt, 0);
    step(kt + 1, 1);
  }
  if (kt < KT) step(kt, 0);
  lt_igemm_store<TM, TN, FULL>(g, acc, out, i0 + wm * WM, j0 + wn * WN, lane);
}

// FP32 form of the pipelined tile (v_mfma_f32_32x32x2_f32) for staged float
// operands, KI % BK == 0.  Same pipeline as lt_igemm_tile_pipe: global loads
// two K tiles ahead, double-buffered LDS rows padded by 16 bytes, one barrier
// per K tile.  Per k16 step each lane reads 8 consecutive k of its fragment
// row (two 16-byte LDS reads): lanes 0-31 hold k 0..7 and lanes 32-63 k 8..15
// of the step, so MFMA m (m = 0..7) adds the products of k = m and k = 8 + m.
// LDS: 2 * (BM + BN) rows of BK + 4 floats.  kt0, nkt: the K tiles to run
// (all by default); add: accumulate into `out` (see lt_igemm_store).
template <int BM, int BN, int BK, bool FULL, int WGM, int WGN, class G>
__device__ __forceinline__ void lt_igemm_tile_f32(const G& g, const float* __restrict__ P,
                                                  const float* __restrict__ Q,
                                                  float* __restrict__ out, long i0, long j0,
                                                  unsigned char* smem, unsigned pbytes,
                                                  unsigned qbytes, int kt0 = 0, int nkt = -1,
                                                  bool add = false) {
  constexpr int NT = 64 * WGM * WGN, LDR = BK + 4;
  constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / 32, TN = WN / 32;
  constexpr int CPR = BK / 4, KK = BK / 16;   // 16-byte chunks per row, k16 steps
  constexpr int VP = BM * CPR / NT, VQ = BN * CPR / NT;
  static_assert(TM >= 1 && TN >= 1 && WM % 32 == 0 && WN % 32 == 0, "wave tile");
  static_assert(VP >= 1 && VQ >= 1 && BM * CPR % NT == 0 && BN * CPR % NT == 0,
                "tile too small for the workgroup");
  static_assert(KK % 2 == 0, "BK must be a multiple of 32");
  const int KIT = g.KI / BK;
  const int KT = nkt < 0 ? g.KO * KIT : nkt;  // K tiles kt0 .. kt0 + KT - 1
  float* sP = (float*)smem;      // [2][BM][LDR]
  float* sQ = sP + 2 * BM * LDR; // [2][BN][LDR]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int cc = (tid % CPR) * 4;
  const auto rsP = __builtin_amdgcn_make_buffer_rsrc((void*)P, (short)0, (int)pbytes, 0x00020000);
  const auto rsQ = __builtin_amdgcn_make_buffer_rsrc((void*)Q, (short)0, (int)qbytes, 0x00020000);
  unsigned pv[VP], qv[VQ];
#pragma unroll
  for (int v = 0; v < VP; ++v) {
    bool ok;
    const long o = g.p_row(i0 + (tid + v * NT) / CPR, ok);
    pv[v] = (FULL || ok) ? (unsigned)((o + cc) * 4) : 0x80000000u;
  }
#pragma unroll
  for (int v = 0; v < VQ; ++v) {
    bool ok;
    const long o = g.q_row(j0 + (tid + v * NT) / CPR, ok);
    qv[v] = (FULL || ok) ? (unsigned)((o + cc) * 4) : 0x80000000u;
  }
  lt_ig_u32x4 r0p[VP], r0q[VQ], r1p[VP], r1q[VQ];
  auto load = [&](int kt, lt_ig_u32x4* rp, lt_ig_u32x4* rq) {
    const int ko = (kt0 + kt) / KIT, kb = ((kt0 + kt) % KIT) * BK;
    const int po = (int)((g.p_koff(ko) + kb) * 4), qo = (int)((g.q_koff(ko) + kb) * 4);
#pragma unroll
    for (int v = 0; v < VP; ++v)
      rp[v] = __builtin_amdgcn_raw_buffer_load_b128(rsP, (int)pv[v], po, 0);
#pragma unroll
    for (int v = 0; v < VQ; ++v)
      rq[v] = __builtin_amdgcn_raw_buffer_load_b128(rsQ, (int)qv[v], qo, 0);
  };
  auto store = [&](int buf, const lt_ig_u32x4* rp, const lt_ig_u32x4* rq) {
#pragma unroll
    for (int v = 0; v < VP; ++v)
      *(lt_ig_u32x4*)(sP + (buf * BM + (tid + v * NT) / CPR) * LDR + cc) = rp[v];
#pragma unroll
    for (int v = 0; v < VQ; ++v)
      *(lt_ig_u32x4*)(sQ + (buf * BN + (tid + v * NT) / CPR) * LDR + cc) = rq[v];
  };
  typedef float F4 __attribute__((ext_vector_type(4)));
  F4 fa[2][TM][2], fb[2][TN][2];
  const int fr = lane & 31, fh = lane >> 5;
  auto frags = [&](int buf, int kk, int s) {
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      const float* p = sP + (buf * BM + wm * WM + a * 32 + fr) * LDR + kk * 16 + fh * 8;
      fa[s][a][0] = *(const F4*)p;
      fa[s][a][1] = *(const F4*)(p + 4);
    }
#pragma unroll
    for (int b = 0; b < TN; ++b) {
      const float* p = sQ + (buf * BN + wn * WN + b * 32 + fr) * LDR + kk * 16 + fh * 8;
      fb[s][b][0] = *(const F4*)p;
      fb[s][b][1] = *(const F4*)(p + 4);
    }
  };
  lt_ig_f32x16 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = (lt_ig_f32x16){};
  auto step = [&](int kt, int B, const lt_ig_u32x4* RNp, const lt_ig_u32x4* RNq,
                  lt_ig_u32x4* RFp, lt_ig_u32x4* RFq, bool ld, bool st) {
    if (ld) load(kt + 2, RFp, RFq);
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
      const int s = kk & 1;
      if (kk + 1 < KK) frags(B, kk + 1, s ^ 1);
      if (kk == KK - 1 && st) store(B ^ 1, RNp, RNq);
#pragma unroll
      for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int b = 0; b < TN; ++b)
            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                fa[s][a][m >> 2][m & 3], fb[s][b][m >> 2][m & 3], acc[a][b], 0, 0, 0);
    }
    if (st) {
      __syncthreads();
      frags(B ^ 1, 0, 0);
    }
  };
  load(0, r0p, r0q);
  if (KT > 1) load(1, r1p, r1q);
  store(0, r0p, r0q);
  __syncthreads();
  frags(0, 0, 0);
  int kt = 0;
  for (; kt + 3 < KT; kt += 2) {
    step(kt, 0, r1p, r1q, r0p, r0q, true, true);
    step(kt + 1, 1, r0p, r0q, r1p, r1q, true, true);
  }
  const int R = KT - kt;
  if (R == 3) {
    step(kt, 0, r1p, r1q, r0p, r0q, true, true);
    step(kt + 1, 1, r0p, r0q, r1p, r1q, false, true);
    step(kt + 2, 0, r1p, r1q, r0p, r0q, false, false);
  } else if (R == 2) {
    step(kt, 0, r1p, r1q, r0p, r0q, false, true);
    step(kt + 1, 1, r0p, r0q, r1p, r1q, false, false);
  } else {
    step(kt, 0, r1p, r1q, r0p, r0q, false, false);
  }
  lt_igemm_store<TM, TN, FULL>(g, acc, out, i0 + wm * WM, j0 + wn * WN, lane, add);
}

#endif  // LT_IGEMM_TILE_H_
