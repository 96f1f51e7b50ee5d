// Kernel library for batches of run-time-sized matmul nests (gfx950).
//
//   lt_gemm_<t>_<BM>x<BN>x<BK>(A, Bt, C, M, N, K)   one GEMM per launch
//   lt_ggemm[_mid|_small]_<t>(items, queue)          persistent grouped launch
//   lt_ggemm_c<cfg>_<t>(items, lists)                the same, one configuration
//   lt_cvt_<t>(src, dst, rows, cols, transpose)      float32 -> 16-bit operands
//
// Every launch form runs the direct-to-LDS pipelined implicit-GEMM tile of
// igemm_tile.h (with its K-tail masking: K is any multiple of 8) on operands
// A[M][ld] and Bt[N][ld], ld = lt_mm_ld(K) (already K-innermost, so no
// staging).  The grouped
// kernels walk a host-built work list of output tiles from many GEMMs, cut
// into one segment per XCD and ordered by decreasing cost within each, so
// that the longest tiles start first, the workgroups finish together, and
// tiles that share operand panels are fetched into one L2.
//  * Mixed configurations (lt_ggemm[_mid|_small]): every item carries its
//    own tile configuration, chosen per shape bucket by the autotuner.  The
//    workgroups of XCD x take the items of segment x, first by index, then
//    from the segment's counter (fetched while they compute the previous
//    item).  The last workgroup to leave resets the counters for the next
//    launch, so graph replays need no memset; launches of one work list must
//    not overlap.
//  * One configuration (lt_ggemm_c<cfg>): each workgroup runs its own list
//    of the segment's items, built on the host for the launch's grid, and
//    overlaps each item's start with the end of the previous one
//    (lt_ggemm_static).
#include <hip/hip_runtime.h>

#include "igemm_tile.h"

// one output tile of one GEMM: the GEMM's operands and sizes, the tile's
// first row and column, its tile configuration
struct LtWorkItem {
  const void* A;
  const void* Bt;
  float* C;
  int M, N, K, m0, n0, cfg;
};

// Row pitch (elements) of the 16-bit operands A[M][pitch] and Bt[N][pitch]:
// K rounded up to 64, so that every row starts on a 128-byte line (with K a
// multiple of 16 only, a K tile's rows would straddle two lines each); the
// columns past K are never read
__host__ __device__ constexpr int lt_mm_ld(int K) { return (K + 63) / 64 * 64; }

// geometry of a plain matmul for the tile: rows of A and Bt are lt_mm_ld(K)
// apart
struct LtMmGeo {
  int KO, KI;
  long MD, ND, N, LD;
  __device__ __forceinline__ long p_row(long i, bool& ok) const {
    ok = i < MD;
    return i * LD;
  }
  __device__ __forceinline__ long q_row(long j, bool& ok) const {
    ok = j < ND;
    return j * LD;
  }
  __device__ __forceinline__ long p_koff(int) const { return 0; }
  __device__ __forceinline__ long q_koff(int) const { return 0; }
  __device__ __forceinline__ long out_row(long i) const { return i * N; }
  __device__ __forceinline__ long out_col(long j) const { return j; }
};

__device__ __forceinline__ LtMmGeo lt_mm_geo(int M, int N, int K) {
  LtMmGeo g;
  g.KO = 1;
  g.KI = K;
  g.MD = M;
  g.ND = N;
  g.N = N;
  g.LD = lt_mm_ld(K);
  return g;
}

// tile configurations: index -> (BM, BN, BK), in three register classes.  A
// grouped kernel holds the code of every configuration of its class, and its
// register count is that of the largest one, which sets how many waves a
// SIMD holds: small (up to 128x64 / 64x128: three), mid (adds 128x128x32:
// two), all (adds 128x128x64, 256x128 and 128x256: one).  The launch takes
// the smallest class that covers the batch; its LDS is dynamic, sized for
// the largest configuration present.  A batch whose items all use one
// configuration runs that configuration's own kernel (lt_ggemm_c<cfg>).
#define LT_CFG_SMALL(_) _(0, 64, 64, 32) _(1, 128, 64, 32) _(2, 64, 128, 32)
#define LT_CFG_MID(_) LT_CFG_SMALL(_) _(6, 128, 128, 32)
#define LT_CFG_LIST(_) LT_CFG_MID(_) _(3, 128, 128, 64) _(4, 256, 128, 32) _(5, 128, 256, 32)

// one GEMM per launch
#define LT_PLAIN(T, TN, BM, BN, BK)                                                          \
  extern "C" __global__ void __launch_bounds__(256)                                         \
      lt_gemm_##TN##_##BM##x##BN##x##BK(const T* __restrict__ A, const T* __restrict__ Bt,  \
                                        float* __restrict__ C, int M, int N, int K) {       \
    __shared__ __attribute__((aligned(16))) unsigned char smem[lt_igemm_dl_lds_bytes<BM, BN, BK>()]; \
    lt_igemm_tile_dl<T, BM, BN, BK, false, 2, 2, true>(                                     \
        lt_mm_geo(M, N, K), A, Bt, C, (long)blockIdx.y * BM, (long)blockIdx.x * BN, smem,  \
        (unsigned)M * lt_mm_ld(K) * 2, (unsigned)N * lt_mm_ld(K) * 2);                      \
  }

#define LT_CASE(i, BM, BN, BK)                                                               \
  case i:                                                                                    \
    lt_igemm_tile_dl<T, BM, BN, BK, false, 2, 2, true>(                                     \
        g, A, Bt, it.C, it.m0, it.n0, smem, (unsigned)it.M * lt_mm_ld(it.K) * 2,            \
        (unsigned)it.N * lt_mm_ld(it.K) * 2);                                               \
    break;

// Persistent loop of a batch whose items all use one tile configuration:
// workgroup w runs the items [wl[w], wl[w + 1]) of a host-built list that
// gives each workgroup its share of one XCD's items (longest first, each to
// the least-loaded workgroup of the XCD), in decreasing cost.  Knowing its
// next item in advance, a workgroup reads the item's record while the
// current one runs and issues the next item's first K tile into the free LDS
// stage during the current item's last K tile: an item starts on operands
// already in LDS, with no queue and no dependent record reads in front of
// it.
template <typename T, int BM, int BN, int BK, int WGM, int WGN>
__device__ __forceinline__ void lt_ggemm_static(const LtWorkItem* __restrict__ items,
                                                const int* __restrict__ wl,
                                                unsigned char* smem) {
  typedef typename lt_ig_vec8<T>::type V;
  constexpr int NT = 64 * WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / 32, TN = WN / 32;
  constexpr int CPR = BK / 8, KK = BK / 16, RP = 128 / BK;
  constexpr int VP = BM * CPR / NT, VQ = BN * CPR / NT;
  constexpr int SP = BM * BK, SQ = BN * BK;
  static_assert(VP >= 1 && VQ >= 1 && KK % 2 == 0, "tile shape");
  T* sP = (T*)smem;     // [2][BM * BK]
  T* sQ = sP + 2 * SP;  // [2][BN * BK]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wave_s = __builtin_amdgcn_readfirstlane(wave);
  const int wm = wave / WGN, wn = wave % WGN, fr = lane & 31, fh = lane >> 5;
  // this lane's chunk of each load round (row, XOR-permuted k), as in
  // lt_igemm_tile_dl: the same for every item
  int pr[VP], pk[VP], qr[VQ], qk[VQ];
#pragma unroll
  for (int v = 0; v < VP; ++v) {
    const int e = v * NT + tid, r = e / CPR;
    pr[v] = r;
    pk[v] = ((e % CPR) ^ ((r / RP) % CPR)) * 8;
  }
#pragma unroll
  for (int v = 0; v < VQ; ++v) {
    const int e = v * NT + tid, r = e / CPR;
    qr[v] = r;
    qk[v] = ((e % CPR) ^ ((r / RP) % CPR)) * 8;
  }
  // this workgroup's items
  int idx = wl[blockIdx.x];
  const int end = wl[blockIdx.x + 1];
  if (idx >= end) return;
  struct Job {
    __amdgpu_buffer_rsrc_t rP, rQ;
    unsigned pv[VP], qv[VQ];
    int K, KT;
  };
  auto job = [&](const LtWorkItem& it) {
    Job j;
    const unsigned ld = (unsigned)lt_mm_ld(it.K);
    j.rP = __builtin_amdgcn_make_buffer_rsrc((void*)it.A, (short)0,
                                             (int)((unsigned)it.M * ld * 2), 0x00020000);
    j.rQ = __builtin_amdgcn_make_buffer_rsrc((void*)it.Bt, (short)0,
                                             (int)((unsigned)it.N * ld * 2), 0x00020000);
#pragma unroll
    for (int v = 0; v < VP; ++v) {
      const int row = it.m0 + pr[v];
      j.pv[v] = row < it.M ? ((unsigned)row * ld + pk[v]) * 2 : 0x80000000u;
    }
#pragma unroll
    for (int v = 0; v < VQ; ++v) {
      const int row = it.n0 + qr[v];
      j.qv[v] = row < it.N ? ((unsigned)row * ld + qk[v]) * 2 : 0x80000000u;
    }
    j.K = it.K;
    j.KT = (it.K + BK - 1) / BK;
    return j;
  };
  // K tile kt of a job into LDS stage `stage` (chunks past K read zeros)
  auto load = [&](const Job& j, int kt, int stage) {
    const int kb = kt * BK;
#pragma unroll
    for (int v = 0; v < VP; ++v)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          j.rP, (lt_ig_lds_void*)(sP + stage * SP + (v * NT + wave_s * 64) * 8), 16,
          (int)(kb + pk[v] >= j.K ? 0x80000000u : j.pv[v]), kb * 2, 0, 0);
#pragma unroll
    for (int v = 0; v < VQ; ++v)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          j.rQ, (lt_ig_lds_void*)(sQ + stage * SQ + (v * NT + wave_s * 64) * 8), 16,
          (int)(kb + qk[v] >= j.K ? 0x80000000u : j.qv[v]), kb * 2, 0, 0);
  };
  V fa[2][TM], fb[2][TN];
  auto frags = [&](int buf, int kk, int sl) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
      fa[sl][a] = *(const V*)(sP + buf * SP +
                              lt_ig_lds_off<BK, true>(wm * WM + a * 32 + fr, 2 * kk + fh));
#pragma unroll
    for (int b = 0; b < TN; ++b)
      fb[sl][b] = *(const V*)(sQ + buf * SQ +
                              lt_ig_lds_off<BK, true>(wn * WN + b * 32 + fr, 2 * kk + fh));
  };
  LtWorkItem it = items[idx];
  Job cur = job(it);
  load(cur, 0, 0);
  int st = 0;  // the LDS stage of the current item's first K tile
  for (;;) {
    const int nidx = idx + 1;
    const bool more = nidx < end;
    LtWorkItem nit = it;
    Job nxt = cur;
    lt_ig_f32x16 acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b) acc[a][b] = (lt_ig_f32x16){};
#pragma unroll 1
    for (int kt = 0; kt < cur.KT; ++kt) {
      const int B = (st + kt) & 1;
      // this wave's loads of tile kt have landed (and the previous item's
      // stores) and its reads of the other stage are done; then everyone's
      lt_ig_lds_barrier<lt_ig_vmcnt(0)>();
      if (kt == 0 && more) nit = items[nidx];  // the next item's record
      const bool last = kt + 1 == cur.KT;
      if (last && more) nxt = job(nit);
      frags(B, 0, 0);
#pragma unroll
      for (int kk = 0; kk < KK; ++kk) {
        const int sl = kk & 1;
        if (kk == 0) {
          if (!last) load(cur, kt + 1, B ^ 1);
          else if (more) load(nxt, 0, B ^ 1);
        }
        if (kk + 1 < KK) frags(B, kk + 1, sl ^ 1);
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int b = 0; b < TN; ++b) acc[a][b] = lt_ig_mfma(fa[sl][a], fb[sl][b], acc[a][b]);
        constexpr int NM = TM * TN, NR = TM + TN, NL = VP + VQ;
#pragma unroll
        for (int i = 0; i < NM; ++i) {
          __builtin_amdgcn_sched_group_barrier(0x0008, 1, 0);  // MFMA
          if (kk + 1 < KK && i < NR) __builtin_amdgcn_sched_group_barrier(0x0100, 1, 0);
          if (kk == 0 && i < NL) __builtin_amdgcn_sched_group_barrier(0x0020, 1, 0);
        }
      }
    }
    lt_igemm_store<TM, TN, false>(lt_mm_geo(it.M, it.N, it.K), acc, it.C, it.m0 + wm * WM,
                                  it.n0 + wn * WN, lane);
    if (!more) break;
    st = (st + cur.KT) & 1;
    idx = nidx;
    it = nit;
    cur = nxt;
  }
}

// queue: [0..8] the 8 segment offsets of the work list, [9..16] one item
// counter per segment, [17] the count of workgroups done.  Workgroup w runs
// on XCD w % 8 and takes items of that XCD's segment: first by index, then
// from the segment's counter.  With fewer than 8 workgroups, workgroup w
// runs segments w, w + grid, ... alone, in order.
#define LT_GROUPED(T_, TN, CLASS, LIST)                                                      \
  extern "C" __global__ void __launch_bounds__(256)                                         \
      lt_ggemm##CLASS##_##TN(const LtWorkItem* __restrict__ items, int* __restrict__ queue) { \
    typedef T_ T;                                                                           \
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];                  \
    __shared__ int next_item;                                                               \
    const int G = (int)gridDim.x;                                                           \
    const bool shared_seg = G >= 8;                                                         \
    for (int x = (int)blockIdx.x % 8; x < 8; x += shared_seg ? 8 : G) {                     \
      const int beg = queue[x], end = queue[x + 1];                                         \
      /* workgroups on this segment, and this one's rank among them */                      \
      const int nx = shared_seg ? (G - x + 7) / 8 : 1;                                      \
      int i = beg + (shared_seg ? (int)blockIdx.x / 8 : 0);                                 \
      while (i < end) {                                                                     \
        /* take the next item now: the atomic's latency hides behind this one */            \
        int nxt = 0;                                                                        \
        if (threadIdx.x == 0)                                                               \
          nxt = shared_seg ? beg + nx + atomicAdd(queue + 9 + x, 1) : i + 1;                \
        const LtWorkItem it = items[i];                                                     \
        const T* A = (const T*)it.A;                                                        \
        const T* Bt = (const T*)it.Bt;                                                      \
        const LtMmGeo g = lt_mm_geo(it.M, it.N, it.K);                                      \
        switch (it.cfg) { LIST(LT_CASE) default: break; }                                   \
        __syncthreads(); /* LDS reuse; every lane has read next_item */                     \
        if (threadIdx.x == 0) next_item = nxt;                                              \
        __syncthreads();                                                                    \
        i = next_item;                                                                      \
      }                                                                                     \
    }                                                                                       \
    if (threadIdx.x == 0) {                                                                 \
      __threadfence(); /* this workgroup's last take precedes its exit count */             \
      if (atomicAdd(queue + 17, 1) == G - 1) {                                              \
        __threadfence(); /* every workgroup has taken its last item */                      \
        for (int x = 0; x < 8; ++x) atomicExch(queue + 9 + x, 0);                           \
        atomicExch(queue + 17, 0);                                                          \
      }                                                                                     \
    }                                                                                       \
  }

#define LT_PLAIN_ALL(T, TN)                                                                \
  LT_PLAIN(T, TN, 64, 64, 32)                                                              \
  LT_PLAIN(T, TN, 128, 64, 32)                                                             \
  LT_PLAIN(T, TN, 64, 128, 32)                                                             \
  LT_PLAIN(T, TN, 128, 128, 64)                                                            \
  LT_PLAIN(T, TN, 256, 128, 32)                                                            \
  LT_PLAIN(T, TN, 128, 256, 32)                                                            \
  LT_PLAIN(T, TN, 128, 128, 32)

LT_PLAIN_ALL(__bf16, bf16)
LT_PLAIN_ALL(_Float16, f16)

#define LT_WPE(n) __attribute__((amdgpu_waves_per_eu(n)))
// one configuration per kernel, for batches whose items all use it: the
// static item loop; its own register count alone sets the occupancy, and
// the waves-per-SIMD hint asks the compiler to keep it there
#define LT_STATIC(T, TN, c, BM, BN, BK, ...)                                                \
  extern "C" __global__ void __launch_bounds__(256) __VA_ARGS__                             \
      lt_ggemm_c##c##_##TN(const LtWorkItem* __restrict__ items, const int* __restrict__ wl) { \
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];                  \
    lt_ggemm_static<T, BM, BN, BK, 2, 2>(items, wl, smem);                                  \
  }
#define LT_GROUPED_ALL(T, TN)                                                               \
  LT_GROUPED(T, TN, , LT_CFG_LIST)                                                          \
  LT_GROUPED(T, TN, _mid, LT_CFG_MID)                                                       \
  LT_GROUPED(T, TN, _small, LT_CFG_SMALL)                                                   \
  LT_STATIC(T, TN, 0, 64, 64, 32, LT_WPE(8))                                                \
  LT_STATIC(T, TN, 1, 128, 64, 32, LT_WPE(5))                                               \
  LT_STATIC(T, TN, 2, 64, 128, 32, LT_WPE(5))                                               \
  LT_STATIC(T, TN, 3, 128, 128, 64, LT_WPE(3))                                              \
  LT_STATIC(T, TN, 4, 256, 128, 32)                                                         \
  LT_STATIC(T, TN, 5, 128, 256, 32)                                                         \
  LT_STATIC(T, TN, 6, 128, 128, 32, LT_WPE(3))

LT_GROUPED_ALL(__bf16, bf16)
LT_GROUPED_ALL(_Float16, f16)

// float32 [rows][cols] -> 16-bit [rows][lt_mm_ld(cols)], or
// [cols][lt_mm_ld(rows)] when transpose (the operand layout of the tiles
// above; the padding columns are not written)
#define LT_CVT(T, TN)                                                                       \
  extern "C" __global__ void __launch_bounds__(256)                                         \
      lt_cvt_##TN(const float* __restrict__ src, T* __restrict__ dst, int rows, int cols,   \
                  int transpose) {                                                          \
    const long i = (long)blockIdx.x * 256 + threadIdx.x;                                    \
    if (i >= (long)rows * cols) return;                                                     \
    const int r = (int)(i / cols), c = (int)(i % cols);                                     \
    const T v = (T)src[i];                                                                  \
    if (transpose)                                                                          \
      dst[(long)c * lt_mm_ld(rows) + r] = v;                                                \
    else                                                                                    \
      dst[(long)r * lt_mm_ld(cols) + c] = v;                                                \
  }

LT_CVT(__bf16, bf16)
LT_CVT(_Float16, f16)
