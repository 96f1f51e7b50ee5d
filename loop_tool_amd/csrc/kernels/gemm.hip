// Kernel library for batches of run-time-sized matmul nests (gfx950).
//
//   lt_gemm_<t>_<BM>x<BN>x<BK>(A, Bt, C, M, N, K)   one GEMM per launch
//   lt_ggemm[_mid|_small]_<t>(descs, items, queue)
//                                                    persistent grouped launch
//   lt_cvt_<t>(src, dst, rows, cols, transpose)      float32 -> 16-bit operands
//
// Both launch forms run the direct-to-LDS pipelined implicit-GEMM tile of
// igemm_tile.h (with its K-tail masking: K is any multiple of 8) on operands
// A[M][K] and Bt[N][K] (already K-innermost, so no staging).  The grouped
// kernel walks a host-built work list of output tiles from many GEMMs, cut
// into one segment per XCD and ordered by decreasing cost within each: the
// workgroups of XCD x take the items of segment x, first by index, then
// from the segment's counter (fetched while they compute the previous
// item), so the longest tiles start first, the workgroups finish together,
// and tiles that share operand panels are fetched into one L2.  The last
// workgroup to leave resets the counters for the next launch, so graph
// replays need no memset; launches of one work list must not overlap.
// Every item carries its own tile configuration, chosen per shape bucket by
// the autotuner, so one launch mixes tile shapes.
#include <hip/hip_runtime.h>

#include "igemm_tile.h"

struct LtGemmDesc {
  const void* A;
  const void* Bt;
  float* C;
  int M, N, K, pad;
};

struct LtWorkItem {
  int gemm, m0, n0, cfg;
};

// geometry of a plain matmul for the tile: rows of A and Bt are K apart
struct LtMmGeo {
  int KO, KI;
  long MD, ND, N;
  __device__ __forceinline__ long p_row(long i, bool& ok) const {
    ok = i < MD;
    return i * KI;
  }
  __device__ __forceinline__ long q_row(long j, bool& ok) const {
    ok = j < ND;
    return j * KI;
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
  return g;
}

// tile configurations: index -> (BM, BN, BK), in three register classes.  A
// grouped kernel holds the code of every configuration of its class, and its
// register count is that of the largest one, which sets how many waves a
// SIMD holds: small (up to 128x64 / 64x128: three), mid (adds 128x128x32:
// two), all (adds 128x128x64, 256x128 and 128x256: one).  The launch takes
// the smallest class that covers the batch; its LDS is dynamic, sized for
// the largest configuration present.
#define LT_CFG_SMALL(_) _(0, 64, 64, 32) _(1, 128, 64, 32) _(2, 64, 128, 32)
#define LT_CFG_MID(_) LT_CFG_SMALL(_) _(6, 128, 128, 32)
#define LT_CFG_LIST(_) LT_CFG_MID(_) _(3, 128, 128, 64) _(4, 256, 128, 32) _(5, 128, 256, 32)
// one configuration per kernel, for batches whose items all use it: its own
// register count alone sets the occupancy, and the waves-per-SIMD hint asks
// the compiler to keep it there
#define LT_CFG_ONE0(_) _(0, 64, 64, 32)
#define LT_CFG_ONE1(_) _(1, 128, 64, 32)
#define LT_CFG_ONE2(_) _(2, 64, 128, 32)
#define LT_CFG_ONE3(_) _(3, 128, 128, 64)
#define LT_CFG_ONE4(_) _(4, 256, 128, 32)
#define LT_CFG_ONE5(_) _(5, 128, 256, 32)
#define LT_CFG_ONE6(_) _(6, 128, 128, 32)

#define LT_PLAIN(T, TN, BM, BN, BK)                                                          \
  extern "C" __global__ void __launch_bounds__(256)                                         \
      lt_gemm_##TN##_##BM##x##BN##x##BK(const T* __restrict__ A, const T* __restrict__ Bt,  \
                                        float* __restrict__ C, int M, int N, int K) {       \
    __shared__ __attribute__((aligned(16))) unsigned char smem[lt_igemm_dl_lds_bytes<BM, BN, BK>()]; \
    lt_igemm_tile_dl<T, BM, BN, BK, false, 2, 2, true>(                                     \
        lt_mm_geo(M, N, K), A, Bt, C, (long)blockIdx.y * BM, (long)blockIdx.x * BN, smem,  \
        (unsigned)M * K * 2, (unsigned)N * K * 2);                                          \
  }

#define LT_CASE(i, BM, BN, BK)                                                               \
  case i:                                                                                    \
    lt_igemm_tile_dl<T, BM, BN, BK, false, 2, 2, true>(                                     \
        g, A, Bt, d.C, it.m0, it.n0, smem, (unsigned)d.M * d.K * 2, (unsigned)d.N * d.K * 2); \
    break;

// queue: [0..8] the 8 segment offsets of the work list, [9..16] one item
// counter per segment, [17] the count of workgroups done.  Workgroup w runs
// on XCD w % 8 and takes items of that XCD's segment: first by index, then
// from the segment's counter.  With fewer than 8 workgroups, workgroup w
// runs segments w, w + grid, ... alone, in order.
#define LT_GROUPED(T_, TN, CLASS, LIST, ...)                                                 \
  extern "C" __global__ void __launch_bounds__(256) __VA_ARGS__                             \
      lt_ggemm##CLASS##_##TN(const LtGemmDesc* __restrict__ descs,                          \
                             const LtWorkItem* __restrict__ items, int* __restrict__ queue) { \
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
        const LtGemmDesc d = descs[it.gemm];                                                \
        const T* A = (const T*)d.A;                                                         \
        const T* Bt = (const T*)d.Bt;                                                       \
        const LtMmGeo g = lt_mm_geo(d.M, d.N, d.K);                                         \
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
#define LT_GROUPED_ALL(T, TN)                                                               \
  LT_GROUPED(T, TN, , LT_CFG_LIST)                                                          \
  LT_GROUPED(T, TN, _mid, LT_CFG_MID)                                                       \
  LT_GROUPED(T, TN, _small, LT_CFG_SMALL)                                                   \
  LT_GROUPED(T, TN, _c0, LT_CFG_ONE0, LT_WPE(8))                                            \
  LT_GROUPED(T, TN, _c1, LT_CFG_ONE1, LT_WPE(5))                                            \
  LT_GROUPED(T, TN, _c2, LT_CFG_ONE2, LT_WPE(5))                                            \
  LT_GROUPED(T, TN, _c3, LT_CFG_ONE3, LT_WPE(3))                                            \
  LT_GROUPED(T, TN, _c4, LT_CFG_ONE4)                                                       \
  LT_GROUPED(T, TN, _c5, LT_CFG_ONE5)                                                       \
  LT_GROUPED(T, TN, _c6, LT_CFG_ONE6, LT_WPE(3))

LT_GROUPED_ALL(__bf16, bf16)
LT_GROUPED_ALL(_Float16, f16)

// float32 [rows][cols] -> 16-bit [rows][cols], or [cols][rows] when transpose
#define LT_CVT(T, TN)                                                                       \
  extern "C" __global__ void __launch_bounds__(256)                                         \
      lt_cvt_##TN(const float* __restrict__ src, T* __restrict__ dst, int rows, int cols,   \
                  int transpose) {                                                          \
    const long i = (long)blockIdx.x * 256 + threadIdx.x;                                    \
    if (i >= (long)rows * cols) return;                                                     \
    const int r = (int)(i / cols), c = (int)(i % cols);                                     \
    const T v = (T)src[i];                                                                  \
    if (transpose)                                                                          \
      dst[(long)c * rows + r] = v;                                                          \
    else                                                                                    \
      dst[i] = v;                                                                           \
  }

LT_CVT(__bf16, bf16)
LT_CVT(_Float16, f16)
