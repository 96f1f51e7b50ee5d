// Kernel library for batches of run-time-sized matmul nests (gfx950).
//
//   lt_gemm_<t>_<BM>x<BN>(A, Bt, C, M, N, K)     one GEMM per launch
//   lt_ggemm_<t>(descs, items, n_items)          persistent grouped launch
//   lt_cvt_<t>(src, dst, rows, cols, transpose)  float32 -> 16-bit operands
//
// The grouped kernel walks a host-built work list of output tiles from many
// GEMMs (ordered by decreasing cost); workgroup w handles items w, w+grid,
// w+2*grid, ...  Every item carries its own tile configuration, chosen per
// shape bucket by the autotuner, so one launch mixes tile shapes.
#include "gemm_tile.h"

struct LtGemmDesc {
  const void* A;
  const void* Bt;
  float* C;
  int M, N, K, pad;
};

struct LtWorkItem {
  int gemm, m0, n0, cfg;
};

// tile configurations: index -> (BM, BN)
#define LT_CFG_LIST(_) _(0, 64, 64) _(1, 128, 64) _(2, 64, 128) _(3, 128, 128) _(4, 256, 128)

constexpr int lt_max_lds() {
  return lt_tile_lds_bytes<256, 128>();
}

#define LT_PLAIN(T, TN, BM, BN)                                                            \
  extern "C" __global__ void __launch_bounds__(256)                                        \
      lt_gemm_##TN##_##BM##x##BN(const T* __restrict__ A, const T* __restrict__ Bt,        \
                                 float* __restrict__ C, int M, int N, int K) {             \
    __shared__ __attribute__((aligned(16))) unsigned char smem[lt_tile_lds_bytes<BM, BN>()]; \
    lt_gemm_tile<T, BM, BN>(A, Bt, C, M, N, K, blockIdx.y * BM, blockIdx.x * BN, smem);    \
  }

#define LT_CASE(T) \
  switch (it.cfg)

#define LT_GROUPED(T, TN)                                                                   \
  extern "C" __global__ void __launch_bounds__(256)                                         \
      lt_ggemm_##TN(const LtGemmDesc* __restrict__ descs, const LtWorkItem* __restrict__ items, \
                    int n_items) {                                                          \
    __shared__ __attribute__((aligned(16))) unsigned char smem[lt_max_lds()];               \
    for (int i = blockIdx.x; i < n_items; i += gridDim.x) {                                 \
      const LtWorkItem it = items[i];                                                       \
      const LtGemmDesc d = descs[it.gemm];                                                  \
      const T* A = (const T*)d.A;                                                           \
      const T* Bt = (const T*)d.Bt;                                                         \
      switch (it.cfg) {                                                                     \
        case 0: lt_gemm_tile<T, 64, 64>(A, Bt, d.C, d.M, d.N, d.K, it.m0, it.n0, smem); break;   \
        case 1: lt_gemm_tile<T, 128, 64>(A, Bt, d.C, d.M, d.N, d.K, it.m0, it.n0, smem); break;  \
        case 2: lt_gemm_tile<T, 64, 128>(A, Bt, d.C, d.M, d.N, d.K, it.m0, it.n0, smem); break;  \
        case 3: lt_gemm_tile<T, 128, 128>(A, Bt, d.C, d.M, d.N, d.K, it.m0, it.n0, smem); break; \
        default: lt_gemm_tile<T, 256, 128>(A, Bt, d.C, d.M, d.N, d.K, it.m0, it.n0, smem); break; \
      }                                                                                     \
      __syncthreads(); /* LDS reuse by the next item */                                    \
    }                                                                                       \
  }

#define LT_ALL(T, TN)        \
  LT_PLAIN(T, TN, 64, 64)    \
  LT_PLAIN(T, TN, 128, 64)   \
  LT_PLAIN(T, TN, 64, 128)   \
  LT_PLAIN(T, TN, 128, 128)  \
  LT_PLAIN(T, TN, 256, 128)  \
  LT_GROUPED(T, TN)

LT_ALL(__bf16, bf16)
LT_ALL(_Float16, f16)

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
