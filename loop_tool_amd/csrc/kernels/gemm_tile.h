// Hand-written gfx950 GEMM tiles for 16-bit operands with float32
// accumulation, used for batches of independent matmul nests
// (C[m,n] = sum_k A[m,k] * B[k,n]) whose sizes are only known at run time.
//
// Layout: A is [M][K] and B is stored transposed as Bt [N][K], both row-major
// 16-bit (bf16 or fp16) in HBM, so both tiles are read as 16-byte K-contiguous
// vectors; C is [M][N] float32.  K must be a multiple of 8 (the batched
// workload uses multiples of 16).
//
// A workgroup of 256 lanes (4 waves, 2 x 2) computes a BM x BN tile: each wave
// owns (BM/2) x (BN/2) as 32x32 MFMA blocks (v_mfma_f32_32x32x16_{bf16,f16}),
// K advances 32 at a time through double-buffered LDS (one barrier per K
// tile; the next tile's global loads are in flight during the MFMAs).
#pragma once
#include <hip/hip_runtime.h>

typedef __bf16 lt_bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 lt_f16x8 __attribute__((ext_vector_type(8)));
typedef float lt_f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short lt_u16x8 __attribute__((ext_vector_type(8)));

#define LT_BK 32
#define LT_LDS_ROW (LT_BK + 8)  // padded row (16-bit elements): 80 bytes

template <typename T> struct lt_vec8;
template <> struct lt_vec8<__bf16> { typedef lt_bf16x8 type; };
template <> struct lt_vec8<_Float16> { typedef lt_f16x8 type; };

__device__ __forceinline__ lt_f32x16 lt_mfma(lt_bf16x8 a, lt_bf16x8 b, lt_f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ lt_f32x16 lt_mfma(lt_f16x8 a, lt_f16x8 b, lt_f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// LDS bytes one tile configuration needs (double-buffered A and Bt tiles)
template <int BM, int BN>
__host__ __device__ constexpr int lt_tile_lds_bytes() {
  return 2 * (BM + BN) * LT_LDS_ROW * 2;
}

// One BM x BN output tile at (m0, n0).  `smem` holds lt_tile_lds_bytes().
template <typename T, int BM, int BN>
__device__ __forceinline__ void lt_gemm_tile(const T* __restrict__ A, const T* __restrict__ Bt,
                                             float* __restrict__ C, int M, int N, int K,
                                             int m0, int n0, unsigned char* smem) {
  typedef typename lt_vec8<T>::type V;
  constexpr int NT = 256;
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 32, TN = WN / 32;
  constexpr int VA = BM * LT_BK / 8 / NT;  // 16-byte vectors per lane per tile
  constexpr int VB = BN * LT_BK / 8 / NT;
  static_assert(VA >= 1 && VB >= 1, "tile too small for 256 lanes");
  T* sA = (T*)smem;                          // [2][BM][LT_LDS_ROW]
  T* sB = sA + 2 * BM * LT_LDS_ROW;          // [2][BN][LT_LDS_ROW]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  lt_u16x8 ra[VA], rb[VB];
  auto load = [&](int k0) {
#pragma unroll
    for (int v = 0; v < VA; ++v) {
      const int e = tid + v * NT;
      const int r = e / (LT_BK / 8), c = (e % (LT_BK / 8)) * 8;
      const int m = m0 + r, k = k0 + c;
      ra[v] = (m < M && k < K) ? *(const lt_u16x8*)(A + (long)m * K + k) : (lt_u16x8){};
    }
#pragma unroll
    for (int v = 0; v < VB; ++v) {
      const int e = tid + v * NT;
      const int r = e / (LT_BK / 8), c = (e % (LT_BK / 8)) * 8;
      const int n = n0 + r, k = k0 + c;
      rb[v] = (n < N && k < K) ? *(const lt_u16x8*)(Bt + (long)n * K + k) : (lt_u16x8){};
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int v = 0; v < VA; ++v) {
      const int e = tid + v * NT;
      const int r = e / (LT_BK / 8), c = (e % (LT_BK / 8)) * 8;
      *(lt_u16x8*)(sA + (buf * BM + r) * LT_LDS_ROW + c) = ra[v];
    }
#pragma unroll
    for (int v = 0; v < VB; ++v) {
      const int e = tid + v * NT;
      const int r = e / (LT_BK / 8), c = (e % (LT_BK / 8)) * 8;
      *(lt_u16x8*)(sB + (buf * BN + r) * LT_LDS_ROW + c) = rb[v];
    }
  };

  lt_f32x16 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = (lt_f32x16){};

  const int nkt = (K + LT_BK - 1) / LT_BK;
  int buf = 0;
  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < nkt; ++kt) {
    const bool more = kt + 1 < nkt;
    if (more) load((kt + 1) * LT_BK);
#pragma unroll
    for (int kk = 0; kk < LT_BK / 16; ++kk) {
      V av[TM], bv[TN];
#pragma unroll
      for (int a = 0; a < TM; ++a)
        av[a] = *(const V*)(sA + (buf * BM + wm * WM + a * 32 + (lane & 31)) * LT_LDS_ROW +
                            kk * 16 + (lane >> 5) * 8);
#pragma unroll
      for (int b = 0; b < TN; ++b)
        bv[b] = *(const V*)(sB + (buf * BN + wn * WN + b * 32 + (lane & 31)) * LT_LDS_ROW +
                            kk * 16 + (lane >> 5) * 8);
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = lt_mfma(av[a], bv[b], acc[a][b]);
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
    const int n = n0 + wn * WN + b * 32 + (lane & 31);
    if (n >= N) continue;
#pragma unroll
    for (int a = 0; a < TM; ++a) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm * WM + a * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (m < M) C[(long)m * N + n] = acc[a][b][r];
      }
    }
  }
}
