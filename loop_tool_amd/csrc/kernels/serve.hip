// Shape-generic kernels of the serving runtime (gfx950).
//
// The hip backend compiles one module per LoopTree with every size baked in;
// a server facing a stream of programs of ever-new shapes cannot compile per
// request.  These kernels take the sizes at run time and serve every program
// whose lazy graph has the structure of one of the families below (the
// server matches the structure with sizes masked):
//
//   lt_sv_fill            inputs materialised in HBM from a seed (the
//                         SplitMix64 uniform of the test programs, bit-exact)
//   lt_sv_softmax_*       row softmax: max-reduce, subtract, exp, sum-reduce,
//                         divide over the last var (one workgroup per row)
//   lt_sv_maxpool2        2x2 max-pool, stride 2, over the last two vars
//   lt_sv_stage_*         permute + zero-halo + round to 16 bits (operand
//                         staging for the implicit GEMM, as the hip backend's
//                         staged lowering does per nest)
//   lt_sv_igemm_*         implicit-GEMM tile (igemm_tile.h) with run-time
//                         geometry: matmul and 3x3 convolution
//   lt_sv_checksum        position-sensitive 64-bit checksum of a buffer
//                         (lt_sv_zero64 clears its word first)
#include <hip/hip_runtime.h>

#include "igemm_tile.h"

// ----------------------------------------------------------------- inputs --
// value i of seed s: SplitMix64 of s + (i + 1) * golden, top 24 bits as a
// uniform in [-1, 1) (exact in float32), times `scale`
extern "C" __global__ void __launch_bounds__(256)
    lt_sv_fill(float* __restrict__ p, long n, unsigned long long seed, float scale) {
  const unsigned long long golden = 0x9E3779B97F4A7C15ULL;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    unsigned long long z = seed + (unsigned long long)(i + 1) * golden;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    const int m = (int)(z >> 40) - (1 << 23);
    p[i] = ((float)m * (1.0f / 8388608.0f)) * scale;
  }
}

// ---------------------------------------------------------------- softmax --
__device__ __forceinline__ float lt_sv_wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float lt_sv_wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// One workgroup of B lanes per row, each lane holding up to PER values of the
// row in registers (column c = lane + p * B): one HBM read and one write.
template <int B, int PER>
__device__ __forceinline__ void lt_sv_softmax_row(const float* __restrict__ X,
                                                  float* __restrict__ Y, long R, int C) {
  __shared__ float red[B / 64];
  const long r = blockIdx.x;
  if (r >= R) return;
  const float* x = X + r * C;
  float* y = Y + r * C;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  float v[PER];
  float m = -__builtin_huge_valf();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int c = t + p * B;
    v[p] = c < C ? __builtin_nontemporal_load(x + c) : -__builtin_huge_valf();
    m = fmaxf(m, v[p]);
  }
  m = lt_sv_wave_max(m);
  if (B > 64) {
    if (lane == 0) red[wave] = m;
    __syncthreads();
    m = red[0];
#pragma unroll
    for (int w = 1; w < B / 64; ++w) m = fmaxf(m, red[w]);
    __syncthreads();
  }
  float s = 0.0f;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int c = t + p * B;
    v[p] = c < C ? expf(v[p] - m) : 0.0f;
    s += v[p];
  }
  s = lt_sv_wave_sum(s);
  if (B > 64) {
    if (lane == 0) red[wave] = s;
    __syncthreads();
    s = 0.0f;
#pragma unroll
    for (int w = 0; w < B / 64; ++w) s += red[w];
  }
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int c = t + p * B;
    if (c < C) __builtin_nontemporal_store(v[p] / s, y + c);
  }
}

#define LT_SV_SOFTMAX(B, PER)                                                            \
  extern "C" __global__ void __launch_bounds__(B)                                       \
      lt_sv_softmax_##B##_##PER(const float* __restrict__ X, float* __restrict__ Y, long R, \
                                int C) {                                                 \
    lt_sv_softmax_row<B, PER>(X, Y, R, C);                                               \
  }
LT_SV_SOFTMAX(64, 4)     // C <= 256
LT_SV_SOFTMAX(64, 16)    // C <= 1024
LT_SV_SOFTMAX(256, 8)    // C <= 2048
LT_SV_SOFTMAX(256, 16)   // C <= 4096
LT_SV_SOFTMAX(256, 32)   // C <= 8192
LT_SV_SOFTMAX(1024, 32)  // C <= 32768

// ---------------------------------------------------------------- maxpool --
// Y[q][ho][wo] = max of X[q][2ho..2ho+1][2wo..2wo+1], q over the leading vars
extern "C" __global__ void __launch_bounds__(256)
    lt_sv_maxpool2(const float* __restrict__ X, float* __restrict__ Y, long Q, int H, int W) {
  const int Ho = H / 2, Wo = W / 2;
  const long n = Q * Ho * Wo;
  for (long o = (long)blockIdx.x * 256 + threadIdx.x; o < n; o += (long)gridDim.x * 256) {
    const int wo = (int)(o % Wo);
    const long t = o / Wo;
    const int ho = (int)(t % Ho);
    const long q = t / Ho;
    const float* x = X + (q * H + 2 * ho) * W + 2 * wo;
    float a = x[0], b = x[1], c = x[W], d = x[W + 1];
    // the reference's select-style max, folded in window order
    float m = a;
    m = b > m ? b : m;
    m = c > m ? c : m;
    m = d > m ? d : m;
    __builtin_nontemporal_store(m, Y + o);
  }
}

// ------------------------------------------------------- operand staging --
// Staged tensor S[s0][s1][s2][s3] (s3 innermost, padded): element s reads the
// source at sum_d (s_d - padlo_d) * src_stride_d when every such index lies in
// [0, src_size_d) and s3 < kin, else 0.  Each lane writes 8 consecutive s3
// elements (16 bytes); lanes walk dim `ld` first so the source reads coalesce.
struct LtSvStage {
  long ssize[4];
  long src_stride[4];
  long src_size[4];
  long padlo[4];
  long kin;
  int ld;
  int pad_;
};

template <typename T>
__device__ __forceinline__ void lt_sv_stage(const LtSvStage& d, const float* __restrict__ in,
                                            T* __restrict__ S) {
  typedef T T8 __attribute__((ext_vector_type(8)));
  const long g3 = d.ssize[3] / 8;
  const long nvec = d.ssize[0] * d.ssize[1] * d.ssize[2] * g3;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < nvec; t += (long)gridDim.x * 256) {
    long s[4];
    long u = t;
    if (d.ld >= 0) {
      s[d.ld] = u % d.ssize[d.ld];
      u /= d.ssize[d.ld];
    }
    s[3] = (u % g3) * 8;
    u /= g3;
#pragma unroll
    for (int q = 2; q >= 0; --q) {
      if (q == d.ld) continue;
      s[q] = u % d.ssize[q];
      u /= d.ssize[q];
    }
    bool ok = true;
    long src = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long i = s[q] - d.padlo[q];
      ok = ok && i >= 0 && i < d.src_size[q];
      src += i * d.src_stride[q];
    }
    const long dst = ((s[0] * d.ssize[1] + s[1]) * d.ssize[2] + s[2]) * d.ssize[3] + s[3];
    T8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e)
      v[e] = (ok && s[3] + e < d.kin) ? (T)in[src + e * d.src_stride[3]] : (T)0.0f;
    *(T8*)(S + dst) = v;
  }
}

extern "C" __global__ void __launch_bounds__(256)
    lt_sv_stage_bf16(LtSvStage d, const float* __restrict__ in, __bf16* __restrict__ S) {
  lt_sv_stage<__bf16>(d, in, S);
}
extern "C" __global__ void __launch_bounds__(256)
    lt_sv_stage_f16(LtSvStage d, const float* __restrict__ in, _Float16* __restrict__ S) {
  lt_sv_stage<_Float16>(d, in, S);
}

// ------------------------------------------------------- implicit GEMM --
// Run-time geometry: P rows i (stride p_rs); Q rows j = (a * qd1 + b) * qd2 + c;
// outer K index ko = x * ko2 + y; output offset out_row(i) + out_col(j).
struct LtSvGeo {
  int KO, KI;
  long MD, ND;
  long p_rs;
  long qd1, qd2, qc0, qc1, qc2, qcst;
  long ko2, pk0, pk1, qk0, qk1;
  long o_rs, oc0, oc1, oc2;
  unsigned pbytes, qbytes;  // staged operand sizes (the tile's buffer range)
  __device__ __forceinline__ long p_row(long i, bool& ok) const {
    ok = i < MD;
    return i * p_rs;
  }
  __device__ __forceinline__ long q_row(long j, bool& ok) const {
    ok = j < ND;
    const long c = j % qd2, t = j / qd2, b = t % qd1, a = t / qd1;
    return a * qc0 + b * qc1 + c * qc2 + qcst;
  }
  __device__ __forceinline__ long p_koff(int ko) const { return (ko / ko2) * pk0 + (ko % ko2) * pk1; }
  __device__ __forceinline__ long q_koff(int ko) const { return (ko / ko2) * qk0 + (ko % ko2) * qk1; }
  __device__ __forceinline__ long out_row(long i) const { return i * o_rs; }
  __device__ __forceinline__ long out_col(long j) const {
    const long c = j % qd2, t = j / qd2, b = t % qd1, a = t / qd1;
    return a * oc0 + b * oc1 + c * oc2;
  }
};

#define LT_SV_IGEMM(T, TN, BM, BN, BK)                                                       \
  extern "C" __global__ void __launch_bounds__(256)                                         \
      lt_sv_igemm_##TN##_##BM##x##BN##_##BK(LtSvGeo g, const T* __restrict__ P,             \
                                            const T* __restrict__ Q, float* __restrict__ out) { \
    __shared__ __attribute__((aligned(16))) unsigned char smem[lt_igemm_lds_bytes<BM, BN, BK>()]; \
    lt_igemm_tile_pipe<T, BM, BN, BK, false, 2, 2, false>(                                  \
        g, P, Q, out, (long)blockIdx.y * BM, (long)lt_igemm_swizzle(blockIdx.x, gridDim.x) * BN, \
        smem, g.pbytes, g.qbytes);                                                          \
  }
#define LT_SV_IGEMM_ALL(T, TN)      \
  LT_SV_IGEMM(T, TN, 64, 64, 32)    \
  LT_SV_IGEMM(T, TN, 128, 64, 32)   \
  LT_SV_IGEMM(T, TN, 64, 128, 32)   \
  LT_SV_IGEMM(T, TN, 128, 128, 64)  \
  LT_SV_IGEMM(T, TN, 128, 256, 32)
LT_SV_IGEMM_ALL(__bf16, bf16)
LT_SV_IGEMM_ALL(_Float16, f16)

// --------------------------------------------------------------- checksum --
extern "C" __global__ void __launch_bounds__(64) lt_sv_zero64(unsigned long long* p) {
  if (threadIdx.x == 0) *p = 0ULL;
}

// sum over i of mix(bits_i, i): order-independent, so the atomic sum is exact
// and equal for bitwise-equal buffers
extern "C" __global__ void __launch_bounds__(256)
    lt_sv_checksum(const unsigned* __restrict__ p, long n, unsigned long long* __restrict__ out) {
  unsigned long long acc = 0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    unsigned long long z = ((unsigned long long)p[i] << 32) ^ (unsigned long long)i;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    acc += z ^ (z >> 31);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, acc);
}
