// BF16x3 split-precision GEMMs on gfx950 MFMA (Henry, Tang & Heinecke 2019,
// "Leveraging the bfloat16 Artificial Intelligence Datatype For
// Higher-Precision Computations"): each FP64 operand is rounded to float and
// split a = hi + lo with hi = bf16(a), lo = bf16(a - hi); the product keeps
// the three terms hi*hi + hi*lo + lo*hi, accumulated in FP32 by
// v_mfma_f32_16x16x16_bf16. Relative error per product ~2^-16; FP32
// accumulation over K-chunks, FP64 sums across chunks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "common.hpp"
#include "gemm.hpp"

namespace bd {

namespace {
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int PAD = 8;  // bf16 elements of padding per LDS row

// Round-to-nearest-even float -> bf16 bits (finite inputs).
__device__ __forceinline__ unsigned short bf16_bits(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(u >> 16);
}
__device__ __forceinline__ float bf16_float(unsigned short b) { return __uint_as_float(unsigned(b) << 16); }

__device__ __forceinline__ void split(double a, unsigned short& hi, unsigned short& lo) {
  const float f = float(a);
  hi = bf16_bits(f);
  lo = bf16_bits(f - bf16_float(hi));
}

// One 64x64 tile of C (or of a K-chunk partial). TA: A is KxM (A^T B).
template <bool TA>
__device__ void tile_bf16x3(int64_t M, int64_t N, int64_t kbeg, int64_t kend, const double* __restrict__ A,
                            int64_t lda, const double* __restrict__ B, int64_t ldb, int64_t row0, int64_t col0,
                            f32x4 acc[2][2]) {
  __shared__ unsigned short Ah[BM][BK + PAD], Al[BM][BK + PAD];
  __shared__ unsigned short Bh[BN][BK + PAD], Bl[BN][BK + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    // A tile: BM x BK.
    for (int e = tid; e < BM * BK; e += 256) {
      int i, k;
      if (TA) {
        k = e % BK;
        i = e / BK;
      } else {
        i = e % BM;
        k = e / BM;
      }
      const int64_t gi = row0 + i, gk = k0 + k;
      double a = 0.0;
      if (gi < M && gk < kend) a = TA ? A[gk + gi * lda] : A[gi + gk * lda];
      split(a, Ah[i][k], Al[i][k]);
    }
    // B tile: BK x BN stored [n][k].
    for (int e = tid; e < BK * BN; e += 256) {
      const int k = e % BK, j = e / BK;
      const int64_t gk = k0 + k, gj = col0 + j;
      const double b = (gk < kend && gj < N) ? B[gk + gj * ldb] : 0.0;
      split(b, Bh[j][k], Bl[j][k]);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int kl = kk + 4 * (lane >> 4);
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) {
        const int r = wm * 32 + bi * 16 + (lane & 15);
        const bf16x4 ah = *reinterpret_cast<const bf16x4*>(&Ah[r][kl]);
        const bf16x4 al = *reinterpret_cast<const bf16x4*>(&Al[r][kl]);
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          const int c = wn * 32 + bj * 16 + (lane & 15);
          const bf16x4 bh = *reinterpret_cast<const bf16x4*>(&Bh[c][kl]);
          const bf16x4 bl = *reinterpret_cast<const bf16x4*>(&Bl[c][kl]);
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc[bi][bj], 0, 0, 0);
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc[bi][bj], 0, 0, 0);
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc[bi][bj], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
}

template <bool TA>
__global__ __launch_bounds__(256) void k_gemm_bf16x3(int64_t M, int64_t N, int64_t K, int64_t chunk, double alpha,
                                                     const double* __restrict__ A, int64_t lda,
                                                     const double* __restrict__ B, int64_t ldb, double beta,
                                                     double* __restrict__ C, int64_t ldc, double* __restrict__ part) {
  const int64_t row0 = int64_t(blockIdx.x) * BM, col0 = int64_t(blockIdx.y) * BN;
  const int64_t kbeg = int64_t(blockIdx.z) * chunk, kend = std::min(K, kbeg + chunk);
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  tile_bf16x3<TA>(M, N, kbeg, kend, A, lda, B, ldb, row0, col0, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  double* out = part ? part + int64_t(blockIdx.z) * M * N : nullptr;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t gi = row0 + wm * 32 + bi * 16 + 4 * (lane >> 4) + i;
        const int64_t gj = col0 + wn * 32 + bj * 16 + (lane & 15);
        if (gi >= M || gj >= N) continue;
        const double v = double(acc[bi][bj][i]);
        if (out) {
          out[gi + gj * M] = v;
        } else {
          double* c = C + gi + gj * ldc;
          *c = beta == 0.0 ? alpha * v : alpha * v + beta * *c;
        }
      }
}

__global__ void k_reduce_parts(int64_t M, int64_t N, int splits, double alpha, const double* __restrict__ part,
                               double beta, double* __restrict__ C, int64_t ldc) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= M * N) return;
  const int64_t i = e % M, j = e / M;
  double s = 0.0;
  for (int p = 0; p < splits; ++p) s += part[int64_t(p) * M * N + e];
  double* c = C + i + j * ldc;
  *c = beta == 0.0 ? alpha * s : alpha * s + beta * *c;
}

int bf_splits(int64_t M, int64_t N, int64_t K) {
  const int64_t tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  int64_t want = std::max<int64_t>(1, 1024 / std::max<int64_t>(tiles, 1));
  int64_t maxs = std::max<int64_t>(1, K / 256);
  return int(std::min<int64_t>({want, maxs, 256}));
}
}  // namespace

void gemm_nn_bf16x3(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                    const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  if (M <= 0 || N <= 0) return;
  dim3 grid(unsigned((M + BM - 1) / BM), unsigned((N + BN - 1) / BN), 1);
  hipLaunchKernelGGL(k_gemm_bf16x3<false>, grid, dim3(256), 0, s, M, N, K, K, alpha, A, lda, B, ldb, beta, C,
                     ldc, nullptr);
}

// Covers every product M' x N' with M' <= M, N' <= N at this K (bf_splits
// is not monotone in the shape): a split count above one is at most
// 1024 / tiles with tiles >= M'N' / (BM * BN), and at most min(K / 256, 256).
int64_t gemm_tn_bf16x3_work(int64_t M, int64_t N, int64_t K) {
  const int64_t smax = std::min<int64_t>(std::max<int64_t>(1, K / 256), 256);
  return std::max(M * N, std::min(int64_t(1024) * BM * BN, smax * M * N));
}

void gemm_tn_bf16x3(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                    const double* B, int64_t ldb, double beta, double* C, int64_t ldc, double* work) {
  if (M <= 0 || N <= 0) return;
  const int splits = bf_splits(M, N, K);
  int64_t chunk = (K + splits - 1) / splits;
  chunk = ((chunk + BK - 1) / BK) * BK;
  const int used = std::max(1, int((K + chunk - 1) / std::max<int64_t>(chunk, 1)));
  dim3 grid(unsigned((M + BM - 1) / BM), unsigned((N + BN - 1) / BN), unsigned(used));
  hipLaunchKernelGGL(k_gemm_bf16x3<true>, grid, dim3(256), 0, s, M, N, K, chunk, 1.0, A, lda, B, ldb, 0.0,
                     nullptr, int64_t(0), work);
  const int64_t total = M * N;
  hipLaunchKernelGGL(k_reduce_parts, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s, M, N, used, alpha,
                     work, beta, C, ldc);
}

// Lane-layout probe: D = A B for one 16x16x16 bf16 MFMA with A, B given as
// bf16 bits (row-major 16x16 each), D float row-major.
__global__ void k_mfma_probe(const unsigned short* __restrict__ A, const unsigned short* __restrict__ B,
                             float* __restrict__ D) {
  const int lane = threadIdx.x;
  bf16x4 a, b;
  for (int i = 0; i < 4; ++i) {
    a[i] = short(A[(lane & 15) * 16 + 4 * (lane >> 4) + i]);
    b[i] = short(B[(4 * (lane >> 4) + i) * 16 + (lane & 15)]);
  }
  f32x4 d = {0.f, 0.f, 0.f, 0.f};
  d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
  for (int i = 0; i < 4; ++i) D[(4 * (lane >> 4) + i) * 16 + (lane & 15)] = d[i];
}

void mfma_probe(hipStream_t s, const unsigned short* A, const unsigned short* B, float* D) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, A, B, D);
}

}  // namespace bd
