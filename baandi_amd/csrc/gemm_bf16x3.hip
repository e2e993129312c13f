// BF16x3 split-precision GEMMs on gfx950 MFMA (Henry, Tang & Heinecke 2019,
// "Leveraging the bfloat16 Artificial Intelligence Datatype For
// Higher-Precision Computations"): each FP64 operand is rounded to float and
// split a = hi + lo with hi = bf16(a), lo = bf16(a - hi); the product keeps
// the three terms hi*hi + hi*lo + lo*hi, accumulated in FP32 by MFMA
// (v_mfma_f32_16x16x16_bf16 for NN, v_mfma_f32_16x16x32_bf16 for TN).
// Relative error per product ~2^-16; FP32 accumulation within a K chunk, FP64
// sums across chunks.
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

// ---------------------------------------------------------------------------
// TN products (the Davidson's border Grams and P^T X: K is the long basis
// dimension): C = A^T B with A K x M, B K x N. A workgroup takes a 128 x 128
// tile of C over one K chunk; the four waves each hold a 64 x 64 block as
// 4 x 4 accumulators of v_mfma_f32_16x16x32_bf16 (gfx950), so every A element
// of a chunk is read by ceil(N / 128) workgroups and every B element by
// ceil(M / 128) (the Si64 Gram 320 x 160: A twice, B three times). FP64 tiles
// are split once per load into hi / lo bf16 halves in LDS. Partials per
// chunk are FP32, summed over the chunks in FP64 in chunk order.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int WM = 128, WN = 128, WK = 32, WPAD = 8;  // a bf16 LDS row: WK + WPAD = 40 elements (80 B)

__device__ __forceinline__ void split2(double a0, double a1, unsigned int& hi, unsigned int& lo) {
  unsigned short h0, l0, h1, l1;
  split(a0, h0, l0);
  split(a1, h1, l1);
  hi = unsigned(h0) | (unsigned(h1) << 16);
  lo = unsigned(l0) | (unsigned(l1) << 16);
}

// Workgroups are dealt to the 8 XCDs round robin (id % 8); the logical
// (tile, chunk) index is laid out so that the tiles of one K chunk run on one
// XCD, one after the other, and share its L2 for the chunk's rows of A and B.
constexpr int kXcds = 8;

__global__ __launch_bounds__(256, 2) void k_gemm_tn_wide(int64_t M, int64_t N, int64_t K, int64_t chunk, int tiles_m,
                                                      int tiles, int total, const double* __restrict__ A,
                                                      int64_t lda, const double* __restrict__ B, int64_t ldb,
                                                      float* __restrict__ part) {
  __shared__ __align__(16) unsigned short sAh[WM][WK + WPAD];
  __shared__ __align__(16) unsigned short sAl[WM][WK + WPAD];
  __shared__ __align__(16) unsigned short sBh[WN][WK + WPAD];
  __shared__ __align__(16) unsigned short sBl[WN][WK + WPAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int per_xcd = (total + kXcds - 1) / kXcds;
  const int logical = (int(blockIdx.x) % kXcds) * per_xcd + int(blockIdx.x) / kXcds;
  if (logical >= total) return;
  const int tile = logical % tiles, kc = logical / tiles;
  const int64_t row0 = int64_t(tile % tiles_m) * WM, col0 = int64_t(tile / tiles_m) * WN;
  const int64_t kbeg = int64_t(kc) * chunk, kend = std::min(K, kbeg + chunk);
  f32x4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  // 128 rows x 32 k of A and of B per step, two consecutive k per thread
  // and row: 16 threads cover one column's 32 k (256 contiguous bytes). The
  // next step's values are loaded into registers (all 32 loads in flight)
  // while the current step's products run.
  double va[8][2], vb[8][2];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int idx = tid + 256 * r, i = idx >> 4, kp = (idx & 15) * 2;
      const int64_t gk = k0 + kp, gi = row0 + i, gj = col0 + i;
      va[r][0] = (gi < M && gk < kend) ? A[gk + gi * lda] : 0.0;
      va[r][1] = (gi < M && gk + 1 < kend) ? A[gk + 1 + gi * lda] : 0.0;
      vb[r][0] = (gj < N && gk < kend) ? B[gk + gj * ldb] : 0.0;
      vb[r][1] = (gj < N && gk + 1 < kend) ? B[gk + 1 + gj * ldb] : 0.0;
    }
  };
  if (kbeg < kend) load(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += WK) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int idx = tid + 256 * r, i = idx >> 4, kp = (idx & 15) * 2;
      unsigned int hi, lo;
      split2(va[r][0], va[r][1], hi, lo);
      *reinterpret_cast<unsigned int*>(&sAh[i][kp]) = hi;
      *reinterpret_cast<unsigned int*>(&sAl[i][kp]) = lo;
      split2(vb[r][0], vb[r][1], hi, lo);
      *reinterpret_cast<unsigned int*>(&sBh[i][kp]) = hi;
      *reinterpret_cast<unsigned int*>(&sBl[i][kp]) = lo;
    }
    __syncthreads();
    if (k0 + WK < kend) load(k0 + WK);
    // Operands of v_mfma_f32_16x16x32_bf16: lane l holds row (l & 15), k
    // 8 (l >> 4) .. + 7 of A^T and column (l & 15), the same k, of B.
    // B's four column blocks stay in registers; A's row blocks are read one
    // at a time (fewer live registers: two workgroups fit a CU).
    const int koff = 8 * (lane >> 4);
    bf16x8 bh[4], bl[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = wn * 64 + q * 16 + (lane & 15);
      bh[q] = *reinterpret_cast<const bf16x8*>(&sBh[c][koff]);
      bl[q] = *reinterpret_cast<const bf16x8*>(&sBl[c][koff]);
    }
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
      const int r = wm * 64 + bi * 16 + (lane & 15);
      const bf16x8 ah = *reinterpret_cast<const bf16x8*>(&sAh[r][koff]);
      const bf16x8 al = *reinterpret_cast<const bf16x8*>(&sAl[r][koff]);
#pragma unroll
      for (int bj = 0; bj < 4; ++bj) {
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[bj], acc[bi][bj], 0, 0, 0);
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[bj], acc[bi][bj], 0, 0, 0);
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[bj], acc[bi][bj], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float* out = part + int64_t(kc) * M * N;
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t gi = row0 + wm * 64 + bi * 16 + 4 * (lane >> 4) + i;
        const int64_t gj = col0 + wn * 64 + bj * 16 + (lane & 15);
        if (gi < M && gj < N) out[gi + gj * M] = acc[bi][bj][i];
      }
}

__global__ void k_reduce_parts_f32(int64_t M, int64_t N, int splits, double alpha, const float* __restrict__ part,
                                   double beta, double* __restrict__ C, int64_t ldc) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= M * N) return;
  const int64_t i = e % M, j = e / M;
  // In chunk order; eight loads in flight at a time.
  double s = 0.0;
  int p = 0;
  for (; p + 8 <= splits; p += 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = part[int64_t(p + u) * M * N + e];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += double(v[u]);
  }
  for (; p < splits; ++p) s += double(part[int64_t(p) * M * N + e]);
  double* c = C + i + j * ldc;
  *c = beta == 0.0 ? alpha * s : alpha * s + beta * *c;
}

// Chunks of K for the wide TN kernel: about 512 workgroups in all, chunks of
// at least 256 rows (a multiple of WK), at most 256 chunks.
int tn_splits(int64_t M, int64_t N, int64_t K) {
  const int64_t tiles = ((M + WM - 1) / WM) * ((N + WN - 1) / WN);
  const int64_t want = (512 + tiles - 1) / std::max<int64_t>(tiles, 1);
  const int64_t maxs = std::max<int64_t>(1, K / 256);
  return int(std::max<int64_t>(1, std::min<int64_t>({want, maxs, 256})));
}
}  // namespace

void gemm_nn_bf16x3(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                    const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  if (M <= 0 || N <= 0) return;
  dim3 grid(unsigned((M + BM - 1) / BM), unsigned((N + BN - 1) / BN), 1);
  hipLaunchKernelGGL(k_gemm_bf16x3<false>, grid, dim3(256), 0, s, M, N, K, K, alpha, A, lda, B, ldb, beta, C,
                     ldc, nullptr);
}

// Doubles of workspace for every TN product M' x N' with M' <= M, N' <= N at
// this K (FP32 partials, two to a double): splits <= min(K / 256, 256) and
// splits <= 512 / tiles + 1 with tiles >= M'N' / (WM WN), so splits x M'N'
// <= min(min(K / 256, 256) MN, 512 WM WN + MN).
int64_t gemm_tn_bf16x3_work(int64_t M, int64_t N, int64_t K) {
  const int64_t smax = std::min<int64_t>(std::max<int64_t>(1, K / 256), 256);
  const int64_t floats = std::min(smax * M * N, int64_t(512) * WM * WN + M * N);
  return std::max(M * N, (floats + 1) / 2 + 1);
}

void gemm_tn_bf16x3(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                    const double* B, int64_t ldb, double beta, double* C, int64_t ldc, double* work) {
  if (M <= 0 || N <= 0) return;
  const int splits = tn_splits(M, N, K);
  int64_t chunk = (K + splits - 1) / splits;
  chunk = ((chunk + WK - 1) / WK) * WK;
  const int used = K > 0 ? std::max(1, int((K + chunk - 1) / std::max<int64_t>(chunk, 1))) : 1;
  float* part = reinterpret_cast<float*>(work);
  const int tiles_m = int((M + WM - 1) / WM), tiles = tiles_m * int((N + WN - 1) / WN);
  const int wgs = tiles * used;
  const int grid = (wgs + kXcds - 1) / kXcds * kXcds;
  hipLaunchKernelGGL(k_gemm_tn_wide, dim3(grid), dim3(256), 0, s, M, N, K, chunk, tiles_m, tiles, wgs, A, lda, B,
                     ldb, part);
  const int64_t total = M * N;
  hipLaunchKernelGGL(k_reduce_parts_f32, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s, M, N, used, alpha,
                     part, beta, C, ldc);
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
