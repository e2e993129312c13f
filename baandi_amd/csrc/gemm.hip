// FP64 tiled GEMM kernels on the matrix cores: 64 x BN output tiles (BN 64;
// 128 is built but slower here), 256 threads (4 waves, a 32 x BN/2 quadrant each), K
// staged through LDS 16 at a time (the next slice loaded into registers while
// the current one is multiplied), products by v_mfma_f64_16x16x4f64 (lane l
// holds A[l%16][l/16], B[l/16][l%16] and D[l/16 + 4v][l%16], v = 0..3;
// pinned by the probe below).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "common.hpp"
#include "gemm.hpp"

namespace bd {

namespace {
constexpr int TM = 64, TK = 16;

typedef double f64x4 __attribute__((ext_vector_type(4)));

// acc[bi][bj] += the wave's quadrant (rows wr*32 + [0, 32), cols wc*BN/2 +
// [0, BN/2)) of As^T Bs for one TK slice.
template <int BN>
__device__ __forceinline__ void mfma_tile(const double (*As)[TM + 1], const double (*Bs)[BN + 1], int lane, int wr,
                                          int wc, f64x4 (&acc)[2][BN / 32]) {
  constexpr int NB = BN / 32;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < TK; kk += 4) {
    double a[2], b[NB];
#pragma unroll
    for (int t = 0; t < 2; ++t) a[t] = As[kk + lk][wr * 32 + t * 16 + li];
#pragma unroll
    for (int t = 0; t < NB; ++t) b[t] = Bs[kk + lk][wc * (BN / 2) + t * 16 + li];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < NB; ++bj)
        acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], b[bj], acc[bi][bj], 0, 0, 0);
  }
}

// Slice loaders: every thread owns four consecutive elements of the A slice
// and 4 * BN/64 of the B slice (consecutive along the contiguous index), so
// the addresses are fixed per thread and advance by a constant per slice.
// Interior slices load unchecked; edge slices check every element (zero
// padding). The next slice is loaded into registers while the current one is
// multiplied out of LDS.
template <int BN>
struct SliceRegs {
  double a[4];
  double b[4 * (BN / 64)];
};

// B slice (TK k x BN columns, k contiguous): thread t holds k = 4 (t & 3) + u
// of column g * 64 + (t >> 2).
template <int BN>
__device__ __forceinline__ void load_b(SliceRegs<BN>& r, const double* __restrict__ B, int64_t ldb, int64_t k0,
                                       int64_t kend, int64_t col0, int64_t N, bool interior) {
  const int tid = threadIdx.x;
  const int kb = (tid & 3) * 4;
#pragma unroll
  for (int g = 0; g < BN / 64; ++g) {
    const int64_t gj = col0 + g * 64 + (tid >> 2);
    const double* p = B + (k0 + kb) + gj * ldb;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      r.b[g * 4 + u] = interior ? p[u] : ((k0 + kb + u < kend && gj < N) ? p[u] : 0.0);
  }
}

template <int BN>
__device__ __forceinline__ void stash_b(const SliceRegs<BN>& r, double (*Bs)[BN + 1]) {
  const int tid = threadIdx.x;
  const int kb = (tid & 3) * 4;
#pragma unroll
  for (int g = 0; g < BN / 64; ++g)
#pragma unroll
    for (int u = 0; u < 4; ++u) Bs[kb + u][g * 64 + (tid >> 2)] = r.b[g * 4 + u];
}

// NN: C = alpha A B + beta C; 64 x BN output tile per workgroup.
template <int BN>
__global__ __launch_bounds__(256) void k_gemm_nn(int64_t M, int64_t N, int64_t K, double alpha,
                                                 const double* __restrict__ A, int64_t lda,
                                                 const double* __restrict__ B, int64_t ldb,
                                                 double beta, double* __restrict__ C, int64_t ldc) {
  constexpr int NB = BN / 32;
  __shared__ double As[TK][TM + 1];
  __shared__ double Bs[TK][BN + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, wr = wave & 1, wc = wave >> 1;
  const int64_t row0 = int64_t(blockIdx.x) * TM, col0 = int64_t(blockIdx.y) * BN;
  f64x4 acc[2][NB];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < NB; ++bj) acc[bi][bj] = f64x4{0.0, 0.0, 0.0, 0.0};
  // A slice (64 rows x TK k, rows contiguous): thread t holds rows
  // 4 (t & 15) + u of k = t >> 4.
  const int ka = tid >> 4, ia = (tid & 15) * 4;
  const bool rows_in = row0 + TM <= M, cols_in = col0 + BN <= N;
  SliceRegs<BN> r;
  auto fetch = [&](int64_t k0) {
    const bool interior = rows_in && cols_in && k0 + TK <= K;
    const double* p = A + (row0 + ia) + (k0 + ka) * lda;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      r.a[u] = interior ? p[u] : ((row0 + ia + u < M && k0 + ka < K) ? p[u] : 0.0);
    load_b<BN>(r, B, ldb, k0, K, col0, N, interior);
  };
  auto stash = [&]() {
#pragma unroll
    for (int u = 0; u < 4; ++u) As[ka][ia + u] = r.a[u];
    stash_b<BN>(r, Bs);
  };
  if (K > 0) {
    fetch(0);
    stash();
    __syncthreads();
  }
  for (int64_t k0 = 0; k0 < K; k0 += TK) {
    const bool more = k0 + TK < K;
    if (more) fetch(k0 + TK);
    mfma_tile<BN>(As, Bs, lane, wr, wc, acc);
    __syncthreads();
    if (more) {
      stash();
      __syncthreads();
    }
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < NB; ++bj) {
      const int64_t gj = col0 + wc * (BN / 2) + bj * 16 + (lane & 15);
      if (gj >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gi = row0 + wr * 32 + bi * 16 + (lane >> 4) + 4 * v;
        if (gi >= M) continue;
        double* p = C + gi + gj * ldc;
        *p = beta == 0.0 ? alpha * acc[bi][bj][v] : alpha * acc[bi][bj][v] + beta * *p;
      }
    }
}

// NN with 32 x 32 output tiles, for products whose 64-wide tiles would fill
// only a fraction of the CUs (the subspace step's m x m and m x bands
// products): four times the workgroups, each wave one 16 x 16 block, so the
// K loop's MFMA chain runs on four times the SIMDs. Same operand layouts and
// K slices as k_gemm_nn.
constexpr int TS = 32;
template <bool TRANS_A>
__global__ __launch_bounds__(256) void k_gemm_s32(int64_t M, int64_t N, int64_t K, double alpha,
                                                  const double* __restrict__ A, int64_t lda,
                                                  const double* __restrict__ B, int64_t ldb, double beta,
                                                  double* __restrict__ C, int64_t ldc) {
  __shared__ double As[TK][TS + 1];
  __shared__ double Bs[TK][TS + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, wr = wave & 1, wc = wave >> 1;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t row0 = int64_t(blockIdx.x) * TS, col0 = int64_t(blockIdx.y) * TS;
  f64x4 acc = f64x4{0.0, 0.0, 0.0, 0.0};
  // A slice (32 rows x TK k): thread t holds rows 2 (t & 15) + u of k = t >> 4.
  // B slice (TK k x 32 columns): thread t holds k = 2 (t & 7) + u of column t >> 3.
  const int ka = tid >> 4, ia = (tid & 15) * 2;
  const int jb = tid >> 3, kb = (tid & 7) * 2;
  double ra[2], rb[2];
  // (TRANS_A: A is K x M, k contiguous; its slice is loaded like B's.)
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (TRANS_A) {
        const int64_t gi = row0 + jb, gk = k0 + kb + u;
        ra[u] = (gi < M && gk < K) ? A[gk + gi * lda] : 0.0;
      } else {
        const int64_t gi = row0 + ia + u, gk = k0 + ka;
        ra[u] = (gi < M && gk < K) ? A[gi + gk * lda] : 0.0;
      }
      const int64_t gj = col0 + jb, gk2 = k0 + kb + u;
      rb[u] = (gj < N && gk2 < K) ? B[gk2 + gj * ldb] : 0.0;
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (TRANS_A)
        As[kb + u][jb] = ra[u];
      else
        As[ka][ia + u] = ra[u];
      Bs[kb + u][jb] = rb[u];
    }
  };
  if (K > 0) {
    fetch(0);
    stash();
    __syncthreads();
  }
  for (int64_t k0 = 0; k0 < K; k0 += TK) {
    const bool more = k0 + TK < K;
    if (more) fetch(k0 + TK);
#pragma unroll
    for (int kk = 0; kk < TK; kk += 4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(As[kk + lk][wr * 16 + li], Bs[kk + lk][wc * 16 + li], acc, 0, 0,
                                                 0);
    __syncthreads();
    if (more) {
      stash();
      __syncthreads();
    }
  }
  const int64_t gj = col0 + wc * 16 + li;
  if (gj >= N) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int64_t gi = row0 + wr * 16 + lk + 4 * v;
    if (gi >= M) continue;
    double* p = C + gi + gj * ldc;
    *p = beta == 0.0 ? alpha * acc[v] : alpha * acc[v] + beta * *p;
  }
}

// TN partial: part[s](MxN) = sum over k in chunk s of A[k,i] B[k,j].
// Workgroups are dealt to the 8 XCDs round robin (id % 8); the logical
// (tile, chunk) index puts the tiles of one K chunk on one XCD, one after the
// other, where they share its L2 for the chunk's rows of A and B.
constexpr int kXcds = 8;

template <int BN>
__global__ __launch_bounds__(256) void k_gemm_tn_part(int64_t M, int64_t N, int64_t K, int64_t chunk, int tiles_m,
                                                      int tiles, int total, const double* __restrict__ A,
                                                      int64_t lda, const double* __restrict__ B, int64_t ldb,
                                                      double* __restrict__ part) {
  constexpr int NB = BN / 32;
  __shared__ double As[TK][TM + 1];
  __shared__ double Bs[TK][BN + 1];
  const int per_xcd = (total + kXcds - 1) / kXcds;
  const int logical = (int(blockIdx.x) % kXcds) * per_xcd + int(blockIdx.x) / kXcds;
  if (logical >= total) return;
  const int tile = logical % tiles, kc = logical / tiles;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, wr = wave & 1, wc = wave >> 1;
  const int64_t row0 = int64_t(tile % tiles_m) * TM, col0 = int64_t(tile / tiles_m) * BN;
  const int64_t kbeg = int64_t(kc) * chunk, kend = std::min(K, kbeg + chunk);
  f64x4 acc[2][NB];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < NB; ++bj) acc[bi][bj] = f64x4{0.0, 0.0, 0.0, 0.0};
  // A slice (TK k x 64 columns i of A, k contiguous): thread t holds
  // k = 4 (t & 3) + u of column i = t >> 2.
  const int ia = tid >> 2, ka = (tid & 3) * 4;
  const bool rows_in = row0 + TM <= M, cols_in = col0 + BN <= N;
  SliceRegs<BN> r;
  auto fetch = [&](int64_t k0) {
    const bool interior = rows_in && cols_in && k0 + TK <= kend;
    const double* p = A + (k0 + ka) + (row0 + ia) * lda;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      r.a[u] = interior ? p[u] : ((k0 + ka + u < kend && row0 + ia < M) ? p[u] : 0.0);
    load_b<BN>(r, B, ldb, k0, kend, col0, N, interior);
  };
  auto stash = [&]() {
#pragma unroll
    for (int u = 0; u < 4; ++u) As[ka + u][ia] = r.a[u];
    stash_b<BN>(r, Bs);
  };
  if (kbeg < kend) {
    fetch(kbeg);
    stash();
    __syncthreads();
  }
  for (int64_t k0 = kbeg; k0 < kend; k0 += TK) {
    const bool more = k0 + TK < kend;
    if (more) fetch(k0 + TK);
    mfma_tile<BN>(As, Bs, lane, wr, wc, acc);
    __syncthreads();
    if (more) {
      stash();
      __syncthreads();
    }
  }
  double* out = part + int64_t(kc) * M * N;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < NB; ++bj) {
      const int64_t gj = col0 + wc * (BN / 2) + bj * 16 + (lane & 15);
      if (gj >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gi = row0 + wr * 32 + bi * 16 + (lane >> 4) + 4 * v;
        if (gi < M) out[gi + gj * M] = acc[bi][bj][v];
      }
    }
}

// Lane-layout probe: D (16x16, row-major) = A (16x4, row-major) B (4x16,
// row-major) by one v_mfma_f64_16x16x4f64 with the operand layout above.
__global__ void k_mfma_f64_probe(const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ D) {
  const int lane = threadIdx.x;
  const double a = A[(lane & 15) * 4 + (lane >> 4)];
  const double b = B[(lane >> 4) * 16 + (lane & 15)];
  f64x4 d = {0.0, 0.0, 0.0, 0.0};
  d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d, 0, 0, 0);
  for (int v = 0; v < 4; ++v) D[((lane >> 4) + 4 * v) * 16 + (lane & 15)] = d[v];
}

__global__ void k_gemm_tn_reduce(int64_t M, int64_t N, int splits, double alpha,
                                 const double* __restrict__ part, double beta,
                                 double* __restrict__ C, int64_t ldc) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= M * N) return;
  const int64_t i = e % M, j = e / M;
  // In chunk order; eight loads in flight at a time.
  double s = 0.0;
  int p = 0;
  for (; p + 8 <= splits; p += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = part[int64_t(p + u) * M * N + e];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; p < splits; ++p) s += part[int64_t(p) * M * N + e];
  double* c = C + i + j * ldc;
  *c = beta == 0.0 ? alpha * s : alpha * s + beta * *c;
}

// Column tile width. 128-wide tiles read A half as often but were slower on
// the Davidson shapes (fewer workgroups, longer B-tile loads): 64 is used.
int col_tile(int64_t) { return 64; }

int tn_splits(int64_t M, int64_t N, int64_t K) {
  const int bn = col_tile(N);
  const int64_t tiles = ((M + TM - 1) / TM) * ((N + bn - 1) / bn);
  // ~512 workgroups (two per CU): more splits cost more in the partial sums
  // than they gain in occupancy on the Davidson shapes.
  int64_t want = std::max<int64_t>(1, 512 / std::max<int64_t>(tiles, 1));
  int64_t maxs = std::max<int64_t>(1, K / 256);
  return int(std::min<int64_t>({want, maxs, 256}));
}
}  // namespace

void gemm_nn(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
             int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  if (M <= 0 || N <= 0) return;
  const int bn = col_tile(N);
  dim3 grid(unsigned((M + TM - 1) / TM), unsigned((N + bn - 1) / bn));
  if (int64_t(grid.x) * grid.y < 128 && K >= TK) {  // under half the CUs busy: quarter-size tiles
    dim3 g32(unsigned((M + TS - 1) / TS), unsigned((N + TS - 1) / TS));
    hipLaunchKernelGGL(k_gemm_s32<false>, g32, dim3(256), 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  } else if (bn == 128)
    hipLaunchKernelGGL(k_gemm_nn<128>, grid, dim3(256), 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  else
    hipLaunchKernelGGL(k_gemm_nn<64>, grid, dim3(256), 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

// Covers the partial sums of every product M' x N' with M' <= M, N' <= N at
// this K, so one buffer serves every call up to its sizing shape:
// tn_splits is not monotone (fewer tiles take more splits). A split count
// above one is at most 512 / tiles with tiles >= M'N' / (TM * 64), and at
// most min(K / 256, 256).
int64_t gemm_tn_work(int64_t M, int64_t N, int64_t K) {
  const int64_t smax = std::min<int64_t>(std::max<int64_t>(1, K / 256), 256);
  return std::max(M * N, std::min(int64_t(512) * TM * col_tile(N), smax * M * N));
}

void gemm_tn(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
             int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
             double* work) {
  if (M <= 0 || N <= 0) return;
  const int splits = tn_splits(M, N, K);
  int64_t chunk = (K + splits - 1) / splits;
  chunk = ((chunk + TK - 1) / TK) * TK;
  const int used = int((K + chunk - 1) / std::max<int64_t>(chunk, 1));
  const int bn = col_tile(N);
  const int tiles_m = int((M + TM - 1) / TM), tiles = tiles_m * int((N + bn - 1) / bn);
  if (used <= 1 && tiles < 128 && K >= TK) {
    // one K chunk and under half the CUs busy: quarter-size tiles, C written
    // directly (no partial sums)
    dim3 g32(unsigned((M + TS - 1) / TS), unsigned((N + TS - 1) / TS));
    hipLaunchKernelGGL(k_gemm_s32<true>, g32, dim3(256), 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    return;
  }
  const int wgs = tiles * std::max(used, 1);
  const dim3 grid(unsigned((wgs + kXcds - 1) / kXcds * kXcds));
  if (bn == 128)
    hipLaunchKernelGGL(k_gemm_tn_part<128>, grid, dim3(256), 0, s, M, N, K, chunk, tiles_m, tiles, wgs, A, lda, B,
                       ldb, work);
  else
    hipLaunchKernelGGL(k_gemm_tn_part<64>, grid, dim3(256), 0, s, M, N, K, chunk, tiles_m, tiles, wgs, A, lda, B,
                       ldb, work);
  const int64_t total = M * N;
  hipLaunchKernelGGL(k_gemm_tn_reduce, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s, M, N,
                     std::max(used, 1), alpha, work, beta, C, ldc);
}

}  // namespace bd

// ---------------------------------------------------------------------------
// C ABI: the GEMMs as stand-alone entry points (row a28 check and tooling).
// ---------------------------------------------------------------------------
extern "C" int64_t bd_gemm_workspace(int32_t mode, int32_t trans_a, int64_t M, int64_t N, int64_t K) {
  if (!trans_a) return 0;
  return mode == 1 ? bd::gemm_tn_bf16x3_work(M, N, K) : bd::gemm_tn_work(M, N, K);
}

extern "C" bd_status bd_gemm_dev(int32_t mode, int32_t trans_a, int64_t M, int64_t N, int64_t K, double alpha,
                                 const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                                 double* C, int64_t ldc, double* work) {
  BD_GUARD({
    if (M < 0 || N < 0 || K < 0 || (mode != 0 && mode != 1)) bd::fail(BD_ERR_INVALID_PARAMETER, "gemm shape/mode");
    if (trans_a && !work && M * N > 0) bd::fail(BD_ERR_INVALID_PARAMETER, "gemm workspace");
    bd::Gemm g;
    g.bf16x3 = mode == 1;
    if (trans_a) g.tn(nullptr, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, work);
    else g.nn(nullptr, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    if (hipDeviceSynchronize() != hipSuccess || hipGetLastError() != hipSuccess)
      bd::fail(BD_ERR_DEVICE, "gemm launch failed");
  });
}

// The same product `reps` times back to back on one stream between two HIP
// events (no host round trip in between): the average milliseconds per
// product in *ms. For measuring the kernels apart from the call overhead.
#define GEMM_HIP(call)                                                               \
  do {                                                                               \
    const hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) bd::fail(BD_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

extern "C" bd_status bd_gemm_time_dev(int32_t mode, int32_t trans_a, int64_t M, int64_t N, int64_t K,
                                      const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                                      int64_t ldc, double* work, int32_t reps, double* ms) {
  BD_GUARD({
    if (M < 0 || N < 0 || K < 0 || (mode != 0 && mode != 1) || reps <= 0 || !ms)
      bd::fail(BD_ERR_INVALID_PARAMETER, "gemm timing shape/mode");
    if (trans_a && !work && M * N > 0) bd::fail(BD_ERR_INVALID_PARAMETER, "gemm workspace");
    bd::Gemm g;
    g.bf16x3 = mode == 1;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    GEMM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    GEMM_HIP(hipEventCreate(&e0));
    GEMM_HIP(hipEventCreate(&e1));
    auto run = [&] {
      if (trans_a) g.tn(s, M, N, K, 1.0, A, lda, B, ldb, 0.0, C, ldc, work);
      else g.nn(s, M, N, K, 1.0, A, lda, B, ldb, 0.0, C, ldc);
    };
    run();
    GEMM_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) run();
    GEMM_HIP(hipEventRecord(e1, s));
    GEMM_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    GEMM_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = double(t) / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
  });
}

extern "C" bd_status bd_mfma_f64_probe_dev(const double* a, const double* b, double* d) {
  BD_GUARD({
    if (!a || !b || !d) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    hipLaunchKernelGGL(bd::k_mfma_f64_probe, dim3(1), dim3(64), 0, nullptr, a, b, d);
    if (hipDeviceSynchronize() != hipSuccess) bd::fail(BD_ERR_DEVICE, "probe failed");
  });
}

extern "C" bd_status bd_mfma_probe_dev(const uint16_t* a_bits, const uint16_t* b_bits, float* d) {
  BD_GUARD({
    bd::mfma_probe(nullptr, a_bits, b_bits, d);
    if (hipDeviceSynchronize() != hipSuccess) bd::fail(BD_ERR_DEVICE, "probe failed");
  });
}
