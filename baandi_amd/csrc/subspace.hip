// Bordered inverse Cholesky of the Davidson overlap and the small matrix
// helpers of the subspace step, on the GPU in FP64.
//
// extend_inverse_cholesky_dev follows crates/dft-solvers/src/lean_davidson.rs:
// 167-331: with T11 = L11^{-1} known for the leading `from` columns of S,
//   W = T11 S12,  G = S22 - W^T W = L22 L22^T,  T22 = L22^{-1},
//   T21 = -T22 W^T T11,
// rejecting when a squared pivot over its overlap diagonal falls below
// eps/1e-12 (or is NaN), or when a new row of T fails the growth test
// eps * (sum_j |T_ij| sqrt(S_jj))^2 <= sqrt(eps).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>

#include "common.hpp"
#include "eigen.hpp"
#include "gemm.hpp"

namespace bd {

namespace {
constexpr int kCT = 1024;
constexpr int kInvRows = 2;  // k_chol_inv_lds: rows per lane of the substitution (k <= 128)
constexpr double kPivotRatioMin = DBL_EPSILON / 1e-12;  // lean_davidson.rs:62-64

// Lane l's value of v (l uniform across the wave).
__device__ __forceinline__ double read_lane(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// G (k x k, ld k) <- S[from:to, from:to].
__global__ void k_copy_block(int k, const double* __restrict__ S, int64_t lds, double* __restrict__ G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < k) G[i + int64_t(j) * k] = S[i + int64_t(j) * lds];
}

// Cholesky of the k x k block G (k <= 64 * kInvRows) with the pivot-ratio
// test, then T22 = L^{-1} into T (ld ldt), zero above the diagonal. The block
// sits in LDS as a packed lower triangle (column-major): an LDL^T elimination
// with one barrier per column (column j is only read while the columns right
// of it are updated), then L = column j over sqrt(d_j), then L^{-1} by
// forward substitution, a wave per column with the right-hand side in
// registers. A flag already set by an earlier block of the same border
// leaves the block untouched; min_ratio keeps the minimum over the blocks.
__global__ __launch_bounds__(kCT) void k_chol_inv_lds(int k, const double* __restrict__ G,
                                                      const double* __restrict__ sdiag, int64_t sstride,
                                                      double* __restrict__ T22, int64_t ldt, int* __restrict__ flag,
                                                      double* __restrict__ min_ratio) {
  extern __shared__ double sh[];
  if (flag[0]) return;  // an earlier block of this border failed: uniform exit
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int nw = kCT / 64;
  double* P = sh;
  double* rhsall = sh + int64_t(k) * (k + 1) / 2;
  auto pk = [k](int i, int l) { return l * k - (l * (l - 1)) / 2 + (i - l); };
  for (int l = wave; l < k; l += nw)
    for (int i = l + lane; i < k; i += 64) P[pk(i, l)] = G[i + int64_t(l) * k];
  __syncthreads();
  double mr = INFINITY;
  for (int j = 0; j < k; ++j) {
    const double dj = P[pk(j, j)];
    const double ratio = dj / sdiag[int64_t(j) * sstride];
    if (isnan(ratio) || ratio < kPivotRatioMin) {
      if (tid == 0) {
        flag[0] = 1;
        min_ratio[0] = isnan(ratio) ? 0.0 : fmin(min_ratio[0], fmin(mr, ratio));
      }
      return;  // every thread read the same pivot: uniform exit
    }
    mr = fmin(mr, ratio);
    const double inv_d = 1.0 / dj;
    for (int l = j + 1 + wave; l < k; l += nw) {
      const double alj = P[pk(l, j)] * inv_d;
      const int base_j = pk(0, j), base_l = pk(0, l);  // column starts (row index offsets follow)
      for (int i = l + lane; i < k; i += 64) P[base_l + i] -= P[base_j + i] * alj;
    }
    __syncthreads();
  }
  for (int l = wave; l < k; l += nw) {
    const double sd = sqrt(P[pk(l, l)]);
    for (int i = l + 1 + lane; i < k; i += 64) P[pk(i, l)] /= sd;
  }
  __syncthreads();
  for (int l = tid; l < k; l += kCT) {
    const double sd = sqrt(P[pk(l, l)]);
    P[pk(l, l)] = sd;
    rhsall[l] = 1.0 / sd;  // reciprocal diagonal for the substitution
  }
  __syncthreads();
  if (tid == 0) min_ratio[0] = fmin(min_ratio[0], mr);
  // Column c of L^{-1}: forward substitution L x = e_c, one wave per column,
  // the right-hand side in registers (lane owns rows lane + 64 t, t <
  // kInvRows), x_l broadcast from its owner lane: the serial chain per row is
  // a lane read, a multiply and the update, with no LDS round trip.
  const double* rd = rhsall;
  for (int c = __builtin_amdgcn_readfirstlane(wave); c < k; c += nw) {
    double r[kInvRows];
#pragma unroll
    for (int t = 0; t < kInvRows; ++t) r[t] = lane + 64 * t == c ? 1.0 : 0.0;
    for (int l = c; l < k; ++l) {
      const int tl = l >> 6, ll = l & 63;
      double own = r[0];
#pragma unroll
      for (int t = 1; t < kInvRows; ++t) own = tl == t ? r[t] : own;
      const double x = read_lane(own, ll) * rd[l];
      const int base_l = pk(0, l);  // P[base_l + i] is in bounds for every i < k
#pragma unroll
      for (int t = 0; t < kInvRows; ++t) {
        const int i = lane + 64 * t;
        const double upd = r[t] - P[base_l + min(i, k - 1)] * x;
        r[t] = i == l ? x : ((i > l && i < k) ? upd : r[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < kInvRows; ++t) {
      const int i = lane + 64 * t;
      if (i < k) T22[i + int64_t(c) * ldt] = i >= c ? r[t] : 0.0;
    }
  }
}

// One wave per new row i of T: finiteness and the growth guard.
__global__ __launch_bounds__(256) void k_growth_guard(const double* __restrict__ S, const double* __restrict__ T,
                                                      int cap, int from, int to, int guard, int* __restrict__ flag) {
  const int lane = threadIdx.x & 63;
  const int i = from + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= to) return;
  double g = 0.0;
  int bad = 0;
  for (int j = lane; j <= i; j += 64) {
    const double v = T[i + int64_t(j) * cap];
    bad |= !isfinite(v);
    g += fabs(v) * sqrt(S[j + int64_t(j) * cap]);
  }
  g = wave_sum(g);
  if (lane == 0 && guard && (bad || !isfinite(g) || DBL_EPSILON * g * g > sqrt(DBL_EPSILON))) flag[0] = 1;
  if (bad) flag[0] = 1;
}

__global__ void k_mirror(double* __restrict__ M, int cap, int from, int to) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = from + blockIdx.y;
  if (j < to && i < j) M[j + int64_t(i) * cap] = M[i + int64_t(j) * cap];
}

__global__ void k_transpose(int m, int n, const double* __restrict__ A, int64_t lda, double* __restrict__ B,
                            int64_t ldb) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
  for (int r = ty; r < 32; r += 8) {
    const int i = bx + tx, j = by + r;
    if (i < m && j < n) tile[r][tx] = A[i + int64_t(j) * lda];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int j = by + tx, i = bx + r;
    if (i < m && j < n) B[j + int64_t(i) * ldb] = tile[tx][r];
  }
}

__global__ void k_symmetrize(int m, double* __restrict__ A, int64_t lda) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < j && j < m) {
    const double v = 0.5 * (A[i + int64_t(j) * lda] + A[j + int64_t(i) * lda]);
    A[i + int64_t(j) * lda] = v;
    A[j + int64_t(i) * lda] = v;
  }
}

__global__ void k_identity(int m, int n, double* __restrict__ A, int64_t lda) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < m && j < n) A[i + int64_t(j) * lda] = i == j ? 1.0 : 0.0;
}

__global__ void k_zero_flag(int* flag, double* min_ratio) {
  flag[0] = 0;
  min_ratio[0] = INFINITY;
}
}  // namespace

// The split-K GEMMs above use at most max(1, K/256) chunks of an M x N tile
// with M, N, K <= cap.
int64_t factor_gwork(int cap) { return int64_t(cap < 256 ? 1 : cap / 256) * cap * cap; }

namespace {
// Widest border block k_chol_inv_lds holds: its rows live in the kInvRows
// registers of a lane, and k(k+1)/2 + k doubles fit the 160 KB of LDS.
constexpr int kFactorBlock = 64 * kInvRows;
// Widest block a border is split into: the elimination costs ~k^3 / 3 in
// barriered column steps, so narrower blocks plus the extension GEMMs cost less.
constexpr int kFactorSplit = 96;

// T[from:to, 0:to] for one block of at most kFactorBlock new columns:
// W = T11 S12, G = S22 - W^T W, T22 = chol(G)^{-1}, T21 = -T22 W^T T11.
void extend_block(hipStream_t s, const double* S, double* T, int cap, int from, int to, const FactorWork& fw) {
  const int k = to - from;
  if (from > 0) gemm_nn(s, from, k, from, 1.0, T, cap, S + int64_t(from) * cap, cap, 0.0, fw.W, from);
  hipLaunchKernelGGL(k_copy_block, dim3((k + 255) / 256, k), dim3(256), 0, s, k,
                     S + from + int64_t(from) * cap, int64_t(cap), fw.G);
  if (from > 0) gemm_tn(s, k, k, from, -1.0, fw.W, from, fw.W, from, 1.0, fw.G, k, fw.gwork);
  double* T22 = T + from + int64_t(from) * cap;
  const size_t lds = (size_t(k) * (k + 1) / 2 + size_t(k)) * 8;
  if (k > kFactorBlock || lds > 160 * 1024) fail(BD_ERR_INVALID_PARAMETER, "inverse Cholesky block too large");
  if (lds > 64 * 1024) allow_full_lds(reinterpret_cast<const void*>(&k_chol_inv_lds), "k_chol_inv_lds");
  hipLaunchKernelGGL(k_chol_inv_lds, dim3(1), dim3(kCT), lds, s, k, fw.G, S + from + int64_t(from) * cap,
                     int64_t(cap) + 1, T22, int64_t(cap), fw.flag, fw.min_ratio);
  if (from > 0) {
    gemm_tn(s, k, from, from, 1.0, fw.W, from, T, cap, 0.0, fw.U, k, fw.gwork);         // U = W^T T11
    gemm_nn(s, k, from, k, -1.0, T22, cap, fw.U, k, 0.0, T + from, cap);                  // T21 = -T22 U
  }
  hipLaunchKernelGGL(k_growth_guard, dim3((k + 3) / 4), dim3(256), 0, s, S, T, cap, from, to, fw.growth_guard,
                     fw.flag);
}
}  // namespace

// A border wider than one block is factored as successive extensions in
// near-equal blocks: the inverse Cholesky factor of S[0:b, 0:b] extended by
// the next columns of S is that of S[0:b', 0:b'] (the pivot of each column is
// the same Schur complement, so the pivot-ratio test is unchanged).
void extend_inverse_cholesky_dev(hipStream_t s, const double* S, double* T, int cap, int from, int to,
                                 const FactorWork& fw) {
  const int k = to - from;
  hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(1), 0, s, fw.flag, fw.min_ratio);
  if (k <= 0) return;
  const int blocks = (k + kFactorSplit - 1) / kFactorSplit;
  const int width = (k + blocks - 1) / blocks;
  for (int a = from; a < to; a += width) extend_block(s, S, T, cap, a, std::min(to, a + width), fw);
}

void mirror_border(hipStream_t s, double* M, int cap, int from, int to) {
  if (to <= from) return;
  hipLaunchKernelGGL(k_mirror, dim3((to + 255) / 256, to - from), dim3(256), 0, s, M, cap, from, to);
}

void transpose(hipStream_t s, int m, int n, const double* A, int64_t lda, double* B, int64_t ldb) {
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL(k_transpose, dim3((m + 31) / 32, (n + 31) / 32), dim3(256), 0, s, m, n, A, lda, B, ldb);
}

void symmetrize(hipStream_t s, int m, double* A, int64_t lda) {
  if (m <= 1) return;
  hipLaunchKernelGGL(k_symmetrize, dim3((m + 255) / 256, m), dim3(256), 0, s, m, A, lda);
}

void set_identity(hipStream_t s, int m, int n, double* A, int64_t lda) {
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL(k_identity, dim3((m + 255) / 256, n), dim3(256), 0, s, m, n, A, lda);
}

}  // namespace bd

namespace bd {
namespace {
// Factor scratch of one call at capacity cap: W, G, U (cap x cap each), the
// split-K partials, then min_ratio and the flag.
constexpr int64_t factor_scratch_doubles(int cap) {
  return 3 * int64_t(cap) * cap + int64_t(cap < 256 ? 1 : cap / 256) * cap * cap + 2;
}

FactorWork factor_layout(double* base, int cap) {
  const int64_t c2 = int64_t(cap) * cap;
  FactorWork fw;
  fw.W = base;
  fw.G = fw.W + c2;
  fw.U = fw.G + c2;
  fw.gwork = fw.U + c2;
  fw.min_ratio = fw.gwork + factor_gwork(cap);
  fw.flag = reinterpret_cast<int*>(fw.min_ratio + 1);
  return fw;
}

// One border extension on the null stream with its scratch at `base`; the
// verdict and the smallest pivot ratio are read back (synchronous).
void factor_on_null_stream(const double* S, double* T, int cap, int from, int to, double* base,
                           int32_t* accepted, double* min_ratio) {
  const FactorWork fw = factor_layout(base, cap);
  extend_inverse_cholesky_dev(nullptr, S, T, cap, from, to, fw);
  int flag = 0;
  if (hipMemcpy(&flag, fw.flag, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(min_ratio, fw.min_ratio, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    fail(BD_ERR_DEVICE, "inverse Cholesky launch failed");
  *accepted = flag == 0 ? 1 : 0;
}

// Allocation-free variant: the scratch of every call on a GPU is this
// file-scope device array (each GPU holds its own copy until the process
// ends). Its host-side name is not a device address: kernels get the address
// from hipGetSymbolAddress, taken once per device with that device current.
// Calls that overlap on one GPU share the array and race.
constexpr int kStaticFactorCap = 640;
constexpr int64_t kStaticFactorDoubles = factor_scratch_doubles(kStaticFactorCap);
static __device__ __align__(256) double scratch[kStaticFactorDoubles];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double* scratch_dev[kMaxDevices] = {};

double* static_factor_scratch() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) fail(BD_ERR_DEVICE, "no current HIP device");
  if (dev < 0 || dev >= kMaxDevices) fail(BD_ERR_DEVICE, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      fail(BD_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    scratch_dev[dev] = static_cast<double*>(ptr);
  }
  return scratch_dev[dev];
}
}  // namespace
}  // namespace bd

extern "C" bd_status bd_inverse_cholesky_dev(const double* S, double* T, int32_t cap, int32_t from, int32_t to,
                                             int32_t* accepted, double* min_ratio) {
  struct Scratch {
    double* p = nullptr;
    ~Scratch() {
      if (p) (void)hipFree(p);
    }
  } buf;
  BD_GUARD({
    if (!S || !T || !accepted || !min_ratio || cap <= 0 || from < 0 || to < from || to > cap)
      bd::fail(BD_ERR_INVALID_PARAMETER, "inverse Cholesky arguments");
    const int64_t total = bd::factor_scratch_doubles(cap);
    if (hipMalloc(&buf.p, size_t(total) * sizeof(double)) != hipSuccess) bd::fail(BD_ERR_DEVICE, "scratch");
    bd::factor_on_null_stream(S, T, cap, from, to, buf.p, accepted, min_ratio);
  });
}

extern "C" bd_status bd_inverse_cholesky_static_dev(const double* S, double* T, int32_t cap, int32_t from,
                                                    int32_t to, int32_t* accepted, double* min_ratio) {
  BD_GUARD({
    if (!S || !T || !accepted || !min_ratio || cap <= 0 || from < 0 || to < from || to > cap)
      bd::fail(BD_ERR_INVALID_PARAMETER, "inverse Cholesky arguments");
    if (cap > bd::kStaticFactorCap)
      bd::fail(BD_ERR_INVALID_PARAMETER, "bd_inverse_cholesky_static_dev: cap " + std::to_string(cap) +
                                             " exceeds the static scratch limit of " +
                                             std::to_string(bd::kStaticFactorCap));
    double* base = bd::static_factor_scratch();
    bd::factor_on_null_stream(S, T, cap, from, to, base, accepted, min_ratio);
  });
}
