// Reference block Davidson (crates/dft-solvers/src/ref_davidson.rs:117-366) on
// the GPU operator. Control flow runs on the host; every matrix lives on the
// GPU: the n-row operations (H·X, border Gram GEMMs, rotations, corrections,
// normalisation, rank repair projections) and the FP64 subspace step (the
// bordered inverse Cholesky of lean_davidson.rs:167-331, A = T H T^T, the
// reduced eigensolve and C = T^T Y) run on the op's stream. Per iteration the
// host reads back only the Ritz values, the correction norms and the factor
// verdict.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "common.hpp"
#include "eigen.hpp"
#include "mlp.hpp"
#include "gemm.hpp"
#include "op.hpp"

namespace bd {

namespace {
constexpr int kT = 256;
constexpr double kEps = std::numeric_limits<double>::epsilon();

__device__ __forceinline__ double block_sum(double v, double* red) {
  // Deterministic tree over the workgroup (fixed shape, fixed order).
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < int(blockDim.x >> 6); ++w) s += red[w];
  return s;  // valid on thread 0
}

// QE smooth denominator in Ha (ref_davidson/stream.rs:16-22).
__device__ __forceinline__ double sqrt_denominator(double h, double e) { return mlp::qe_denominator(h, e); }

// Teter-Payne-Allan K(T/<T>) (crates/dft-solvers/src/lib.rs:197-207).
__device__ __forceinline__ double tpa_weight(double kinetic, double scale) {
  const double x = kinetic / scale;
  if (x <= 1.0) {
    const double p = 27.0 + x * (18.0 + x * (12.0 + x * 8.0));
    return p / (p + 16.0 * x * x * x * x);
  }
  const double y = scale / kinetic;
  const double p = y * (8.0 + y * (12.0 + y * (18.0 + y * 27.0)));
  return p / (p + 16.0);
}

// One workgroup per active column j: r = hx - theta x, e = r / d (QE) or
// -r K (TPA); writes e and |e|^2 (ref_davidson/stream.rs:34-67).
__global__ __launch_bounds__(kT) void k_corrections(const double* __restrict__ X, const double* __restrict__ HX,
                                                    int64_t n, const double* __restrict__ theta,
                                                    const double* __restrict__ diag,
                                                    const double* __restrict__ kinetic, int mode,
                                                    double* __restrict__ E, double* __restrict__ norm2) {
  __shared__ double red[16];
  __shared__ double bcast;
  const int j = blockIdx.x;
  const double* x = X + int64_t(j) * n;
  const double* hx = HX + int64_t(j) * n;
  double* e = E + int64_t(j) * n;
  const double th = theta[j];
  double scale = 0.0;
  if (mode == 1) {
    double a = 0.0, b = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
      const double w = x[i] * x[i];
      a += w * kinetic[i];
      b += w;
    }
    const double sa = block_sum(a, red);
    const double sb = block_sum(b, red);
    if (threadIdx.x == 0) bcast = fmax(sa / sb, 0.1);  // kinetic_scale, lib.rs:178-193
    __syncthreads();
    scale = bcast;
  }
  // Four rows per thread and pass, their loads issued before any is used
  // (one workgroup per column: the loads' latency, not the arithmetic,
  // bounds the pass); the sum keeps the row order i, i + T, i + 2T, ...
  double acc = 0.0;
  const int64_t T = blockDim.x;
  int64_t i = threadIdx.x;
  for (; i + 3 * T < n; i += 4 * T) {
    double xv[4], hv[4], wv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xv[u] = x[i + u * T];
      hv[u] = hx[i + u * T];
      wv[u] = mode == 1 ? kinetic[i + u * T] : diag[i + u * T];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double r = hv[u] - th * xv[u];
      const double v = mode == 1 ? -r * tpa_weight(wv[u], scale) : r / sqrt_denominator(wv[u], th);
      e[i + u * T] = v;
      acc += v * v;
    }
  }
  for (; i < n; i += T) {
    const double r = hx[i] - th * x[i];
    const double v = mode == 1 ? -r * tpa_weight(kinetic[i], scale) : r / sqrt_denominator(diag[i], th);
    e[i] = v;
    acc += v * v;
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) norm2[j] = s;
}

// dst[:, j] = src[:, idx[j]] * scale[j].
__global__ void k_gather_scale(const double* __restrict__ src, int64_t n, const int* __restrict__ idx,
                               const double* __restrict__ scale, double* __restrict__ dst) {
  const int j = blockIdx.y;
  const double* s = src + int64_t(idx[j]) * n;
  double* d = dst + int64_t(j) * n;
  const double c = scale ? scale[j] : 1.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    d[i] = s[i] * c;
}

__global__ void k_finite(const double* __restrict__ p, int64_t count, int* __restrict__ bad) {
  int local = 0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
    local |= !isfinite(p[i]);
  if (local) *bad = 1;
}

__global__ void k_zero_int(int* p) { *p = 0; }

// |V[:, j]|^2 per column, one workgroup per column (fixed order).
__global__ __launch_bounds__(kT) void k_norm2_cols(const double* __restrict__ V, int64_t n, double* __restrict__ out) {
  __shared__ double red[16];
  const double* v = V + int64_t(blockIdx.x) * n;
  double acc = 0.0;
  const int64_t T = blockDim.x;
  int64_t i = threadIdx.x;
  for (; i + 3 * T < n; i += 4 * T) {  // loads first, sum in row order
    double a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = v[i + u * T];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += a[u] * a[u];
  }
  for (; i < n; i += T) acc += v[i] * v[i];
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// M[0:k, 0:k] = diag(values) (ld ldm).
__global__ void k_set_diag(int k, const double* __restrict__ values, double* __restrict__ M, int64_t ldm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < k) M[i + int64_t(j) * ldm] = i == j ? values[i] : 0.0;
}

// max |S[i, j] - delta_ij| over the leading k x k block (one workgroup; a
// non-finite entry gives +inf).
__global__ __launch_bounds__(kT) void k_identity_defect(int k, const double* __restrict__ S, int64_t lds,
                                                        double* __restrict__ out) {
  __shared__ double red[16];
  double mx = 0.0;
  for (int64_t t = threadIdx.x; t < int64_t(k) * k; t += blockDim.x) {
    const int i = int(t % k), j = int(t / k);
    const double v = S[i + int64_t(j) * lds] - (i == j ? 1.0 : 0.0);
    mx = isfinite(v) ? fmax(mx, fabs(v)) : INFINITY;
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < int(blockDim.x >> 6); ++w) r = fmax(r, red[w]);
    out[0] = r;
  }
}

// Restart guard of ref_davidson/ideal_reset.rs: column norms within this of 1
// allow the ideal reset; a measured |S - I| above it disables it.
constexpr double kOrthogonalityLimit = 1e-12;
// Every REFRESH_INTERVAL-th contraction is a measured refresh.
constexpr int kRefreshInterval = 4;

struct Dev {
  double* p = nullptr;
  void alloc(size_t count) { BD_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(double))); }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
};
}  // namespace

// Davidson workspace cached on the op per (n, bands, cap), like the reference's
// per-thread row cache (ref_davidson/workspace.rs:21-31).
struct DavidsonWork {
  int64_t n = 0;
  int nb = 0, cap = 0;
  Dev V, HV, X, HX, E, E2, small1, small2, small3, gwork, theta, norm2, rnorm2, scale, mlpw;
  // Subspace state (cap x cap, ld cap): overlap S, projected H, T = L^{-1}.
  Dev S, H, T, Tt, A, C, Y, vals, eigw, fW, fG, fU, fgwork, fmin;
  int* idx = nullptr;
  int* flag = nullptr;
  int* info = nullptr;  // [0] eigensolver, [1] factor verdict
  FactorWork fw;
  DavidsonWork(int64_t n_, int nb_, int cap_) : n(n_), nb(nb_), cap(cap_) {
    const size_t c2 = size_t(cap) * cap;
    S.alloc(c2);
    H.alloc(c2);
    T.alloc(c2);
    Tt.alloc(c2);
    A.alloc(c2);
    C.alloc(size_t(cap) * nb);
    Y.alloc(size_t(cap) * nb);
    vals.alloc(nb);
    eigw.alloc(size_t(sym_eigen_work(cap, nb)));
    fW.alloc(c2);
    fG.alloc(c2);
    fU.alloc(c2);
    fgwork.alloc(size_t(factor_gwork(cap)));
    fmin.alloc(1);
    BD_HIP(hipMalloc(&info, 2 * sizeof(int)));
    fw.W = fW.p;
    fw.G = fG.p;
    fw.U = fU.p;
    fw.gwork = fgwork.p;
    fw.flag = info + 1;
    fw.min_ratio = fmin.p;
    V.alloc(size_t(n) * cap);
    HV.alloc(size_t(n) * cap);
    X.alloc(size_t(n) * cap);
    HX.alloc(size_t(n) * cap);
    E.alloc(size_t(n) * nb);
    E2.alloc(size_t(n) * nb);
    small1.alloc(size_t(cap) * cap);
    small2.alloc(size_t(cap) * cap);
    small3.alloc(size_t(cap) * cap);
    gwork.alloc(size_t(std::max(Gemm::tn_work(cap, cap, n), int64_t(cap) * cap)));
    theta.alloc(nb);
    norm2.alloc(nb);
    rnorm2.alloc(nb);
    mlpw.alloc(size_t(mlp::correct_work(n, nb)));
    scale.alloc(nb);
    BD_HIP(hipMalloc(&idx, sizeof(int) * std::max(nb, 1)));
    BD_HIP(hipMalloc(&flag, sizeof(int)));
  }
  ~DavidsonWork() {
    if (idx) (void)hipFree(idx);
    if (flag) (void)hipFree(flag);
    if (info) (void)hipFree(info);
  }
};

namespace {
struct Solver {
  GammaOp& op;
  DavidsonWork& w;
  hipStream_t s;
  int64_t n;
  int nb, cap;
  const bd_davidson_options& o;
  const double* diag;
  double* V;
  double* HV;
  double* X;
  double* HX;
  // Ritz values of the current subspace (host copy of the device values).
  std::vector<double> values, old;
  // Learned-preconditioner state: switched off after a regression; last
  // residual norm^2 per band (NaN before the first correction).
  bool mlp_off = false;
  int strikes = 0;
  int step_now = 0;
  std::vector<double> prev_r2;
  bd_davidson_record rec{};
  double t_apply = 0.0, t_dense = 0.0;
  // Restart state (workspace.rs:772-935): contractions of this solve, and
  // whether a guard trip has switched the ideal reset off for its rest.
  int contractions = 0;
  bool reset_disabled = false;
  hipEvent_t ev0, ev1;

  Solver(GammaOp& op_, DavidsonWork& w_, const bd_davidson_options& o_, const double* diag_)
      : op(op_), w(w_), s(op_.stream), n(w_.n), nb(w_.nb), cap(w_.cap), o(o_), diag(diag_) {
    V = w.V.p;
    HV = w.HV.p;
    X = w.X.p;
    HX = w.HX.p;
    values.assign(nb, 0.0);
    prev_r2.assign(nb, std::numeric_limits<double>::quiet_NaN());
    rec.mlp_fallback = -1;
    old.assign(nb, 0.0);
    BD_HIP(hipEventCreate(&ev0));
    BD_HIP(hipEventCreate(&ev1));
  }
  ~Solver() {
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
  }

  double band_threshold(int i) const {
    if (o.first_step || i < o.occupied) return o.threshold;
    return std::max(5.0 * o.threshold, 5e-6);
  }

  void sync() { BD_HIP(hipStreamSynchronize(s)); }
  // Host sources are often temporaries: every upload completes before return.
  void up_bytes(void* d, const void* h, size_t bytes) {
    BD_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
    sync();
  }
  void up(double* d, const double* h, size_t count) { up_bytes(d, h, count * 8); }
  void down(double* h, const double* d, size_t count) {
    BD_HIP(hipMemcpyAsync(h, d, count * 8, hipMemcpyDeviceToHost, s));
    sync();
  }

  // H applied to k columns at src into dst, with the finiteness agreement of
  // ref_davidson/workspace.rs:1013-1036.
  void apply(const double* src, int k, double* dst) {
    BD_HIP(hipEventRecord(ev0, s));
    op.apply(src, n, k, dst, n);
    BD_HIP(hipEventRecord(ev1, s));
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, w.flag);
    hipLaunchKernelGGL(k_finite, dim3(256), dim3(256), 0, s, dst, n * k, w.flag);
    int bad = 0;
    BD_HIP(hipMemcpyAsync(&bad, w.flag, sizeof(int), hipMemcpyDeviceToHost, s));
    sync();
    float ms = 0.f;
    BD_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    t_apply += ms * 1e-3;
    rec.h_calls += 1;
    rec.h_columns += k;
    if (bad) fail(BD_ERR_APPLICATION, "reference H application failed (non-finite output)");
  }

  // Border columns [from, to) of S (is_h false) or H_r, mirrored
  // (ref_davidson/workspace.rs:300-381).
  void border(int from, int to, bool is_h) {
    const int k = to - from;
    if (k <= 0) return;
    const double* B = (is_h ? HV : V) + int64_t(from) * n;
    double* M = is_h ? w.H.p : w.S.p;
    op.gemm.tn(s, to, k, n, 1.0, V, n, B, n, 0.0, M + int64_t(from) * cap, cap, w.gwork.p);
    mirror_border(s, M, cap, from, to);
  }

  bool factor(int from, int to) {
    auto t0 = std::chrono::steady_clock::now();
    extend_inverse_cholesky_dev(s, w.S.p, w.T.p, cap, from, to, w.fw);
    int verdict = 0;
    BD_HIP(hipMemcpyAsync(&verdict, w.fw.flag, sizeof(int), hipMemcpyDeviceToHost, s));
    sync();
    t_dense += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return verdict == 0;
  }

  // A = T H_r T^T, lowest nb eigenpairs, C = T^T Y (workspace.rs:435-499).
  // If the inverse iteration cannot separate the vectors of a tight cluster
  // (the exact pairs of a realified k-point matrix can land so), A is formed
  // again and reduced by the one-step tridiagonalisation, whose rounding
  // differs; only a second failure is an error.
  void eigen(int m) {
    auto t0 = std::chrono::steady_clock::now();
    int bad = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
      transpose(s, m, m, w.T.p, cap, w.Tt.p, m);
      gemm_nn(s, m, m, m, 1.0, w.H.p, cap, w.Tt.p, m, 0.0, w.small3.p, m);  // H_r T^T
      gemm_nn(s, m, m, m, 1.0, w.T.p, cap, w.small3.p, m, 0.0, w.A.p, m);   // A
      symmetrize(s, m, w.A.p, m);
      sym_eigen_lowest(s, m, nb, w.A.p, m, w.vals.p, w.Y.p, m, w.eigw.p, w.info, attempt == 0);
      gemm_nn(s, m, nb, m, 1.0, w.Tt.p, m, w.Y.p, m, 0.0, w.C.p, m);  // C
      BD_HIP(hipMemcpyAsync(values.data(), w.vals.p, size_t(nb) * 8, hipMemcpyDeviceToHost, s));
      BD_HIP(hipMemcpyAsync(&bad, w.info, sizeof(int), hipMemcpyDeviceToHost, s));
      sync();
      if (!bad) break;
      rec.eigen_retries += 1;
    }
    t_dense += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (bad) fail(BD_ERR_APPLICATION, "reference reduced eigensolve failed");
  }

  void refresh(int m) {
    border(0, m, false);
    border(0, m, true);
    if (!factor(0, m)) fail(BD_ERR_APPLICATION, "reference contracted basis is dependent");
  }

  // Column norms^2 of the nb restarted columns all within 1e-12 of 1
  // (reset_norm_guard, workspace.rs:895-918).
  bool norm_guard() {
    hipLaunchKernelGGL(k_norm2_cols, dim3(nb), dim3(kT), 0, s, V, n, w.norm2.p);
    std::vector<double> nr(nb);
    down(nr.data(), w.norm2.p, nb);
    for (double v : nr)
      if (!std::isfinite(v) || std::fabs(v - 1.0) > kOrthogonalityLimit) return false;
    return true;
  }

  // V <- V C (nb columns), HV <- HV C, then C = I and either the ideal reset
  // (S = I, H_r = diag(theta), T = I: the restarted columns are the Ritz
  // vectors) or a measured refresh (workspace.rs:772-935, ideal_reset.rs):
  // measured every kRefreshInterval-th contraction, and for the rest of the
  // solve once a guard trips (column norms off 1, or a measured |S - I| above
  // 1e-12).
  void contract(int m) {
    op.gemm.nn(s, n, nb, m, 1.0, V, n, w.C.p, m, 0.0, X, n);
    op.gemm.nn(s, n, nb, m, 1.0, HV, n, w.C.p, m, 0.0, HX, n);
    std::swap(V, X);
    std::swap(HV, HX);
    contractions += 1;
    bool ideal = false;
    if (reset_disabled) {
    } else if (contractions % kRefreshInterval == 0) {
      rec.periodic_refreshes += 1;
    } else if (norm_guard()) {
      ideal = true;
      rec.ideal_resets += 1;
    } else {
      reset_disabled = true;
      rec.guard_fallbacks += 1;
    }
    set_identity(s, nb, nb, w.C.p, nb);
    if (ideal) {
      auto t0 = std::chrono::steady_clock::now();
      set_identity(s, nb, nb, w.S.p, cap);
      set_identity(s, nb, nb, w.T.p, cap);
      hipLaunchKernelGGL(k_set_diag, dim3((nb + 255) / 256, nb), dim3(256), 0, s, nb, w.vals.p, w.H.p,
                         int64_t(cap));
      BD_HIP(hipGetLastError());
      t_dense += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    } else {
      refresh(nb);
      if (!reset_disabled) {
        hipLaunchKernelGGL(k_identity_defect, dim3(1), dim3(kT), 0, s, nb, w.S.p, int64_t(cap), w.norm2.p);
        double defect = 0.0;
        down(&defect, w.norm2.p, 1);
        if (!(defect <= kOrthogonalityLimit)) {
          reset_disabled = true;
          rec.guard_fallbacks += 1;
        }
      }
    }
    rec.restarts += 1;
  }

  // Corrections of the active bands into w.E; returns host |e|^2.
  std::vector<double> corrections(int m, const std::vector<int>& active, bool contracted) {
    const int k = int(active.size());
    std::vector<double> th(k);
    for (int j = 0; j < k; ++j) th[j] = values[active[j]];
    up(w.theta.p, th.data(), k);
    up_bytes(w.idx, active.data(), sizeof(int) * k);
    if (contracted) {
      hipLaunchKernelGGL(k_gather_scale, dim3(64, k), dim3(256), 0, s, V, n, w.idx, nullptr, X);
      hipLaunchKernelGGL(k_gather_scale, dim3(64, k), dim3(256), 0, s, HV, n, w.idx, nullptr, HX);
    } else {
      hipLaunchKernelGGL(k_gather_scale, dim3(1, k), dim3(256), 0, s, w.C.p, int64_t(m), w.idx, nullptr, w.small1.p);
      op.gemm.nn(s, n, k, m, 1.0, V, n, w.small1.p, m, 0.0, X, n);
      op.gemm.nn(s, n, k, m, 1.0, HV, n, w.small1.p, m, 0.0, HX, n);
    }
    std::vector<double> nrm(k);
    if (o.preconditioner == BD_PRECOND_MLP && !mlp_off) {
      mlp::correct(s, n, k, X, HX, w.theta.p, op.d_kinetic.p, diag, o.mlp_vbar, o.mlp_blob, w.E.p, w.norm2.p,
                   w.rnorm2.p, w.mlpw.p);
      std::vector<double> rn(k);
      BD_HIP(hipMemcpyAsync(rn.data(), w.rnorm2.p, size_t(k) * 8, hipMemcpyDeviceToHost, s));
      down(nrm.data(), w.norm2.p, k);
      bool finite = true;
      for (double v : nrm) finite = finite && std::isfinite(v);
      if (!finite || mlp_regressed(active, rn)) {
        // QE fallback for the rest of this solve (the network's K = 1); a
        // non-finite network correction is replaced in this iteration already.
        mlp_off = true;
        rec.mlp_fallback = step_now;
        if (finite) return nrm;
      } else {
        return nrm;
      }
    }
    const int mode = (o.preconditioner == BD_PRECOND_TPA) ? 1 : 0;
    hipLaunchKernelGGL(k_corrections, dim3(k), dim3(kT), 0, s, X, HX, n, w.theta.p, diag, op.d_kinetic.p, mode,
                       w.E.p, w.norm2.p);
    down(nrm.data(), w.norm2.p, k);
    return nrm;
  }

  // The fallback rule: the residual-reduction ratio over the columns active
  // in this and the previous correction step, sqrt(sum |r_new|^2 / sum
  // |r_old|^2), at or above 1 (or not finite) in two consecutive steps.
  bool mlp_regressed(const std::vector<int>& active, const std::vector<double>& rn) {
    double num = 0.0, den = 0.0;
    for (size_t j = 0; j < active.size(); ++j) {
      const int b = active[j];
      if (std::isfinite(prev_r2[b])) {
        num += rn[j];
        den += prev_r2[b];
      }
      prev_r2[b] = rn[j];
    }
    if (den <= 0.0) return false;
    const double ratio = std::sqrt(num / den);
    strikes = (!std::isfinite(ratio) || ratio >= 1.0) ? strikes + 1 : 0;
    return strikes >= 2;
  }

  // Drop |e| <= 32 eps, scale the rest into dst (stream.rs:81-143); returns kept.
  int normalize(const std::vector<double>& norm2, double* dst) {
    std::vector<int> idx;
    std::vector<double> sc;
    for (int j = 0; j < int(norm2.size()); ++j) {
      const double nrm = std::sqrt(norm2[j]);
      if (!std::isfinite(nrm)) fail(BD_ERR_APPLICATION, "nonfinite reference correction");
      if (nrm <= 32.0 * kEps) continue;
      idx.push_back(j);
      sc.push_back(1.0 / nrm);
    }
    const int kept = int(idx.size());
    if (kept) {
      up_bytes(w.idx, idx.data(), sizeof(int) * kept);
      up(w.scale.p, sc.data(), kept);
      hipLaunchKernelGGL(k_gather_scale, dim3(64, kept), dim3(256), 0, s, w.E.p, n, w.idx, w.scale.p, dst);
    }
    return kept;
  }

  // Rank repair (workspace.rs:586-771) on k normalised corrections in w.E2.
  int repair(int m, int k) {
    rec.repairs += 1;
    // Coordinates of the Ritz block in the whitened basis: C <- T S C.
    gemm_nn(s, m, nb, m, 1.0, w.S.p, cap, w.C.p, m, 0.0, w.small2.p, m);
    gemm_nn(s, m, nb, m, 1.0, w.T.p, cap, w.small2.p, m, 0.0, w.C.p, m);
    transpose(s, m, m, w.T.p, cap, w.small1.p, m);  // T^T
    op.gemm.nn(s, n, m, m, 1.0, V, n, w.small1.p, m, 0.0, X, n);
    op.gemm.nn(s, n, m, m, 1.0, HV, n, w.small1.p, m, 0.0, HX, n);
    std::swap(V, X);
    std::swap(HV, HX);
    refresh(m);
    double* E = w.E2.p;
    for (int pass = 0; pass < 2; ++pass) {
      op.gemm.tn(s, m, k, n, 1.0, V, n, E, n, 0.0, w.small1.p, m, w.gwork.p);
      op.gemm.nn(s, n, k, m, -1.0, V, n, w.small1.p, m, 1.0, E, n);
    }
    int kept = 0;
    for (int j = 0; j < k; ++j) {
      double* ej = E + int64_t(j) * n;
      for (int pass = 0; pass < 2 && kept > 0; ++pass) {
        gemm_tn(s, kept, 1, n, 1.0, E, n, ej, n, 0.0, w.small1.p, kept, w.gwork.p);
        gemm_nn(s, n, 1, kept, -1.0, E, n, w.small1.p, kept, 1.0, ej, n);
      }
      gemm_tn(s, 1, 1, n, 1.0, ej, n, ej, n, 0.0, w.small1.p, 1, w.gwork.p);
      double nn2 = 0.0;
      down(&nn2, w.small1.p, 1);
      const double nrm = std::sqrt(nn2);
      if (nrm <= 1e-12 || m + kept == cap) continue;  // DEPENDENCE_TOLERANCE
      const int jj = j;
      const double inv = 1.0 / nrm;
      up_bytes(w.idx, &jj, sizeof(int));
      up(w.scale.p, &inv, 1);
      hipLaunchKernelGGL(k_gather_scale, dim3(64, 1), dim3(256), 0, s, E, n, w.idx, w.scale.p,
                         E + int64_t(kept) * n);
      sync();
      ++kept;
    }
    if (kept > 0) {
      BD_HIP(hipMemcpyAsync(V + int64_t(m) * n, E, size_t(n) * kept * 8, hipMemcpyDeviceToDevice, s));
      border(m, m + kept, false);
      if (!factor(m, m + kept)) fail(BD_ERR_APPLICATION, "reference overlap remains dependent after repair");
    }
    return kept;
  }

  bd_status run(const double* x0, int64_t ldx0, const double* hx0, double* eig, double* xout, double* hxout,
                int64_t ldx, int32_t* conv) {
    auto t_start = std::chrono::steady_clock::now();
    zero_doubles(s, w.T.p, int64_t(cap) * cap);  // T stays lower triangular from here on
    BD_HIP(hipMemcpy2DAsync(V, n * 8, x0, ldx0 * 8, n * 8, nb, hipMemcpyDeviceToDevice, s));
    if (hx0) BD_HIP(hipMemcpy2DAsync(HV, n * 8, hx0, ldx0 * 8, n * 8, nb, hipMemcpyDeviceToDevice, s));
    else apply(V, nb, HV);
    border(0, nb, false);
    if (!factor(0, nb)) fail(BD_ERR_INVALID_PARAMETER, "reference Davidson start is numerically dependent");
    border(0, nb, true);
    eigen(nb);
    int m = nb;
    std::vector<int> active(nb);
    std::iota(active.begin(), active.end(), 0);
    std::vector<char> converged(nb, 0);
    bool contracted = false;
    int iterations = 0;
    for (int step = 0; step < o.max_iterations; ++step) {
      if (m + int(active.size()) > cap && m > nb) {
        contract(m);
        m = nb;
        contracted = true;
      }
      step_now = step;
      std::vector<double> nrm = corrections(m, active, contracted);
      contracted = false;
      int k = 0;
      bool staged = false;
      {
        int kk = 0;
        for (double v : nrm) kk += (std::sqrt(v) > 32.0 * kEps) ? 1 : 0;
        staged = m + kk <= cap && kk > 0;
        k = normalize(nrm, staged ? V + int64_t(m) * n : w.E2.p);
      }
      if (!staged || k == 0) {
        k = k > 0 ? repair(m, k) : 0;
      } else {
        border(m, m + k, false);
        if (!factor(m, m + k)) {
          BD_HIP(hipMemcpyAsync(w.E2.p, V + int64_t(m) * n, size_t(n) * k * 8, hipMemcpyDeviceToDevice, s));
          k = repair(m, k);
        }
      }
      std::copy(values.begin(), values.begin() + nb, old.begin());
      if (k > 0) {
        apply(V + int64_t(m) * n, k, HV + int64_t(m) * n);
        border(m, m + k, true);
        eigen(m + k);
        m += k;
      }
      active.clear();
      for (int i = 0; i < nb; ++i) {
        converged[i] = std::fabs(values[i] - old[i]) < band_threshold(i);
        if (!converged[i]) active.push_back(i);
      }
      if (o.profile) {
        // Solution::ritz_changes_ha / ritz_values_ha (ref_davidson.rs:86-89).
        for (int i = 0; i < nb; ++i) {
          const int64_t at = int64_t(step) * nb + i;
          if (o.ritz_changes_ha) o.ritz_changes_ha[at] = std::fabs(values[i] - old[i]);
          if (o.ritz_values_ha) {
            o.ritz_values_ha[2 * at] = old[i];
            o.ritz_values_ha[2 * at + 1] = values[i];
          }
        }
      }
      iterations = step + 1;
      if (active.empty() || iterations == o.max_iterations) break;
      if (m + int(active.size()) > cap) {
        contract(m);
        m = nb;
        contracted = true;
      }
    }
    // Return X = V C, HX = HV C (workspace.rs:936-980).
    op.gemm.nn(s, n, nb, m, 1.0, V, n, w.C.p, m, 0.0, xout, ldx);
    if (hxout) op.gemm.nn(s, n, nb, m, 1.0, HV, n, w.C.p, m, 0.0, hxout, ldx);
    sync();
    bool all = true;
    for (int i = 0; i < nb; ++i) {
      eig[i] = values[i];
      if (conv) conv[i] = converged[i];
      all = all && converged[i];
    }
    rec.iterations = iterations;
    rec.all_converged = all ? 1 : 0;
    rec.seconds_apply = t_apply;
    rec.seconds_dense = t_dense;
    rec.seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return all ? BD_OK : BD_NOT_CONVERGED;
  }
};
}  // namespace

// ---------------------------------------------------------------------------
// Complex reference Davidson (ref_davidson_complex.rs:641-1113) for k-point
// ops: the columns are complex, stored as interleaved (re, im) rows (the k
// op's own row layout, n = 2 n_k doubles per column), so H·X, the
// corrections and the normalisation are the real kernels on complex columns.
// The Hermitian Grams V^H W come from two real products over the
// interleaved rows (Re = V^T W, Im = -V^T (i W)); a rotation V C with
// complex C is one real product V [Re C, Im C] and a combination. The small
// subspace state (overlap, projected H, T = L^{-1}, the reduced matrix, the
// Ritz coefficients) is kept in its real 2x2-block form (z -> [[re, -im],
// [im, re]]), where the bordered inverse Cholesky and the products are the
// real ones; the reduced eigenproblem is solved as the complex Hermitian
// matrix it is (herm_eigen_lowest), so no eigenvalue is doubled.
// ---------------------------------------------------------------------------
namespace {
// dst[:, j] = i src[:, j] (interleaved complex columns of n doubles).
__global__ void k_times_i(const double* __restrict__ src, int64_t n, int k, double* __restrict__ dst) {
  const int j = blockIdx.y;
  const double2* a = reinterpret_cast<const double2*>(src + int64_t(j) * n);
  double2* b = reinterpret_cast<double2*>(dst + int64_t(j) * n);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n / 2; i += int64_t(gridDim.x) * blockDim.x) {
    const double2 v = a[i];
    b[i] = make_double2(-v.y, v.x);
  }
}

// Realified columns [2 from, 2 to) of M (ld ldm), rows [0, 2 to), from the
// complex Gram G_ij = P[i, j] - i P[to k + i, j] (P: to x k real parts, then
// to x k negated imaginary parts, ld to); a diagonal entry's imaginary part
// is zero (G is Hermitian).
__global__ void k_scatter_herm(const double* __restrict__ P, int to, int k, int from, double* __restrict__ M,
                               int64_t ldm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= to) return;
  const double gr = P[i + int64_t(j) * to];
  const double gi = (i == from + j) ? 0.0 : -P[int64_t(to) * k + i + int64_t(j) * to];
  double* c0 = M + int64_t(2 * (from + j)) * ldm;
  double* c1 = c0 + ldm;
  c0[2 * i] = gr;
  c0[2 * i + 1] = gi;
  c1[2 * i] = -gi;
  c1[2 * i + 1] = gr;
}

// Complex A (m x m, ld m) from its symmetric realified form AR (ld ldr).
__global__ void k_extract_herm(const double* __restrict__ AR, int64_t ldr, int m, double2* __restrict__ A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= m) return;
  const double* c0 = AR + int64_t(2 * j) * ldr;
  const double* c1 = c0 + ldr;
  const double re = 0.5 * (c0[2 * i] + c1[2 * i + 1]);
  const double im = i == j ? 0.0 : 0.5 * (c0[2 * i + 1] - c1[2 * i]);
  A[i + int64_t(j) * m] = make_double2(re, im);
}

// Realified YR (2m x 2k, ld ldr) of the complex Y (m x k, ld m).
__global__ void k_realify(const double2* __restrict__ Y, int m, int k, double* __restrict__ YR, int64_t ldr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= m) return;
  const double2 y = Y[i + int64_t(j) * m];
  double* c0 = YR + int64_t(2 * j) * ldr;
  double* c1 = c0 + ldr;
  c0[2 * i] = y.x;
  c0[2 * i + 1] = y.y;
  c1[2 * i] = -y.y;
  c1[2 * i + 1] = y.x;
}

// [Re C, Im C] (m x 2k, ld m) of the complex columns cols[j] (or j when cols
// is null) of the realified CR (ld ldr).
__global__ void k_coef_parts(const double* __restrict__ CR, int64_t ldr, int m, int k, const int* __restrict__ cols,
                             double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= m) return;
  const int c = cols ? cols[j] : j;
  const double* c0 = CR + int64_t(2 * c) * ldr;
  out[i + int64_t(j) * m] = c0[2 * i];
  out[i + int64_t(k + j) * m] = c0[2 * i + 1];
}

// p[0:count] = -p[0:count].
__global__ void k_negate(double* __restrict__ p, int64_t count) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) p[i] = -p[i];
}

// Y = beta Y + alpha (T1 + i T2) for interleaved complex columns (T12: n x 2k,
// ld n; the first k columns T1, then T2).
__global__ void k_combine(const double* __restrict__ T12, int64_t n, int k, double alpha, double beta,
                          double* __restrict__ Y, int64_t ldy) {
  const int j = blockIdx.y;
  const double2* t1 = reinterpret_cast<const double2*>(T12 + int64_t(j) * n);
  const double2* t2 = reinterpret_cast<const double2*>(T12 + int64_t(k + j) * n);
  double2* y = reinterpret_cast<double2*>(Y + int64_t(j) * ldy);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n / 2; i += int64_t(gridDim.x) * blockDim.x) {
    const double2 a = t1[i], b = t2[i];
    const double2 v = make_double2(alpha * (a.x - b.y), alpha * (a.y + b.x));
    y[i] = beta == 0.0 ? v : make_double2(beta * y[i].x + v.x, beta * y[i].y + v.y);
  }
}

// out[2 i] = out[2 i + 1] = v[i]: realified diagonal values.
__global__ void k_dup(const double* __restrict__ v, int k, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) {
    out[2 * i] = v[i];
    out[2 * i + 1] = v[i];
  }
}
}  // namespace

// Complex Davidson workspace: cap complex columns of n doubles for the
// n-row blocks; the small state in realified form (capR = 2 cap).
struct ComplexWork {
  int64_t n = 0;
  int nb = 0, cap = 0, capR = 0;
  Dev V, HV, X, HX, E, E2, W2, T12, P, gwork, theta, norm2, scale, coef;
  Dev S, H, T, Tt, A, C, YR, small1, small2, small3, vals, vals2, eigw, Ac, Yc, fW, fG, fU, fgwork, fmin;
  int* idx = nullptr;
  int* flag = nullptr;
  int* info = nullptr;
  FactorWork fw;
  ComplexWork(int64_t n_, int nb_, int cap_) : n(n_), nb(nb_), cap(cap_), capR(2 * cap_) {
    const size_t c2 = size_t(capR) * capR;
    V.alloc(size_t(n) * cap);
    HV.alloc(size_t(n) * cap);
    X.alloc(size_t(n) * cap);
    HX.alloc(size_t(n) * cap);
    E.alloc(size_t(n) * nb);
    E2.alloc(size_t(n) * nb);
    W2.alloc(size_t(n) * cap);
    T12.alloc(size_t(n) * 2 * cap);
    P.alloc(size_t(2) * cap * cap);
    gwork.alloc(size_t(std::max(Gemm::tn_work(cap, cap, n), int64_t(cap) * cap)));
    theta.alloc(nb);
    norm2.alloc(cap);
    scale.alloc(nb);
    coef.alloc(size_t(2) * cap * cap);
    S.alloc(c2);
    H.alloc(c2);
    T.alloc(c2);
    Tt.alloc(c2);
    A.alloc(c2);
    C.alloc(size_t(capR) * 2 * nb);
    YR.alloc(size_t(capR) * 2 * nb);
    small1.alloc(c2);
    small2.alloc(c2);
    small3.alloc(c2);
    vals.alloc(nb);
    vals2.alloc(2 * nb);
    eigw.alloc(size_t(sym_eigen_work(cap, nb)) + 1);
    Ac.alloc(size_t(2) * cap * cap);
    Yc.alloc(size_t(2) * cap * nb);
    fW.alloc(c2);
    fG.alloc(c2);
    fU.alloc(c2);
    fgwork.alloc(size_t(factor_gwork(capR)));
    fmin.alloc(1);
    BD_HIP(hipMalloc(&info, 2 * sizeof(int)));
    BD_HIP(hipMalloc(&idx, sizeof(int) * std::max(cap, 1)));
    BD_HIP(hipMalloc(&flag, sizeof(int)));
    fw.W = fW.p;
    fw.G = fG.p;
    fw.U = fU.p;
    fw.gwork = fgwork.p;
    fw.flag = info + 1;
    fw.min_ratio = fmin.p;
    fw.growth_guard = 0;
  }
  ~ComplexWork() {
    if (idx) (void)hipFree(idx);
    if (flag) (void)hipFree(flag);
    if (info) (void)hipFree(info);
  }
};

ComplexWork* complex_work(GammaOp& op, int64_t n, int nb, int cap) {
  if (!op.cdav || op.cdav->n != n || op.cdav->nb != nb || op.cdav->cap != cap) {
    delete op.cdav;
    op.cdav = nullptr;
    op.cdav = new ComplexWork(n, nb, cap);
  }
  return op.cdav;
}
void destroy_complex_work(ComplexWork* w) { delete w; }

namespace {
struct ComplexSolver {
  GammaOp& op;
  ComplexWork& w;
  hipStream_t s;
  int64_t n;
  int nb, cap, capR;
  const bd_davidson_options& o;
  const double* diag;
  double* V;
  double* HV;
  double* X;
  double* HX;
  std::vector<double> values, old;
  bd_davidson_record rec{};
  double t_apply = 0.0, t_dense = 0.0;
  int contractions = 0;
  bool reset_disabled = false;
  hipEvent_t ev0, ev1;

  ComplexSolver(GammaOp& op_, ComplexWork& w_, const bd_davidson_options& o_, const double* diag_)
      : op(op_), w(w_), s(op_.stream), n(w_.n), nb(w_.nb), cap(w_.cap), capR(w_.capR), o(o_), diag(diag_) {
    V = w.V.p;
    HV = w.HV.p;
    X = w.X.p;
    HX = w.HX.p;
    values.assign(nb, 0.0);
    old.assign(nb, 0.0);
    rec.mlp_fallback = -1;
    BD_HIP(hipEventCreate(&ev0));
    BD_HIP(hipEventCreate(&ev1));
  }
  ~ComplexSolver() {
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
  }

  double band_threshold(int i) const {
    if (o.first_step || i < o.occupied) return o.threshold;
    return std::max(5.0 * o.threshold, 5e-6);
  }
  void sync() { BD_HIP(hipStreamSynchronize(s)); }
  void up_bytes(void* d, const void* h, size_t bytes) {
    BD_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
    sync();
  }
  void down(double* h, const double* d, size_t count) {
    BD_HIP(hipMemcpyAsync(h, d, count * 8, hipMemcpyDeviceToHost, s));
    sync();
  }
  unsigned row_blocks() const { return unsigned(std::min<int64_t>((n / 2 + 255) / 256, 256)); }

  void apply(const double* src, int k, double* dst) {
    BD_HIP(hipEventRecord(ev0, s));
    op.apply(src, n, k, dst, n);
    BD_HIP(hipEventRecord(ev1, s));
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, w.flag);
    hipLaunchKernelGGL(k_finite, dim3(256), dim3(256), 0, s, dst, n * k, w.flag);
    int bad = 0;
    BD_HIP(hipMemcpyAsync(&bad, w.flag, sizeof(int), hipMemcpyDeviceToHost, s));
    sync();
    float ms = 0.f;
    BD_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    t_apply += ms * 1e-3;
    rec.h_calls += 1;
    rec.h_columns += k;
    if (bad) fail(BD_ERR_APPLICATION, "complex H application failed (non-finite output)");
  }

  // P = [Re(V[:, :rows]^H B), -Im(...)] (rows x k each, ld rows) for k
  // complex columns B.
  void gram(int rows, const double* B, int k, double* P) {
    hipLaunchKernelGGL(k_times_i, dim3(row_blocks(), k), dim3(256), 0, s, B, n, k, w.W2.p);
    op.gemm.tn(s, rows, k, n, 1.0, V, n, B, n, 0.0, P, rows, w.gwork.p);
    op.gemm.tn(s, rows, k, n, 1.0, V, n, w.W2.p, n, 0.0, P + int64_t(rows) * k, rows, w.gwork.p);
  }

  // Y (k complex columns, ld ldy) = alpha src[:, :m] C + beta Y, with C the
  // realified coefficients CR (ld ldr) at complex columns cols (or 0..k-1).
  void rotate(const double* src, int m, const double* CR, int64_t ldr, int k, const int* cols, double alpha,
              double beta, double* Y, int64_t ldy) {
    if (k <= 0) return;
    hipLaunchKernelGGL(k_coef_parts, dim3((m + 255) / 256, k), dim3(256), 0, s, CR, ldr, m, k, cols, w.coef.p);
    op.gemm.nn(s, n, 2 * k, m, 1.0, src, n, w.coef.p, m, 0.0, w.T12.p, n);
    hipLaunchKernelGGL(k_combine, dim3(row_blocks(), k), dim3(256), 0, s, w.T12.p, n, k, alpha, beta, Y, ldy);
  }

  // Border columns [from, to) of S (is_h false) or H_r, realified.
  void border(int from, int to, bool is_h) {
    const int k = to - from;
    if (k <= 0) return;
    const double* B = (is_h ? HV : V) + int64_t(from) * n;
    double* M = is_h ? w.H.p : w.S.p;
    gram(to, B, k, w.P.p);
    hipLaunchKernelGGL(k_scatter_herm, dim3((to + 255) / 256, k), dim3(256), 0, s, w.P.p, to, k, from, M,
                       int64_t(capR));
    mirror_border(s, M, capR, 2 * from, 2 * to);
  }

  bool factor(int from, int to) {
    auto t0 = std::chrono::steady_clock::now();
    extend_inverse_cholesky_dev(s, w.S.p, w.T.p, capR, 2 * from, 2 * to, w.fw);
    int verdict = 0;
    BD_HIP(hipMemcpyAsync(&verdict, w.fw.flag, sizeof(int), hipMemcpyDeviceToHost, s));
    sync();
    t_dense += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return verdict == 0;
  }

  // A = T H_r T^H (realified), the lowest nb eigenpairs of the complex A,
  // C = T^H Y (realified).
  void eigen(int m) {
    auto t0 = std::chrono::steady_clock::now();
    const int mR = 2 * m, nbR = 2 * nb;
    transpose(s, mR, mR, w.T.p, capR, w.Tt.p, mR);
    gemm_nn(s, mR, mR, mR, 1.0, w.H.p, capR, w.Tt.p, mR, 0.0, w.small3.p, mR);
    gemm_nn(s, mR, mR, mR, 1.0, w.T.p, capR, w.small3.p, mR, 0.0, w.A.p, mR);
    symmetrize(s, mR, w.A.p, mR);
    double2* Ac = reinterpret_cast<double2*>(w.Ac.p);
    hipLaunchKernelGGL(k_extract_herm, dim3((m + 255) / 256, m), dim3(256), 0, s, w.A.p, int64_t(mR), m, Ac);
    herm_eigen_lowest(s, m, nb, w.Ac.p, m, w.vals.p, w.Yc.p, m, w.eigw.p, w.info);
    hipLaunchKernelGGL(k_realify, dim3((m + 255) / 256, nb), dim3(256), 0, s,
                       reinterpret_cast<const double2*>(w.Yc.p), m, nb, w.YR.p, int64_t(mR));
    gemm_nn(s, mR, nbR, mR, 1.0, w.Tt.p, mR, w.YR.p, mR, 0.0, w.C.p, mR);
    int bad = 0;
    BD_HIP(hipMemcpyAsync(values.data(), w.vals.p, size_t(nb) * 8, hipMemcpyDeviceToHost, s));
    BD_HIP(hipMemcpyAsync(&bad, w.info, sizeof(int), hipMemcpyDeviceToHost, s));
    sync();
    t_dense += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (bad) fail(BD_ERR_APPLICATION, "complex reduced eigensolve failed");
  }

  void refresh(int m) {
    border(0, m, false);
    border(0, m, true);
    if (!factor(0, m)) fail(BD_ERR_APPLICATION, "complex contracted basis is dependent");
  }

  bool norm_guard() {
    hipLaunchKernelGGL(k_norm2_cols, dim3(nb), dim3(kT), 0, s, V, n, w.norm2.p);
    std::vector<double> nr(nb);
    down(nr.data(), w.norm2.p, nb);
    for (double v : nr)
      if (!std::isfinite(v) || std::fabs(v - 1.0) > kOrthogonalityLimit) return false;
    return true;
  }

  void contract(int m) {
    rotate(V, m, w.C.p, 2 * m, nb, nullptr, 1.0, 0.0, X, n);
    rotate(HV, m, w.C.p, 2 * m, nb, nullptr, 1.0, 0.0, HX, n);
    std::swap(V, X);
    std::swap(HV, HX);
    contractions += 1;
    bool ideal = false;
    if (reset_disabled) {
    } else if (contractions % kRefreshInterval == 0) {
      rec.periodic_refreshes += 1;
    } else if (norm_guard()) {
      ideal = true;
      rec.ideal_resets += 1;
    } else {
      reset_disabled = true;
      rec.guard_fallbacks += 1;
    }
    const int nbR = 2 * nb;
    set_identity(s, nbR, nbR, w.C.p, nbR);
    if (ideal) {
      set_identity(s, nbR, nbR, w.S.p, capR);
      set_identity(s, nbR, nbR, w.T.p, capR);
      hipLaunchKernelGGL(k_dup, dim3((nb + 255) / 256), dim3(256), 0, s, w.vals.p, nb, w.vals2.p);
      hipLaunchKernelGGL(k_set_diag, dim3((nbR + 255) / 256, nbR), dim3(256), 0, s, nbR, w.vals2.p, w.H.p,
                         int64_t(capR));
      BD_HIP(hipGetLastError());
    } else {
      refresh(nb);
      if (!reset_disabled) {
        hipLaunchKernelGGL(k_identity_defect, dim3(1), dim3(kT), 0, s, nbR, w.S.p, int64_t(capR), w.norm2.p);
        double defect = 0.0;
        down(&defect, w.norm2.p, 1);
        if (!(defect <= kOrthogonalityLimit)) {
          reset_disabled = true;
          rec.guard_fallbacks += 1;
        }
      }
    }
    rec.restarts += 1;
  }

  std::vector<double> corrections(int m, const std::vector<int>& active, bool contracted) {
    const int k = int(active.size());
    std::vector<double> th(k);
    for (int j = 0; j < k; ++j) th[j] = values[active[j]];
    up_bytes(w.theta.p, th.data(), size_t(k) * 8);
    up_bytes(w.idx, active.data(), sizeof(int) * k);
    if (contracted) {
      hipLaunchKernelGGL(k_gather_scale, dim3(64, k), dim3(256), 0, s, V, n, w.idx, nullptr, X);
      hipLaunchKernelGGL(k_gather_scale, dim3(64, k), dim3(256), 0, s, HV, n, w.idx, nullptr, HX);
    } else {
      rotate(V, m, w.C.p, 2 * m, k, w.idx, 1.0, 0.0, X, n);
      rotate(HV, m, w.C.p, 2 * m, k, w.idx, 1.0, 0.0, HX, n);
    }
    const int mode = (o.preconditioner == BD_PRECOND_QE) ? 0 : 1;
    hipLaunchKernelGGL(k_corrections, dim3(k), dim3(kT), 0, s, X, HX, n, w.theta.p, diag, op.d_kinetic.p, mode,
                       w.E.p, w.norm2.p);
    std::vector<double> nrm(k);
    down(nrm.data(), w.norm2.p, k);
    return nrm;
  }

  int normalize(const std::vector<double>& norm2, double* dst) {
    std::vector<int> idx;
    std::vector<double> sc;
    for (int j = 0; j < int(norm2.size()); ++j) {
      const double nrm = std::sqrt(norm2[j]);
      if (!std::isfinite(nrm)) fail(BD_ERR_APPLICATION, "nonfinite complex correction");
      if (nrm <= 32.0 * kEps) continue;
      idx.push_back(j);
      sc.push_back(1.0 / nrm);
    }
    const int kept = int(idx.size());
    if (kept) {
      up_bytes(w.idx, idx.data(), sizeof(int) * kept);
      up_bytes(w.scale.p, sc.data(), size_t(kept) * 8);
      hipLaunchKernelGGL(k_gather_scale, dim3(64, kept), dim3(256), 0, s, w.E.p, n, w.idx, w.scale.p, dst);
    }
    return kept;
  }

  // E[:, cols] -= V[:, :m] (V^H E) for k complex columns at E.
  void project_out(int m, double* E, int k) {
    gram(m, E, k, w.P.p);
    // [Re G, Im G] (m x 2k, ld m): the Gram's second half negated.
    hipLaunchKernelGGL(k_negate, dim3((m * k + 255) / 256), dim3(256), 0, s, w.P.p + int64_t(m) * k,
                       int64_t(m) * k);
    op.gemm.nn(s, n, 2 * k, m, 1.0, V, n, w.P.p, m, 0.0, w.T12.p, n);
    hipLaunchKernelGGL(k_combine, dim3(row_blocks(), k), dim3(256), 0, s, w.T12.p, n, k, -1.0, 1.0, E, n);
  }

  // Rank repair (workspace.rs:586-771) on k normalised corrections in w.E2.
  int repair(int m, int k) {
    rec.repairs += 1;
    const int mR = 2 * m, nbR = 2 * nb;
    gemm_nn(s, mR, nbR, mR, 1.0, w.S.p, capR, w.C.p, mR, 0.0, w.small2.p, mR);
    gemm_nn(s, mR, nbR, mR, 1.0, w.T.p, capR, w.small2.p, mR, 0.0, w.C.p, mR);
    transpose(s, mR, mR, w.T.p, capR, w.small1.p, mR);  // T^T (realified T^H)
    rotate(V, m, w.small1.p, mR, m, nullptr, 1.0, 0.0, X, n);
    rotate(HV, m, w.small1.p, mR, m, nullptr, 1.0, 0.0, HX, n);
    std::swap(V, X);
    std::swap(HV, HX);
    refresh(m);
    double* E = w.E2.p;
    for (int pass = 0; pass < 2; ++pass) project_out(m, E, k);
    int kept = 0;
    for (int j = 0; j < k; ++j) {
      double* ej = E + int64_t(j) * n;
      if (kept > 0) {
        // Against the corrections kept so far (complex, two passes).
        double* Vsave = V;
        V = E;
        for (int pass = 0; pass < 2; ++pass) project_out(kept, ej, 1);
        V = Vsave;
      }
      gemm_tn(s, 1, 1, n, 1.0, ej, n, ej, n, 0.0, w.small1.p, 1, w.gwork.p);
      double nn2 = 0.0;
      down(&nn2, w.small1.p, 1);
      const double nrm = std::sqrt(nn2);
      if (nrm <= 1e-12 || m + kept == cap) continue;
      const double inv = 1.0 / nrm;
      up_bytes(w.idx, &j, sizeof(int));
      up_bytes(w.scale.p, &inv, 8);
      hipLaunchKernelGGL(k_gather_scale, dim3(64, 1), dim3(256), 0, s, E, n, w.idx, w.scale.p, E + int64_t(kept) * n);
      sync();
      ++kept;
    }
    if (kept > 0) {
      BD_HIP(hipMemcpyAsync(V + int64_t(m) * n, E, size_t(n) * kept * 8, hipMemcpyDeviceToDevice, s));
      border(m, m + kept, false);
      if (!factor(m, m + kept)) fail(BD_ERR_APPLICATION, "complex overlap remains dependent after repair");
    }
    return kept;
  }

  bd_status run(const double* x0, int64_t ldx0, const double* hx0, double* eig, double* xout, double* hxout,
                int64_t ldx, int32_t* conv) {
    auto t_start = std::chrono::steady_clock::now();
    zero_doubles(s, w.T.p, int64_t(capR) * capR);
    BD_HIP(hipMemcpy2DAsync(V, n * 8, x0, ldx0 * 8, n * 8, nb, hipMemcpyDeviceToDevice, s));
    if (hx0) BD_HIP(hipMemcpy2DAsync(HV, n * 8, hx0, ldx0 * 8, n * 8, nb, hipMemcpyDeviceToDevice, s));
    else apply(V, nb, HV);
    border(0, nb, false);
    if (!factor(0, nb)) fail(BD_ERR_INVALID_PARAMETER, "complex Davidson start is numerically dependent");
    border(0, nb, true);
    eigen(nb);
    int m = nb;
    std::vector<int> active(nb);
    std::iota(active.begin(), active.end(), 0);
    std::vector<char> converged(nb, 0);
    bool contracted = false;
    int iterations = 0;
    for (int step = 0; step < o.max_iterations; ++step) {
      if (m + int(active.size()) > cap && m > nb) {
        contract(m);
        m = nb;
        contracted = true;
      }
      std::vector<double> nrm = corrections(m, active, contracted);
      contracted = false;
      int k = 0;
      bool staged = false;
      {
        int kk = 0;
        for (double v : nrm) kk += (std::sqrt(v) > 32.0 * kEps) ? 1 : 0;
        staged = m + kk <= cap && kk > 0;
        k = normalize(nrm, staged ? V + int64_t(m) * n : w.E2.p);
      }
      if (!staged || k == 0) {
        k = k > 0 ? repair(m, k) : 0;
      } else {
        border(m, m + k, false);
        if (!factor(m, m + k)) {
          BD_HIP(hipMemcpyAsync(w.E2.p, V + int64_t(m) * n, size_t(n) * k * 8, hipMemcpyDeviceToDevice, s));
          k = repair(m, k);
        }
      }
      std::copy(values.begin(), values.begin() + nb, old.begin());
      if (k > 0) {
        apply(V + int64_t(m) * n, k, HV + int64_t(m) * n);
        border(m, m + k, true);
        eigen(m + k);
        m += k;
      }
      active.clear();
      for (int i = 0; i < nb; ++i) {
        converged[i] = std::fabs(values[i] - old[i]) < band_threshold(i);
        if (!converged[i]) active.push_back(i);
      }
      if (o.profile) {
        for (int i = 0; i < nb; ++i) {
          const int64_t at = int64_t(step) * nb + i;
          if (o.ritz_changes_ha) o.ritz_changes_ha[at] = std::fabs(values[i] - old[i]);
          if (o.ritz_values_ha) {
            o.ritz_values_ha[2 * at] = old[i];
            o.ritz_values_ha[2 * at + 1] = values[i];
          }
        }
      }
      iterations = step + 1;
      if (active.empty() || iterations == o.max_iterations) break;
      if (m + int(active.size()) > cap) {
        contract(m);
        m = nb;
        contracted = true;
      }
    }
    rotate(V, m, w.C.p, 2 * m, nb, nullptr, 1.0, 0.0, xout, ldx);
    if (hxout) rotate(HV, m, w.C.p, 2 * m, nb, nullptr, 1.0, 0.0, hxout, ldx);
    sync();
    bool all = true;
    for (int i = 0; i < nb; ++i) {
      eig[i] = values[i];
      if (conv) conv[i] = converged[i];
      all = all && converged[i];
    }
    rec.iterations = iterations;
    rec.all_converged = all ? 1 : 0;
    rec.seconds_apply = t_apply;
    rec.seconds_dense = t_dense;
    rec.seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return all ? BD_OK : BD_NOT_CONVERGED;
  }
};
}  // namespace

DavidsonWork* davidson_work(GammaOp& op, int64_t n, int nb, int cap);

}  // namespace bd

namespace bd {
DavidsonWork* davidson_work(GammaOp& op, int64_t n, int nb, int cap) {
  if (!op.dav || op.dav->n != n || op.dav->nb != nb || op.dav->cap != cap) {
    delete op.dav;
    op.dav = nullptr;
    op.dav = new DavidsonWork(n, nb, cap);
  }
  return op.dav;
}
void destroy_davidson_work(DavidsonWork* w) { delete w; }
}  // namespace bd

namespace bd {
namespace {
// dst = src, count doubles (a kernel, so that it may sit in a captured graph).
__global__ void k_copy_d(const double* __restrict__ src, int64_t count, double* __restrict__ dst) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

// |E[:, j]|^2 per column, one workgroup per column.
__global__ __launch_bounds__(kT) void k_col_norm2(const double* __restrict__ E, int64_t n, double* __restrict__ out) {
  __shared__ double red[16];
  const double* e = E + int64_t(blockIdx.x) * n;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc += e[i] * e[i];
  const double v = block_sum(acc, red);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// The correction of a converged band (|e| < 1e-9, whose rounding-level
// residual carries no direction) or one lying in span(X) (|e_perp|^2 <=
// 1e-20 |e|^2) is replaced by a fixed pseudo-random direction, so that
// [X, W] stays independent with fixed shapes.
__global__ __launch_bounds__(kT) void k_fix_weak(double* __restrict__ E, int64_t n, const double* __restrict__ before,
                                                 const double* __restrict__ after) {
  const int j = blockIdx.x;
  if (before[j] >= 1e-18 && after[j] > 1e-20 * before[j]) return;
  double* e = E + int64_t(j) * n;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    uint64_t h = (uint64_t(i) * 0x9E3779B97F4A7C15ull) ^ (uint64_t(j + 1) * 0xC2B2AE3D27D4EB4Full);
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    e[i] = double(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

}  // namespace

// Canonical orthonormalisation of a small Gram matrix: given G = U diag(l) U^T
// (ascending l), Tc[:, i] = U[:, i] / sqrt(l_i) for l_i > kCanonRel l_max,
// else 0 (the dropped direction); dropped[i] = 1 for those.
constexpr double kCanonRel = 1e-10;
__global__ void k_canon(const double* __restrict__ U, const double* __restrict__ l, int k, double* __restrict__ Tc,
                        int* __restrict__ dropped) {
  const int i = blockIdx.y;
  const double lmax = l[k - 1];
  const bool keep = l[i] > kCanonRel * lmax && lmax > 0.0;
  const double sc = keep ? 1.0 / sqrt(l[i]) : 0.0;
  for (int r = threadIdx.x; r < k; r += blockDim.x) Tc[int64_t(i) * k + r] = U[int64_t(i) * k + r] * sc;
  if (threadIdx.x == 0) dropped[i] = keep ? 0 : 1;
}

// A[off + i, off + i] += big for every dropped direction i: a zero column of
// the basis then sits far above the wanted eigenvalues.
__global__ void k_lift_dropped(double* __restrict__ A, int64_t lda, int off, const int* __restrict__ dropped, int k,
                               double big) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k && dropped[i]) A[int64_t(off + i) * lda + off + i] += big;
}

// One fixed-shape block Davidson iteration: the nb Ritz vectors X (with HX
// and their Ritz values) and their nb preconditioned residual directions W
// span V = [X, W]. W is made orthogonal to X (two projections) and
// orthonormalised canonically (directions of a near-singular Gram matrix are
// dropped), so V is orthonormal; Rayleigh-Ritz on V (A = V^T HV, lowest nb
// eigenpairs Y) gives X = V Y, HX = HV Y. Every step is device work on the
// op's stream with fixed shapes and no host round trip, so the iteration is
// captured once as a HIP graph and replayed. It is the reference iteration
// (ref_davidson.rs:117-366) with every band active and a restart to the Ritz
// block after each step.
struct BlockGraph {
  GammaOp& op;
  int64_t n;
  int nb, m2;
  int mode;  // 0 QE, 1 TPA
  Dev X, HX, V, HV, E, theta, norm2, norm2b, diag, G, lam, U, Tc, A, Y, small, eigw, gwork;
  int* info = nullptr;     // [0] eigensolver of A, [1] eigensolver of G
  int* dropped = nullptr;  // nb
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  BlockGraph(GammaOp& o, int nb_, int precond) : op(o), n(o.n), nb(nb_), m2(2 * nb_), mode(precond) {
    const size_t c2 = size_t(m2) * m2;
    X.alloc(size_t(n) * nb);
    HX.alloc(size_t(n) * nb);
    V.alloc(size_t(n) * m2);
    HV.alloc(size_t(n) * m2);
    E.alloc(size_t(n) * nb);
    theta.alloc(nb);
    norm2.alloc(nb);
    norm2b.alloc(nb);
    diag.alloc(size_t(n));
    G.alloc(size_t(nb) * nb);
    lam.alloc(nb);
    U.alloc(size_t(nb) * nb);
    Tc.alloc(size_t(nb) * nb);
    A.alloc(c2);
    Y.alloc(size_t(m2) * nb);
    small.alloc(size_t(nb) * nb);
    eigw.alloc(size_t(std::max(sym_eigen_work(m2, nb), sym_eigen_work(nb, nb))));
    gwork.alloc(size_t(std::max(Gemm::tn_work(m2, m2, n), int64_t(c2))));
    BD_HIP(hipMalloc(&info, 2 * sizeof(int)));
    BD_HIP(hipMalloc(&dropped, sizeof(int) * nb));
  }
  ~BlockGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (info) (void)hipFree(info);
    if (dropped) (void)hipFree(dropped);
  }

  // dst (n x nb) = src Tc with Tc from the canonical orthonormalisation of
  // src^T src (src is overwritten by nothing; dst must differ from src).
  void canonical(const double* src, double* dst) {
    hipStream_t s = op.stream;
    op.gemm.tn(s, nb, nb, n, 1.0, src, n, src, n, 0.0, G.p, nb, gwork.p);
    symmetrize(s, nb, G.p, nb);
    sym_eigen_lowest(s, nb, nb, G.p, nb, lam.p, U.p, nb, eigw.p, info + 1);
    hipLaunchKernelGGL(k_canon, dim3(1, nb), dim3(64), 0, s, U.p, lam.p, nb, Tc.p, dropped);
    op.gemm.nn(s, n, nb, nb, 1.0, src, n, Tc.p, nb, 0.0, dst, n);
  }

  // Rayleigh-Ritz on the first m columns of the orthonormal V (HV = H V).
  void ritz(int m) {
    hipStream_t s = op.stream;
    op.gemm.tn(s, m, m, n, 1.0, V.p, n, HV.p, n, 0.0, A.p, m, gwork.p);
    symmetrize(s, m, A.p, m);
    hipLaunchKernelGGL(k_lift_dropped, dim3(unsigned((nb + 255) / 256)), dim3(256), 0, s, A.p, int64_t(m), m - nb,
                       dropped, nb, 1e6);
    sym_eigen_lowest(s, m, nb, A.p, m, theta.p, Y.p, m, eigw.p, info);
    op.gemm.nn(s, n, nb, m, 1.0, V.p, n, Y.p, m, 0.0, X.p, n);
    op.gemm.nn(s, n, nb, m, 1.0, HV.p, n, Y.p, m, 0.0, HX.p, n);
  }

  // The iteration's device work before H·W: corrections, W = E with span(X)
  // projected out (twice) and weak columns replaced, canonical W into the
  // second half of V.
  void before_apply() {
    hipStream_t s = op.stream;
    const unsigned g = unsigned(std::min<int64_t>((n * nb + 255) / 256, 4096));
    hipLaunchKernelGGL(k_corrections, dim3(nb), dim3(kT), 0, s, X.p, HX.p, n, theta.p, diag.p, op.d_kinetic.p, mode,
                       E.p, norm2.p);
    // W: corrections with span(X) projected out (twice), weak ones replaced.
    auto project = [&]() {
      op.gemm.tn(s, nb, nb, n, 1.0, X.p, n, E.p, n, 0.0, small.p, nb, gwork.p);
      op.gemm.nn(s, n, nb, nb, -1.0, X.p, n, small.p, nb, 1.0, E.p, n);
    };
    project();
    hipLaunchKernelGGL(k_col_norm2, dim3(nb), dim3(kT), 0, s, E.p, n, norm2b.p);
    hipLaunchKernelGGL(k_fix_weak, dim3(nb), dim3(kT), 0, s, E.p, n, norm2.p, norm2b.p);
    project();
    project();
    hipLaunchKernelGGL(k_copy_d, dim3(g), dim3(256), 0, s, X.p, n * nb, V.p);
    hipLaunchKernelGGL(k_copy_d, dim3(g), dim3(256), 0, s, HX.p, n * nb, HV.p);
    canonical(E.p, V.p + n * nb);
  }

  // The iteration's device work (captured once, then replayed).
  void enqueue() {
    before_apply();
    op.apply(V.p + n * nb, n, nb, HV.p + n * nb, n);
    ritz(m2);
  }

  void start_before_apply(const double* x0, int64_t ldx0) {
    hipStream_t s = op.stream;
    BD_HIP(hipMemcpy2DAsync(E.p, n * 8, x0, ldx0 * 8, n * 8, nb, hipMemcpyDeviceToDevice, s));
    canonical(E.p, V.p);
  }

  void start(const double* x0, int64_t ldx0) {
    start_before_apply(x0, ldx0);
    op.apply(V.p, n, nb, HV.p, n);
    ritz(nb);
    check();
  }

  void step(bool use_graph) {
    if (!use_graph) {
      enqueue();
    } else {
      if (!exec) {
        if (op.profiling) fail(BD_ERR_INVALID_PARAMETER, "graph capture needs profiling off");
        BD_HIP(hipStreamBeginCapture(op.stream, hipStreamCaptureModeThreadLocal));
        enqueue();
        BD_HIP(hipStreamEndCapture(op.stream, &graph));
        BD_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      }
      BD_HIP(hipGraphLaunch(exec, op.stream));
    }
  }

  // Verdicts of the last iteration (synchronises).
  void check() {
    int v[2] = {0, 0};
    BD_HIP(hipMemcpyAsync(v, info, sizeof(v), hipMemcpyDeviceToHost, op.stream));
    BD_HIP(hipStreamSynchronize(op.stream));
    if (v[0] || v[1]) fail(BD_ERR_APPLICATION, "block iteration: reduced eigensolve failed");
  }
};
// The block iterations of every k point of a map stepped together. Each
// iteration forks from one stream into the k points' own streams for the
// per-k work (parallel branches), joins for one fused launch of H·W over all
// columns of all k points, and forks again for the per-k Rayleigh-Ritz; the
// whole iteration is captured once as one HIP graph and replayed. The work
// tables of the two batched applies (start: H·V, step: H·W) hold fixed
// buffers, so they are uploaded once.
struct KBlocks {
  std::vector<GammaOp*> ops;
  std::vector<std::unique_ptr<BlockGraph>> blocks;
  int nb;
  hipStream_t main = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  DevBuf<KWork> wstart, wstep;
  FusedShape f;
  size_t lds = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  KBlocks(const std::vector<GammaOp*>& o, int nb_, int mode) : ops(o), nb(nb_) {
    const GammaOp& o0 = *ops[0];
    f = o0.fshape;
    for (GammaOp* op : ops) {
      if (!op->kb || !op->fused) fail(BD_ERR_INVALID_PARAMETER, "k blocks need k-point ops on the fused path");
      if (op->sys != o0.sys || op->precision != o0.precision)
        fail(BD_ERR_INVALID_PARAMETER, "k blocks must share the system and precision");
      if (!op->fused_valid(f)) fail(BD_ERR_INVALID_PARAMETER, "k blocks need a common fused shape");
      if (nb <= 0 || 2 * int64_t(nb) > op->n || nb > op->capacity)
        fail(BD_ERR_INVALID_PARAMETER, "block iteration: bands vs rows and capacity");
      lds = std::max(lds, op->fused_lds(f));
      blocks.emplace_back(new BlockGraph(*op, nb, mode));
    }
    const int count = int(ops.size());
    std::vector<KWork> ws(count), wt(count);
    for (int k = 0; k < count; ++k) {
      BlockGraph& b = *blocks[k];
      ws[k] = ops[k]->kwork(b.V.p, b.n, nb, b.HV.p, b.n, k * nb);
      wt[k] = ops[k]->kwork(b.V.p + b.n * nb, b.n, nb, b.HV.p + b.n * nb, b.n, k * nb);
    }
    wstart.alloc(size_t(count));
    wstep.alloc(size_t(count));
    BD_HIP(hipMemcpy(wstart.p, ws.data(), sizeof(KWork) * count, hipMemcpyHostToDevice));
    BD_HIP(hipMemcpy(wstep.p, wt.data(), sizeof(KWork) * count, hipMemcpyHostToDevice));
    BD_HIP(hipStreamCreateWithFlags(&main, hipStreamNonBlocking));
    BD_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    joins.resize(count, nullptr);
    for (auto& e : joins) BD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  ~KBlocks() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto e : joins)
      if (e) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (main) (void)hipStreamDestroy(main);
  }

  // fn(k) on every k point's stream after the work queued on `main`; `main`
  // then waits for all of them.
  template <class F>
  void each(F fn) {
    BD_HIP(hipEventRecord(fork, main));
    for (size_t k = 0; k < ops.size(); ++k) {
      BD_HIP(hipStreamWaitEvent(ops[k]->stream, fork, 0));
      fn(int(k));
      BD_HIP(hipEventRecord(joins[k], ops[k]->stream));
    }
    for (size_t k = 0; k < ops.size(); ++k) BD_HIP(hipStreamWaitEvent(main, joins[k], 0));
  }

  // `main` after everything queued on the k points' streams (potentials,
  // diagonals set since the last call).
  void order_after_ops() {
    for (size_t k = 0; k < ops.size(); ++k) {
      BD_HIP(hipEventRecord(joins[k], ops[k]->stream));
      BD_HIP(hipStreamWaitEvent(main, joins[k], 0));
    }
  }

  void batched_apply(const DevBuf<KWork>& w) {
    const KWork none{};
    fused_launch(*ops[0], f, lds, w.p, none, int(ops.size()), int(ops.size()) * nb, main);
  }

  void enqueue() {
    each([&](int k) { blocks[k]->before_apply(); });
    batched_apply(wstep);
    each([&](int k) { blocks[k]->ritz(2 * nb); });
  }

  void start(const double* const* x0, const int64_t* ldx0) {
    order_after_ops();
    each([&](int k) { blocks[k]->start_before_apply(x0[k], ldx0[k]); });
    batched_apply(wstart);
    each([&](int k) { blocks[k]->ritz(nb); });
    BD_HIP(hipStreamSynchronize(main));
    for (auto& b : blocks) b->check();
  }

  void step(bool use_graph) {
    if (!use_graph) {
      enqueue();
      return;
    }
    if (!exec) {
      for (GammaOp* op : ops)
        if (op->profiling) fail(BD_ERR_INVALID_PARAMETER, "graph capture needs profiling off");
      BD_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
      enqueue();
      BD_HIP(hipStreamEndCapture(main, &graph));
      BD_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    BD_HIP(hipGraphLaunch(exec, main));
  }

  // Ritz values of every k point (count x nb, host) after the queued work;
  // fails when any reduced eigensolve failed.
  void values(double* out) {
    for (size_t k = 0; k < ops.size(); ++k)
      BD_HIP(hipMemcpyAsync(out + k * nb, blocks[k]->theta.p, size_t(nb) * 8, hipMemcpyDeviceToHost, main));
    BD_HIP(hipStreamSynchronize(main));
    for (auto& b : blocks) b->check();
  }
};
}  // namespace bd

struct bd_block {
  bd::BlockGraph g;
  bd_block(bd::GammaOp& op, int nb, int mode) : g(op, nb, mode) {}
};

struct bd_kblocks {
  bd::KBlocks g;
  bd_kblocks(const std::vector<bd::GammaOp*>& ops, int nb, int mode) : g(ops, nb, mode) {}
};

namespace {
// One solve on one op (the body of bd_davidson_solve_dev); errors as status
// plus message.
bd_status solve_one(bd_gamma_op* h, const double* x0, int64_t ldx0, const double* hx0, const double* diag,
                    const bd_davidson_options* o, double* eig, double* xout, double* hxout, int64_t ldx,
                    int32_t* conv, bd_davidson_record* record, std::string* msg) {
  try {
    if (!h || !x0 || !diag || !o || !eig || !xout) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& op = h->op;
    const int64_t n = op.n;
    const int nb = o->bands;
    if (o->complex_columns) {
      // Complex columns: n / 2 complex rows per column.
      if (!op.kb) bd::fail(BD_ERR_INVALID_PARAMETER, "complex Davidson needs a k-point op");
      const int64_t nc = n / 2;
      const int cap = int(std::min<int64_t>(o->capacity == 0 ? 2 * int64_t(nb) : o->capacity, nc));
      if (nb <= 0 || nb > nc || cap < std::min<int64_t>(2 * int64_t(nb), nc) || o->occupied > nb ||
          o->max_iterations <= 0 || !(o->threshold > 0.0) || !std::isfinite(o->threshold) || ldx0 < n ||
          ldx < n || (o->return_hx && !hxout))
        bd::fail(BD_ERR_INVALID_PARAMETER, "complex Davidson dimensions/threshold/start");
      if (nb > op.capacity) bd::fail(BD_ERR_INVALID_PARAMETER, "bands exceed the operator capacity");
      if (o->preconditioner != BD_PRECOND_QE && o->preconditioner != BD_PRECOND_TPA)
        bd::fail(BD_ERR_INVALID_PARAMETER, "complex Davidson preconditioner (QE or TPA)");
      if (!op.has_potential) bd::fail(BD_ERR_INVALID_PARAMETER, "no potential installed");
      bd::ComplexWork* w = bd::complex_work(op, n, nb, cap);
      bd::ComplexSolver solver(op, *w, *o, diag);
      bd_status st = solver.run(x0, ldx0, hx0, eig, xout, o->return_hx ? hxout : nullptr, ldx, conv);
      if (record) *record = solver.rec;
      if (st == BD_NOT_CONVERGED) *msg = "complex Davidson reached the iteration limit";
      return st;
    }
    const int cap = int(std::min<int64_t>(o->capacity == 0 ? 2 * int64_t(nb) : o->capacity, n));
    if (nb <= 0 || nb > n || cap < std::min<int64_t>(2 * int64_t(nb), n) || o->occupied > nb ||
        o->max_iterations <= 0 || !(o->threshold > 0.0) || !std::isfinite(o->threshold) || ldx0 < n ||
        ldx < n || (o->return_hx && !hxout))
      bd::fail(BD_ERR_INVALID_PARAMETER, "reference Davidson dimensions/threshold/start");
    if (nb > op.capacity) bd::fail(BD_ERR_INVALID_PARAMETER, "bands exceed the operator capacity");
    if (o->preconditioner < BD_PRECOND_QE || o->preconditioner > BD_PRECOND_MLP ||
        (o->preconditioner == BD_PRECOND_MLP && (!o->mlp_blob || !std::isfinite(o->mlp_vbar))))
      bd::fail(BD_ERR_INVALID_PARAMETER, "reference Davidson preconditioner");
    if (!op.has_potential) bd::fail(BD_ERR_INVALID_PARAMETER, "no potential installed");
    bd::DavidsonWork* w = bd::davidson_work(op, n, nb, cap);
    bd::Solver solver(op, *w, *o, diag);
    bd_status st = solver.run(x0, ldx0, hx0, eig, xout, o->return_hx ? hxout : nullptr, ldx, conv);
    if (record) *record = solver.rec;
    if (st == BD_NOT_CONVERGED) *msg = "reference Davidson reached the iteration limit";
    return st;
  } catch (const bd::Error& e) {
    *msg = e.what();
    return e.status;
  } catch (const std::exception& e) {
    *msg = e.what();
    return BD_ERR_APPLICATION;
  }
}
}  // namespace

extern "C" bd_status bd_davidson_solve_dev(bd_gamma_op* h, const double* x0, int64_t ldx0, const double* hx0,
                                           const double* diag, const bd_davidson_options* o, double* eig,
                                           double* xout, double* hxout, int64_t ldx, int32_t* conv,
                                           bd_davidson_record* record) {
  bd::clear_last_error();
  std::string msg;
  const bd_status st = solve_one(h, x0, ldx0, hx0, diag, o, eig, xout, hxout, ldx, conv, record, &msg);
  if (st != BD_OK) bd::set_last_error(msg);
  return st;
}

// Independent solves on `count` distinct ops, `threads` host threads at a
// time, each op on its own stream (the k points of an SCF map). statuses[i]
// per solve; the return value is the first non-OK status other than
// BD_NOT_CONVERGED (then its message is the last error), else BD_NOT_CONVERGED
// if any solve hit its iteration limit, else BD_OK.
extern "C" bd_status bd_davidson_solve_many_dev(int32_t count, bd_gamma_op* const* ops, const double* const* x0,
                                                const int64_t* ldx0, const double* const* diag,
                                                const bd_davidson_options* options, double* const* eig,
                                                double* const* xout, double* const* hxout, const int64_t* ldx,
                                                int32_t* const* conv, bd_davidson_record* records,
                                                bd_status* statuses, int32_t threads) {
  bd::clear_last_error();
  if (count < 0 || !ops || !x0 || !ldx0 || !diag || !options || !eig || !xout || !ldx || !statuses) {
    bd::set_last_error("null argument");
    return BD_ERR_INVALID_PARAMETER;
  }
  for (int i = 0; i < count; ++i)
    for (int j = 0; j < i; ++j)
      if (ops[i] == ops[j]) {
        bd::set_last_error("bd_davidson_solve_many_dev: an op appears twice");
        return BD_ERR_INVALID_PARAMETER;
      }
  std::vector<std::string> msgs(count);
  std::atomic<int> next{0};
  int dev = 0;
  (void)hipGetDevice(&dev);
  auto worker = [&]() {
    (void)hipSetDevice(dev);
    for (int i = next.fetch_add(1); i < count; i = next.fetch_add(1))
      statuses[i] = solve_one(ops[i], x0[i], ldx0[i], nullptr, diag[i], &options[i], eig[i], xout[i],
                              hxout ? hxout[i] : nullptr, ldx[i], conv ? conv[i] : nullptr,
                              records ? &records[i] : nullptr, &msgs[i]);
  };
  const int nt = std::max(1, std::min<int>(threads, count));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
  bd_status result = BD_OK;
  for (int i = 0; i < count; ++i) {
    if (statuses[i] != BD_OK && statuses[i] != BD_NOT_CONVERGED) {
      bd::set_last_error("k " + std::to_string(i) + ": " + msgs[i]);
      return statuses[i];
    }
    if (statuses[i] == BD_NOT_CONVERGED) result = BD_NOT_CONVERGED;
  }
  return result;
}

// ---------------------------------------------------------------------------
// Fixed-shape block iteration, eager or replayed from a captured HIP graph
// ---------------------------------------------------------------------------
extern "C" bd_status bd_block_create(bd_gamma_op* h, int32_t bands, int32_t preconditioner, bd_block** out) {
  BD_GUARD({
    if (!h || !out) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& op = h->op;
    if (bands <= 0 || 2 * int64_t(bands) > op.n || bands > op.capacity)
      bd::fail(BD_ERR_INVALID_PARAMETER, "block iteration: bands vs rows and capacity");
    if (preconditioner != BD_PRECOND_QE && preconditioner != BD_PRECOND_TPA)
      bd::fail(BD_ERR_INVALID_PARAMETER, "block iteration: QE or TPA preconditioner");
    *out = new bd_block(op, bands, preconditioner == BD_PRECOND_QE ? 0 : 1);
  });
}

extern "C" void bd_block_destroy(bd_block* b) { delete b; }

// The Davidson diagonal (device, n rows); read by every later step.
extern "C" bd_status bd_block_set_diagonal_dev(bd_block* b, const double* diag) {
  BD_GUARD({
    if (!b || !diag) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    BD_HIP(hipMemcpyAsync(b->g.diag.p, diag, size_t(b->g.n) * 8, hipMemcpyDeviceToDevice, b->g.op.stream));
    BD_HIP(hipStreamSynchronize(b->g.op.stream));
  });
}

// Rayleigh-Ritz of the start block (n x bands, ldx0): X, HX and the values.
extern "C" bd_status bd_block_start_dev(bd_block* b, const double* x0, int64_t ldx0, double* values) {
  BD_GUARD({
    if (!b || !x0 || !values) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& g = b->g;
    if (ldx0 < g.n) bd::fail(BD_ERR_INVALID_PARAMETER, "leading dimension smaller than rows");
    if (!g.op.has_potential) bd::fail(BD_ERR_INVALID_PARAMETER, "no potential installed");
    g.start(x0, ldx0);
    BD_HIP(hipMemcpy(values, g.theta.p, size_t(g.nb) * 8, hipMemcpyDeviceToHost));
  });
}

// `steps` iterations (use_graph: captured on the first call, then replayed);
// values (host, bands) of the last one. Synchronises once at the end.
extern "C" bd_status bd_block_step(bd_block* b, int32_t steps, int32_t use_graph, double* values) {
  BD_GUARD({
    if (!b || !values || steps <= 0) bd::fail(BD_ERR_INVALID_PARAMETER, "block step arguments");
    auto& g = b->g;
    for (int i = 0; i < steps; ++i) g.step(use_graph != 0);
    BD_HIP(hipMemcpyAsync(values, g.theta.p, size_t(g.nb) * 8, hipMemcpyDeviceToHost, g.op.stream));
    g.check();
  });
}

// Enqueue `steps` iterations without waiting (several ops' streams can then
// run side by side); bd_block_finish waits and reads the values.
extern "C" bd_status bd_block_step_async(bd_block* b, int32_t steps, int32_t use_graph) {
  BD_GUARD({
    if (!b || steps <= 0) bd::fail(BD_ERR_INVALID_PARAMETER, "block step arguments");
    for (int i = 0; i < steps; ++i) b->g.step(use_graph != 0);
  });
}

extern "C" bd_status bd_block_finish(bd_block* b, double* values) {
  BD_GUARD({
    if (!b || !values) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& g = b->g;
    BD_HIP(hipMemcpyAsync(values, g.theta.p, size_t(g.nb) * 8, hipMemcpyDeviceToHost, g.op.stream));
    g.check();
  });
}

// Copy X (and HX when hx is not NULL) out, n x bands with leading dimension ld.
extern "C" bd_status bd_block_vectors_dev(bd_block* b, double* x, double* hx, int64_t ld) {
  BD_GUARD({
    if (!b || !x) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& g = b->g;
    if (ld < g.n) bd::fail(BD_ERR_INVALID_PARAMETER, "leading dimension smaller than rows");
    BD_HIP(hipMemcpy2DAsync(x, ld * 8, g.X.p, g.n * 8, g.n * 8, g.nb, hipMemcpyDeviceToDevice, g.op.stream));
    if (hx)
      BD_HIP(hipMemcpy2DAsync(hx, ld * 8, g.HX.p, g.n * 8, g.n * 8, g.nb, hipMemcpyDeviceToDevice, g.op.stream));
    BD_HIP(hipStreamSynchronize(g.op.stream));
  });
}

// ---------------------------------------------------------------------------
// Block iterations of all k points of a map: per-k work on the k points'
// streams, one fused H·W launch for all of them, one captured graph
// ---------------------------------------------------------------------------
extern "C" bd_status bd_kblocks_create(bd_gamma_op* const* ops, int32_t count, int32_t bands, int32_t preconditioner,
                                       bd_kblocks** out) {
  BD_GUARD({
    if (!ops || !out || count <= 0) bd::fail(BD_ERR_INVALID_PARAMETER, "empty k blocks");
    if (preconditioner != BD_PRECOND_QE && preconditioner != BD_PRECOND_TPA)
      bd::fail(BD_ERR_INVALID_PARAMETER, "block iteration: QE or TPA preconditioner");
    std::vector<bd::GammaOp*> v;
    for (int i = 0; i < count; ++i) {
      if (!ops[i]) bd::fail(BD_ERR_INVALID_PARAMETER, "null op in k blocks");
      for (int j = 0; j < i; ++j)
        if (ops[j] == ops[i]) bd::fail(BD_ERR_INVALID_PARAMETER, "an op appears twice in k blocks");
      v.push_back(&ops[i]->op);
    }
    *out = new bd_kblocks(v, bands, preconditioner == BD_PRECOND_QE ? 0 : 1);
  });
}

extern "C" void bd_kblocks_destroy(bd_kblocks* b) { delete b; }

// The Davidson diagonals of every k point (device, n_k rows each).
extern "C" bd_status bd_kblocks_set_diagonals_dev(bd_kblocks* b, const double* const* diag) {
  BD_GUARD({
    if (!b || !diag) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& g = b->g;
    for (size_t k = 0; k < g.ops.size(); ++k) {
      if (!diag[k]) bd::fail(BD_ERR_INVALID_PARAMETER, "null diagonal");
      BD_HIP(hipMemcpyAsync(g.blocks[k]->diag.p, diag[k], size_t(g.blocks[k]->n) * 8, hipMemcpyDeviceToDevice,
                            g.ops[k]->stream));
    }
    for (auto* op : g.ops) BD_HIP(hipStreamSynchronize(op->stream));
  });
}

// Rayleigh-Ritz of every start block (n_k x bands, ldx0[k]); values: count x
// bands (host).
extern "C" bd_status bd_kblocks_start_dev(bd_kblocks* b, const double* const* x0, const int64_t* ldx0,
                                          double* values) {
  BD_GUARD({
    if (!b || !x0 || !ldx0 || !values) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& g = b->g;
    for (size_t k = 0; k < g.ops.size(); ++k) {
      if (!x0[k] || ldx0[k] < g.ops[k]->n) bd::fail(BD_ERR_INVALID_PARAMETER, "start block");
      if (!g.ops[k]->has_potential) bd::fail(BD_ERR_INVALID_PARAMETER, "no potential installed");
    }
    g.start(x0, ldx0);
    g.values(values);
  });
}

// `steps` iterations of all k points (use_graph: captured on the first call,
// then replayed); values (count x bands, host) of the last one.
extern "C" bd_status bd_kblocks_step(bd_kblocks* b, int32_t steps, int32_t use_graph, double* values) {
  BD_GUARD({
    if (!b || !values || steps <= 0) bd::fail(BD_ERR_INVALID_PARAMETER, "block step arguments");
    auto& g = b->g;
    g.order_after_ops();
    for (int i = 0; i < steps; ++i) g.step(use_graph != 0);
    g.values(values);
  });
}

// X (and HX when hx is not NULL) of every k point, n_k x bands with leading
// dimension ld[k].
extern "C" bd_status bd_kblocks_vectors_dev(bd_kblocks* b, double* const* x, double* const* hx, const int64_t* ld) {
  BD_GUARD({
    if (!b || !x || !ld) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& g = b->g;
    for (size_t k = 0; k < g.ops.size(); ++k) {
      auto& bl = *g.blocks[k];
      if (!x[k] || ld[k] < bl.n) bd::fail(BD_ERR_INVALID_PARAMETER, "vector block");
      BD_HIP(hipMemcpy2DAsync(x[k], ld[k] * 8, bl.X.p, bl.n * 8, bl.n * 8, bl.nb, hipMemcpyDeviceToDevice, g.main));
      if (hx && hx[k])
        BD_HIP(hipMemcpy2DAsync(hx[k], ld[k] * 8, bl.HX.p, bl.n * 8, bl.n * 8, bl.nb, hipMemcpyDeviceToDevice,
                                g.main));
    }
    BD_HIP(hipStreamSynchronize(g.main));
  });
}
