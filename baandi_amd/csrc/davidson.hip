// Reference block Davidson (crates/dft-solvers/src/ref_davidson.rs:117-366) on
// the GPU operator. Control flow runs on the host; every n-row operation
// (H·X, border Gram GEMMs, rotations, corrections, normalisation, rank
// repair projections) runs on the op's stream. The small FP64 subspace steps
// run on the host: the bordered inverse Cholesky (lean_davidson.rs:167-331)
// and the reduced eigensolve; the m x m products around them (A = T H T^T,
// C = T^T Y) run on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <utility>
#include <vector>

#include "common.hpp"
#include "dense.hpp"
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
__device__ __forceinline__ double sqrt_denominator(double h, double e) {
  const double x = 2.0 * (h - e);
  const double z = fabs(x - 1.0);
  return 0.25 * (2.0 * fmax(x, 1.0) + 1.0 / (sqrt(1.0 + z * z) + z));
}

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
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
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
  Dev V, HV, X, HX, E, E2, small1, small2, small3, gwork, theta, norm2, scale;
  int* idx = nullptr;
  int* flag = nullptr;
  DavidsonWork(int64_t n_, int nb_, int cap_) : n(n_), nb(nb_), cap(cap_) {
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
    scale.alloc(nb);
    BD_HIP(hipMalloc(&idx, sizeof(int) * std::max(nb, 1)));
    BD_HIP(hipMalloc(&flag, sizeof(int)));
  }
  ~DavidsonWork() {
    if (idx) (void)hipFree(idx);
    if (flag) (void)hipFree(flag);
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
  // host subspace state (column-major, ld = cap)
  std::vector<double> S, Hr, T, C, values, old;
  bd_davidson_record rec{};
  double t_apply = 0.0, t_dense = 0.0;
  hipEvent_t ev0, ev1;

  Solver(GammaOp& op_, DavidsonWork& w_, const bd_davidson_options& o_, const double* diag_)
      : op(op_), w(w_), s(op_.stream), n(w_.n), nb(w_.nb), cap(w_.cap), o(o_), diag(diag_) {
    V = w.V.p;
    HV = w.HV.p;
    X = w.X.p;
    HX = w.HX.p;
    S.assign(size_t(cap) * cap, 0.0);
    Hr.assign(size_t(cap) * cap, 0.0);
    T.assign(size_t(cap) * cap, 0.0);
    values.assign(cap, 0.0);
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
    op.gemm.tn(s, to, k, n, 1.0, V, n, B, n, 0.0, w.small1.p, to, w.gwork.p);
    std::vector<double> h(size_t(to) * k);
    down(h.data(), w.small1.p, h.size());
    std::vector<double>& M = is_h ? Hr : S;
    for (int jj = 0; jj < k; ++jj) {
      const int j = from + jj;
      for (int i = 0; i <= j; ++i) {
        const double v = h[i + size_t(jj) * to];
        M[i + size_t(j) * cap] = v;
        M[j + size_t(i) * cap] = v;
      }
    }
  }

  bool factor(int from, int to) {
    auto t0 = std::chrono::steady_clock::now();
    double ratio;
    bool ok = extend_inverse_cholesky(S.data(), cap, from, to, T.data(), &ratio);
    t_dense += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return ok;
  }

  // Host (m x m, ld cap) -> device compact (ld m).
  void up_square(double* d, const std::vector<double>& h, int m, bool transpose) {
    std::vector<double> c(size_t(m) * m);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) c[i + size_t(j) * m] = transpose ? h[j + size_t(i) * cap] : h[i + size_t(j) * cap];
    up(d, c.data(), c.size());
  }

  // A = T H_r T^T, EVD, C = T^T Y[:, :nb] (workspace.rs:435-499).
  void eigen(int m) {
    up_square(w.small1.p, T, m, true);   // T^T
    up_square(w.small2.p, Hr, m, false); // H_r
    gemm_nn(s, m, m, m, 1.0, w.small2.p, m, w.small1.p, m, 0.0, w.small3.p, m);  // H_r T^T
    up_square(w.small2.p, T, m, false);  // T
    gemm_nn(s, m, m, m, 1.0, w.small2.p, m, w.small3.p, m, 0.0, w.small1.p, m);  // A
    std::vector<double> A(size_t(m) * m), Y(size_t(m) * m);
    down(A.data(), w.small1.p, A.size());
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < j; ++i) {
        const double v = 0.5 * (A[i + size_t(j) * m] + A[j + size_t(i) * m]);
        A[i + size_t(j) * m] = A[j + size_t(i) * m] = v;
      }
    auto t0 = std::chrono::steady_clock::now();
    if (!symmetric_eigen(A.data(), m, m, values.data(), Y.data()))
      fail(BD_ERR_APPLICATION, "reference reduced eigensolve failed");
    t_dense += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    up_square(w.small2.p, T, m, true);  // T^T
    up(w.small3.p, Y.data(), size_t(m) * nb);
    gemm_nn(s, m, nb, m, 1.0, w.small2.p, m, w.small3.p, m, 0.0, w.small1.p, m);  // C
    C.assign(size_t(m) * nb, 0.0);
    down(C.data(), w.small1.p, C.size());
  }

  void refresh(int m) {
    border(0, m, false);
    border(0, m, true);
    if (!factor(0, m)) fail(BD_ERR_APPLICATION, "reference contracted basis is dependent");
  }

  // V <- V C (nb columns), HV <- HV C, measured refresh, C = I (workspace.rs:772-935).
  void contract(int m) {
    up(w.small1.p, C.data(), size_t(m) * nb);
    op.gemm.nn(s, n, nb, m, 1.0, V, n, w.small1.p, m, 0.0, X, n);
    op.gemm.nn(s, n, nb, m, 1.0, HV, n, w.small1.p, m, 0.0, HX, n);
    std::swap(V, X);
    std::swap(HV, HX);
    refresh(nb);
    C.assign(size_t(nb) * nb, 0.0);
    for (int i = 0; i < nb; ++i) C[i + size_t(i) * nb] = 1.0;
    rec.restarts += 1;
  }

  // Corrections of the active bands into w.E; returns host |e|^2.
  std::vector<double> corrections(int m, const std::vector<int>& active, bool contracted) {
    const int k = int(active.size());
    std::vector<double> th(k);
    for (int j = 0; j < k; ++j) th[j] = values[active[j]];
    up(w.theta.p, th.data(), k);
    if (contracted) {
      std::vector<int> idx(active);
      up_bytes(w.idx, idx.data(), sizeof(int) * k);
      hipLaunchKernelGGL(k_gather_scale, dim3(64, k), dim3(256), 0, s, V, n, w.idx, nullptr, X);
      hipLaunchKernelGGL(k_gather_scale, dim3(64, k), dim3(256), 0, s, HV, n, w.idx, nullptr, HX);
    } else {
      std::vector<double> ca(size_t(m) * k);
      for (int j = 0; j < k; ++j)
        std::copy(&C[size_t(active[j]) * m], &C[size_t(active[j]) * m] + m, &ca[size_t(j) * m]);
      up(w.small1.p, ca.data(), ca.size());
      op.gemm.nn(s, n, k, m, 1.0, V, n, w.small1.p, m, 0.0, X, n);
      op.gemm.nn(s, n, k, m, 1.0, HV, n, w.small1.p, m, 0.0, HX, n);
    }
    hipLaunchKernelGGL(k_corrections, dim3(k), dim3(kT), 0, s, X, HX, n, w.theta.p, diag, op.d_kinetic.p,
                       o.preconditioner == BD_PRECOND_TPA ? 1 : 0, w.E.p, w.norm2.p);
    std::vector<double> nrm(k);
    down(nrm.data(), w.norm2.p, k);
    return nrm;
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
    std::vector<double> SC(size_t(m) * nb, 0.0), TSC(size_t(m) * nb, 0.0);
    for (int c = 0; c < nb; ++c)
      for (int j = 0; j < m; ++j) {
        const double v = C[j + size_t(c) * m];
        for (int i = 0; i < m; ++i) SC[i + size_t(c) * m] += S[i + size_t(j) * cap] * v;
      }
    for (int c = 0; c < nb; ++c)
      for (int j = 0; j < m; ++j) {
        const double v = SC[j + size_t(c) * m];
        for (int i = j; i < m; ++i) TSC[i + size_t(c) * m] += T[i + size_t(j) * cap] * v;
      }
    up_square(w.small1.p, T, m, true);  // T^T
    op.gemm.nn(s, n, m, m, 1.0, V, n, w.small1.p, m, 0.0, X, n);
    op.gemm.nn(s, n, m, m, 1.0, HV, n, w.small1.p, m, 0.0, HX, n);
    std::swap(V, X);
    std::swap(HV, HX);
    C = TSC;
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
      iterations = step + 1;
      if (active.empty() || iterations == o.max_iterations) break;
      if (m + int(active.size()) > cap) {
        contract(m);
        m = nb;
        contracted = true;
      }
    }
    // Return X = V C, HX = HV C (workspace.rs:936-980).
    up(w.small1.p, C.data(), size_t(m) * nb);
    op.gemm.nn(s, n, nb, m, 1.0, V, n, w.small1.p, m, 0.0, xout, ldx);
    if (hxout) op.gemm.nn(s, n, nb, m, 1.0, HV, n, w.small1.p, m, 0.0, hxout, ldx);
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

extern "C" bd_status bd_davidson_solve_dev(bd_gamma_op* h, const double* x0, int64_t ldx0, const double* hx0,
                                           const double* diag, const bd_davidson_options* o, double* eig,
                                           double* xout, double* hxout, int64_t ldx, int32_t* conv,
                                           bd_davidson_record* record) {
  bd::clear_last_error();
  try {
    if (!h || !x0 || !diag || !o || !eig || !xout) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& op = h->op;
    const int64_t n = op.n;
    const int nb = o->bands;
    const int cap = int(std::min<int64_t>(o->capacity == 0 ? 2 * int64_t(nb) : o->capacity, n));
    if (nb <= 0 || nb > n || cap < std::min<int64_t>(2 * int64_t(nb), n) || o->occupied > nb ||
        o->max_iterations <= 0 || !(o->threshold > 0.0) || !std::isfinite(o->threshold) || ldx0 < n ||
        ldx < n || (o->return_hx && !hxout))
      bd::fail(BD_ERR_INVALID_PARAMETER, "reference Davidson dimensions/threshold/start");
    if (nb > op.capacity) bd::fail(BD_ERR_INVALID_PARAMETER, "bands exceed the operator capacity");
    if (!op.has_potential) bd::fail(BD_ERR_INVALID_PARAMETER, "no potential installed");
    bd::DavidsonWork* w = bd::davidson_work(op, n, nb, cap);
    bd::Solver solver(op, *w, *o, diag);
    bd_status st = solver.run(x0, ldx0, hx0, eig, xout, o->return_hx ? hxout : nullptr, ldx, conv);
    if (record) *record = solver.rec;
    if (st == BD_NOT_CONVERGED) bd::set_last_error("reference Davidson reached the iteration limit");
    return st;
  } catch (const bd::Error& e) {
    bd::set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    bd::set_last_error(e.what());
    return BD_ERR_APPLICATION;
  }
}
