// The complex Davidson of every k point of a map in lock step
// (ref_davidson_complex::solve_with_initial_hx,
// crates/dft-solvers/src/ref_davidson_complex.rs:641-1113, run for the k
// points of crates/dft-engine/src/refscf/k_map.rs:181-208 at once).
//
// Each k point follows its own Davidson: its own basis size, active bands,
// convergence, contractions and ideal resets. What is shared is the launch:
// every stage of an iteration is one launch over all k points that take part
// in it (a workgroup row per k point), and H·W of every k point's new columns
// is one launch of the fused k-point kernel. The subspace step of a k point
// (bordered Cholesky of its complex overlap, L^-1 H L^-H, a complex Jacobi
// eigensolve, C = L^-H Y) runs in one workgroup with the small matrices in
// LDS. The host reads back two small status blocks per iteration (the kept
// corrections, then the Ritz values and verdicts), for all k points together.
//
// Layout per k point: V, HV, X, HX (n x cap doubles: cap complex columns of
// n = 2 n_k interleaved (re, im) rows), E (n x bands), and the complex small
// matrices S, H, L (cap x cap) and C (cap x bands), column-major.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "eigen.hpp"
#include "op.hpp"

namespace bd {
namespace {

constexpr int kKB = 24;    // bands per k point
constexpr int kKCap = 48;  // complex basis columns per k point
constexpr int kKT = 256;
constexpr int kEigT = 256;  // k_kd_eigen workgroup (64 measured twice as slow: the rounds are work-bound)
constexpr double kEps = std::numeric_limits<double>::epsilon();
constexpr double kPivotMin = kEps / 1e-12;  // PIVOT_RATIO_MIN, lean_davidson.rs:64
constexpr double kOrthoLimit = 1e-12;       // ORTHOGONALITY_LIMIT, ref_davidson/ideal_reset.rs:7
constexpr int kRefreshEvery = 4;            // REFRESH_INTERVAL, ref_davidson/ideal_reset.rs:8

struct KDesc {
  double* V;
  double* HV;
  double* X;
  double* HX;
  double* E;
  double2* S;
  double2* H;
  double2* L;
  double2* C;
  const double* diag;
  const double* kinetic;
  const double* in_x;  // start block (ld ldi) and optional H·start
  const double* in_hx;
  double* out_x;  // final Ritz block (ld ldo) and optional H·X
  double* out_hx;
  int64_t n;  // doubles per column
  int64_t ldi, ldo;
  int m;         // basis columns
  int kk;        // border columns [m, m + kk)
  int nact;      // active bands
  int live;      // corrections stage
  int stage2;    // border / factor / H / eigen stage
  int contract;  // contraction at the start of the step
  int act[kKB];
};

struct KStat {
  int kept, nonfinite_e, factor_bad, eig_bad, nonfinite_h, refresh, refresh_bad, repair;
  int contractions, disabled, ideal_resets, periodic, guard_fallbacks, repair_bad, kept2, pad;
  double values[kKB];
};

// The border / factor range [from, to) of one k point for a stage, or false
// when it takes no part. 0: the new corrections [m, m + kept) of live k
// points; 1: the restarted block [0, nb) of contracting k points that measure
// it; 2: [0, nb) of every live k point (the start); 3: the whitened basis
// [0, m) of k points in rank repair; 4: their repaired corrections
// [m, m + kept2); 5: the columns H was applied to, [m, m + kept) or
// [m, m + kept2) after a repair.
__device__ __forceinline__ bool kd_range(const KDesc& d, const KStat& st, int mode, int nb, int& from, int& to) {
  switch (mode) {
    case 0:
      if (!d.live || st.kept <= 0) return false;
      from = d.m;
      to = d.m + st.kept;
      return true;
    case 1:
      if (!d.contract || !st.refresh) return false;
      from = 0;
      to = nb;
      return true;
    case 2:
      if (!d.live) return false;
      from = 0;
      to = nb;
      return true;
    case 3:
      if (!d.stage2 || !st.repair) return false;
      from = 0;
      to = d.m;
      return true;
    case 4:
      if (!d.stage2 || !st.repair || st.kept2 <= 0) return false;
      from = d.m;
      to = d.m + st.kept2;
      return true;
    case 5: {
      if (!d.stage2) return false;
      const int kk = st.repair ? st.kept2 : st.kept;
      if (kk <= 0) return false;
      from = d.m;
      to = d.m + kk;
      return true;
    }
  }
  return false;
}

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) b
__device__ __forceinline__ double2 zcmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double zabs2(double2 a) { return a.x * a.x + a.y * a.y; }

// Deterministic workgroup sum (fixed tree), valid on every thread.
__device__ double kd_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < int(blockDim.x >> 6); ++w) s += red[w];
  __syncthreads();
  return s;
}

// Copy the start block (and H·start) of every k point into V (HV).
__global__ void k_kd_copy_in(const KDesc* __restrict__ D, int nb) {
  const KDesc& d = D[blockIdx.y];
  const int64_t total = d.n * nb;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t j = t / d.n, i = t % d.n;
    d.V[t] = d.in_x[i + j * d.ldi];
    if (d.in_hx) d.HV[t] = d.in_hx[i + j * d.ldi];
  }
}

// Y[:, j] = V[:, :m] C[:, col_j], HY likewise from HV. mode 0: the active
// bands into X / HX (live k points); 1: every band into X / HX (contracting k
// points); 2: every band into the caller's outputs (live k points); 3: the
// whitening V L^-H (coefficients in S) into X / HX (k points in rank repair);
// 4: V[:, m + j] -= V[:, :m] G[:, j] for the repair's projection (G in E,
// ld m).
__global__ __launch_bounds__(kKT) void k_kd_rotate(const KDesc* __restrict__ D, const KStat* __restrict__ St,
                                                   int mode, int nb, int cap) {
  __shared__ double2 cs[kKCap][8];
  const KDesc& d = D[blockIdx.z];
  const KStat& st = St[blockIdx.z];
  const bool on = mode == 1 ? d.contract : mode >= 3 ? (d.stage2 && st.repair) : d.live;
  if (!on) return;
  const int ncols = mode == 0 ? d.nact : mode == 3 ? d.m : mode == 4 ? st.kept : nb;
  const int j0 = blockIdx.y * 8;
  if (j0 >= ncols) return;
  const int m = d.m;
  const double2* G = reinterpret_cast<const double2*>(d.E);
  for (int t = threadIdx.x; t < m * 8; t += blockDim.x) {
    const int l = t >> 3, jj = t & 7, j = j0 + jj;
    double2 c = make_double2(0.0, 0.0);
    if (j < ncols) {
      if (mode == 3) c = d.S[l + int64_t(j) * cap];
      else if (mode == 4) c = G[l + int64_t(j) * m];
      else c = d.C[l + int64_t(mode == 0 ? d.act[j] : j) * cap];
    }
    cs[l][jj] = c;
  }
  __syncthreads();
  const int64_t nc = d.n / 2;
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= nc) return;
  const double2* V = reinterpret_cast<const double2*>(d.V);
  const double2* HV = reinterpret_cast<const double2*>(d.HV);
  double2 ax[8], ah[8];
#pragma unroll
  for (int jj = 0; jj < 8; ++jj) ax[jj] = ah[jj] = make_double2(0.0, 0.0);
  for (int l = 0; l < m; ++l) {
    const double2 v = V[r + l * nc], hv = HV[r + l * nc];
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      ax[jj] = zadd(ax[jj], zmul(v, cs[l][jj]));
      ah[jj] = zadd(ah[jj], zmul(hv, cs[l][jj]));
    }
  }
  if (mode == 4) {
    double2* E = reinterpret_cast<double2*>(d.V) + int64_t(m) * nc;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj)
      if (j0 + jj < ncols) E[r + (j0 + jj) * nc] = zsub(E[r + (j0 + jj) * nc], ax[jj]);
    return;
  }
  double2* Y = reinterpret_cast<double2*>(mode == 2 ? d.out_x : d.X);
  double2* HY = reinterpret_cast<double2*>(mode == 2 ? d.out_hx : d.HX);
  const int64_t ldy = mode == 2 ? d.ldo / 2 : nc;
#pragma unroll
  for (int jj = 0; jj < 8; ++jj) {
    if (j0 + jj < ncols) {
      Y[r + (j0 + jj) * ldy] = ax[jj];
      if (HY) HY[r + (j0 + jj) * ldy] = ah[jj];
    }
  }
}

// V[:, :c] = X[:, :c], HV[:, :c] = HX[:, :c]: c = nb for contracting k points
// (mode 0), c = m for k points in rank repair (mode 1).
__global__ void k_kd_take(const KDesc* __restrict__ D, const KStat* __restrict__ St, int mode, int nb) {
  const KDesc& d = D[blockIdx.y];
  const KStat& st = St[blockIdx.y];
  if (mode == 0 ? !d.contract : !(d.stage2 && st.repair)) return;
  const int64_t total = d.n * (mode == 0 ? nb : d.m);
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += int64_t(gridDim.x) * blockDim.x) {
    d.V[t] = d.X[t];
    d.HV[t] = d.HX[t];
  }
}

// QE smooth denominator (ref_davidson/stream.rs:16-22) and the TPA weight
// (crates/dft-solvers/src/lib.rs:197-207), as in the Gamma solver.
__device__ __forceinline__ double kd_sqrt_denominator(double h, double e) {
  const double x = 2.0 * (h - e);
  const double z = fabs(x - 1.0);
  return 0.25 * (2.0 * fmax(x, 1.0) + 1.0 / (sqrt(1.0 + z * z) + z));
}
__device__ __forceinline__ double kd_tpa(double kinetic, double scale) {
  const double x = kinetic / scale;
  if (x <= 1.0) {
    const double p = 27.0 + x * (18.0 + x * (12.0 + x * 8.0));
    return p / (p + 16.0 * x * x * x * x);
  }
  const double y = scale / kinetic;
  const double p = y * (8.0 + y * (12.0 + y * (18.0 + y * 27.0)));
  return p / (p + 16.0);
}

// One workgroup per (active column j, k point): e = (hx - theta x) / d (QE)
// or -(hx - theta x) K (TPA) into E[:, j], |e|^2 into norms.
__global__ __launch_bounds__(kKT) void k_kd_corrections(const KDesc* __restrict__ D, const KStat* __restrict__ St,
                                                        int mode, double* __restrict__ norms) {
  __shared__ double red[kKT / 64];
  const KDesc& d = D[blockIdx.y];
  const int j = blockIdx.x;
  if (!d.live || j >= d.nact) return;
  const double th = St[blockIdx.y].values[d.act[j]];
  const int64_t n = d.n;
  const double* x = d.X + j * n;
  const double* hx = d.HX + j * n;
  double* e = d.E + j * n;
  double scale = 0.0;
  if (mode == 1) {
    double a = 0.0, b = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
      const double w = x[i] * x[i];
      a += w * d.kinetic[i];
      b += w;
    }
    const double sa = kd_sum(a, red), sb = kd_sum(b, red);
    scale = fmax(sa / sb, 0.1);
  }
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double r = hx[i] - th * x[i];
    const double v = mode == 1 ? -r * kd_tpa(d.kinetic[i], scale) : r / kd_sqrt_denominator(d.diag[i], th);
    e[i] = v;
    acc += v * v;
  }
  const double s = kd_sum(acc, red);
  if (threadIdx.x == 0) norms[int64_t(blockIdx.y) * kKB + j] = s;
}

// One workgroup per k point: drop corrections with |e| <= 32 eps
// (ref_davidson/stream.rs:81-143), scale the rest and append them to V at m
// when they fit the capacity. stat.kept = the count kept.
__global__ __launch_bounds__(kKT) void k_kd_normalize(const KDesc* __restrict__ D, KStat* __restrict__ St,
                                                      const double* __restrict__ norms, int cap) {
  __shared__ int list[kKB];
  __shared__ double scale[kKB];
  __shared__ int kept;
  const KDesc& d = D[blockIdx.x];
  if (!d.live) return;
  KStat& st = St[blockIdx.x];
  if (threadIdx.x == 0) {
    int k = 0, bad = 0;
    for (int j = 0; j < d.nact; ++j) {
      const double nrm = sqrt(norms[int64_t(blockIdx.x) * kKB + j]);
      if (!isfinite(nrm)) {
        bad = 1;
        continue;
      }
      if (nrm <= 32.0 * kEps) continue;
      list[k] = j;
      scale[k] = 1.0 / nrm;
      ++k;
    }
    kept = k;
    st.kept = k;
    st.nonfinite_e = bad;
  }
  __syncthreads();
  if (d.m + kept > cap) return;
  const int64_t n = d.n;
  for (int q = 0; q < kept; ++q) {
    const double* e = d.E + list[q] * n;
    double* v = d.V + (d.m + q) * n;
    const double c = scale[q];
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) v[i] = e[i] * c;
  }
}

// Hermitian Gram border (ref_davidson_complex.rs:235-276): G = V[:, :to]^H
// B[:, from:to] with B = V (which 0: S) or HV (which 1: H), written to both
// triangles of M (exactly Hermitian, real diagonal), over the range of
// kd_range(mode). which 2: the rank repair's projection G = V[:, :m]^H
// V[:, m:m + kept] into E (ld m, k points in repair). A 16 x 16 tile of G per
// workgroup, 64 rows at a time through LDS; for which 1 a non-finite H·V
// entry sets stat.nonfinite_h.
__global__ __launch_bounds__(kKT) void k_kd_border(const KDesc* __restrict__ D, KStat* __restrict__ St, int which,
                                                   int mode, int nb, int cap) {
  constexpr int R = 64;
  __shared__ double2 vs[16][R + 1];
  __shared__ double2 bs[16][R + 1];
  const KDesc& d = D[blockIdx.z];
  KStat& st = St[blockIdx.z];
  int from, to, rows;
  if (which == 2) {
    if (!d.stage2 || !st.repair || st.kept <= 0) return;
    from = d.m;
    to = d.m + st.kept;
    rows = d.m;
  } else {
    if (!kd_range(d, st, mode, nb, from, to)) return;
    rows = to;
  }
  const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
  if (i0 >= rows || j0 >= to - from) return;
  const int ti = threadIdx.x & 15, tj = threadIdx.x >> 4;
  const int64_t nc = d.n / 2;
  const double2* V = reinterpret_cast<const double2*>(d.V);
  const double2* B = reinterpret_cast<const double2*>(which == 1 ? d.HV : d.V);
  double2 acc = make_double2(0.0, 0.0);
  int bad = 0;
  for (int64_t r0 = 0; r0 < nc; r0 += R) {
    for (int t = threadIdx.x; t < 16 * R; t += blockDim.x) {
      const int c = t / R, rr = t % R;
      const int64_t r = r0 + rr;
      const int iv = i0 + c, jb = from + j0 + c;
      vs[c][rr] = (iv < rows && r < nc) ? V[r + iv * nc] : make_double2(0.0, 0.0);
      double2 b = (jb < to && r < nc) ? B[r + jb * nc] : make_double2(0.0, 0.0);
      bad |= !(isfinite(b.x) && isfinite(b.y));
      bs[c][rr] = b;
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < R; ++rr) acc = zadd(acc, zcmul(vs[ti][rr], bs[tj][rr]));
    __syncthreads();
  }
  if (which == 1 && bad) st.nonfinite_h = 1;
  const int i = i0 + ti, jc = from + j0 + tj;
  if (i >= rows || jc >= to) return;
  if (which == 2) {
    reinterpret_cast<double2*>(d.E)[i + int64_t(jc - from) * rows] = acc;
    return;
  }
  double2* M = which == 1 ? d.H : d.S;
  if (i < jc) {
    M[i + int64_t(jc) * cap] = acc;
    M[jc + int64_t(i) * cap] = zconj(acc);
  } else if (i == jc) {
    M[i + int64_t(i) * cap] = make_double2(acc.x, 0.0);
  }
}

// Bordered complex Cholesky, one workgroup per k point
// (ComplexOverlapFactor::extend / factor, crates/dft-numerics/src/linalg.rs:
// 234-340): L21 = (L11^-1 S12)^H, L22 = chol(S22 - L21 L21^H); rejected when
// a new diagonal of S is not positive, the Cholesky fails, a pivot ratio
// |L_ii|^2 / S_ii falls below eps / 1e-12 or an entry is not finite. Rows
// [from, to) of L are written; the leading factor is untouched. mode as in
// k_kd_border. The verdict goes to stat.factor_bad.
__global__ __launch_bounds__(kKT) void k_kd_factor(const KDesc* __restrict__ D, KStat* __restrict__ St, int mode,
                                                   int nb, int cap) {
  extern __shared__ double2 fsh[];
  const KDesc& d = D[blockIdx.x];
  KStat& st = St[blockIdx.x];
  int from, to;
  if (!kd_range(d, st, mode, nb, from, to)) return;
  const int k = to - from;
  double2* Y = fsh;            // from x k: L11^-1 S12, column-major (ld from)
  double2* Z = Y + from * k;   // k x k: the Schur complement, then L22 (ld k)
  __shared__ int bad;
  const double2* S = d.S;
  double2* L = d.L;
  if (threadIdx.x == 0) {
    int b = 0;
    for (int i = from; i < to; ++i) b |= !(S[i + int64_t(i) * cap].x > 0.0);
    bad = b;
  }
  // Forward substitution, one thread per column of S12.
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    for (int i = 0; i < from; ++i) {
      double2 z = S[i + int64_t(from + j) * cap];
      for (int p = 0; p < i; ++p) z = zsub(z, zmul(L[i + int64_t(p) * cap], Y[p + j * from]));
      const double lii = L[i + int64_t(i) * cap].x;
      Y[i + j * from] = make_double2(z.x / lii, z.y / lii);
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < k * k; t += blockDim.x) {
    const int i = t % k, j = t / k;
    double2 z = S[from + i + int64_t(from + j) * cap];
    for (int p = 0; p < from; ++p) z = zsub(z, zcmul(Y[p + i * from], Y[p + j * from]));
    Z[i + j * k] = z;
  }
  __syncthreads();
  // Right-looking Cholesky of Z (lower triangle), a column per step.
  for (int j = 0; j < k; ++j) {
    if (threadIdx.x == 0) {
      const double djj = Z[j + j * k].x;
      const double l = sqrt(djj);
      const double sjj = S[from + j + int64_t(from + j) * cap].x;
      if (!(djj > 0.0) || !isfinite(l) || !(l * l / sjj >= kPivotMin)) bad = 1;
      Z[j + j * k] = make_double2(l, 0.0);
    }
    __syncthreads();
    const double l = Z[j + j * k].x;
    for (int i = j + 1 + threadIdx.x; i < k; i += blockDim.x) Z[i + j * k] = make_double2(Z[i + j * k].x / l, Z[i + j * k].y / l);
    __syncthreads();
    for (int t = threadIdx.x; t < (k - j - 1) * (k - j - 1); t += blockDim.x) {
      const int i = j + 1 + t % (k - j - 1), c = j + 1 + t / (k - j - 1);
      if (i >= c) Z[i + c * k] = zsub(Z[i + c * k], zmul(Z[i + j * k], zconj(Z[c + j * k])));
    }
    __syncthreads();
  }
  // L21 = Y^H, L22 = Z's lower triangle; finiteness of every new entry.
  int nf = 0;
  for (int t = threadIdx.x; t < k * to; t += blockDim.x) {
    const int i = t % k, j = t / k;  // row from + i, column j
    double2 v;
    if (j < from) v = zconj(Y[j + i * from]);
    else if (j - from <= i) v = Z[i + (j - from) * k];
    else v = make_double2(0.0, 0.0);
    nf |= !(isfinite(v.x) && isfinite(v.y));
    L[from + i + int64_t(j) * cap] = v;
  }
  if (nf) bad = 1;
  __syncthreads();
  if (threadIdx.x == 0) {
    if (mode == 1) st.refresh_bad = bad;
    else if (mode >= 3) st.repair_bad |= bad;
    else st.factor_bad = bad;
  }
}

// The reduced problem of one k point per workgroup
// (ComplexOverlapFactor::solve_lowest_into, linalg.rs:380-440): A = L^-1 H
// L^-H Hermitised, its eigenpairs by cyclic complex Jacobi rotations in
// parallel (round-robin) order until a sweep finds no entry above eps ||A||_F
// (a fixed, deterministic order; the rotations keep the eigenvectors
// orthonormal, equal eigenvalues included), the lowest nb values ascending
// into stat.values and C = L^-H Y. Skipped when the factor was rejected.
__global__ __launch_bounds__(kKT) void k_kd_eigen(const KDesc* __restrict__ D, KStat* __restrict__ St, int mode,
                                                  int nb, int cap) {
  extern __shared__ double2 esh[];
  __shared__ double red[kKT / 64];
  __shared__ double rot[kKCap / 2][4];
  __shared__ int rp[kKCap / 2][2];
  __shared__ int nrot, perm[kKCap];
  const KDesc& d = D[blockIdx.x];
  KStat& st = St[blockIdx.x];
  int m;
  if (mode == 0) {
    int from;
    if (!kd_range(d, st, 5, nb, from, m)) return;  // nothing new (a repair that kept no correction)
  } else {
    if (!d.live) return;
    m = nb;
  }
  if (st.factor_bad || st.nonfinite_h || st.repair_bad) return;
  // Leading dimension m + 1 (odd): a row of a matrix (stride la) spreads over
  // the LDS banks.
  const int la = m + 1;
  double2* Ls = esh;          // m x m
  double2* A = Ls + m * la;   // m x m
  double2* W = A + m * la;    // m x m: L^-1 H, then the eigenvectors Q
  for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
    const int i = t % m, j = t / m;
    Ls[i + j * la] = d.L[i + int64_t(j) * cap];
    W[i + j * la] = d.H[i + int64_t(j) * cap];
  }
  __syncthreads();
  // W = L^-1 H (forward substitution per column), then A = L^-1 W^H.
  for (int j = threadIdx.x; j < m; j += blockDim.x)
    for (int i = 0; i < m; ++i) {
      double2 z = W[i + j * la];
      for (int p = 0; p < i; ++p) z = zsub(z, zmul(Ls[i + p * la], W[p + j * la]));
      W[i + j * la] = zscale(1.0 / Ls[i + i * la].x, z);
    }
  __syncthreads();
  for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
    const int i = t % m, j = t / m;
    A[i + j * la] = zconj(W[j + i * la]);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < m; j += blockDim.x)
    for (int i = 0; i < m; ++i) {
      double2 z = A[i + j * la];
      for (int p = 0; p < i; ++p) z = zsub(z, zmul(Ls[i + p * la], A[p + j * la]));
      A[i + j * la] = zscale(1.0 / Ls[i + i * la].x, z);
    }
  __syncthreads();
  // Hermitise (in W), then back into A; Q = I in W.
  for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
    const int i = t % m, j = t / m;
    const double2 a = A[i + j * la], b = zconj(A[j + i * la]);
    W[i + j * la] = i == j ? make_double2(0.5 * (a.x + b.x), 0.0) : zscale(0.5, zadd(a, b));
  }
  __syncthreads();
  double fro = 0.0;
  for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
    const int i = t % m, j = t / m;
    A[i + j * la] = W[i + j * la];
    W[i + j * la] = i == j ? make_double2(1.0, 0.0) : make_double2(0.0, 0.0);
    fro += zabs2(A[i + j * la]);
  }
  const double thr = kEps * sqrt(kd_sum(fro, red));
  const int mp = m + (m & 1);  // players; an odd m has one idle player per round
  const int pairs = mp / 2;
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (threadIdx.x == 0) nrot = 0;
    __syncthreads();
    for (int round = 0; round < mp - 1; ++round) {
      if (threadIdx.x < pairs) {
        const int t = threadIdx.x;
        const int a = t == 0 ? mp - 1 : (t - 1 + round) % (mp - 1);
        const int b = (mp - 2 - t + round) % (mp - 1);
        int p = min(a, b), q = max(a, b);
        double c = 1.0, s = 0.0, er = 1.0, ei = 0.0;
        if (q < m) {
          const double2 apq = A[p + q * la];
          const double r = sqrt(zabs2(apq));
          if (r > thr) {
            const double app = A[p + p * la].x, aqq = A[q + q * la].x;
            const double ir = 1.0 / r;
            const double th = 0.5 * (aqq - app) * ir;
            const double tt = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(tt * tt + 1.0);
            s = tt * c;
            er = apq.x * ir;
            ei = apq.y * ir;
            atomicAdd(&nrot, 1);
          } else {
            q = -1;
          }
        } else {
          q = -1;
        }
        rot[t][0] = c;
        rot[t][1] = s;
        rot[t][2] = er;
        rot[t][3] = ei;
        rp[t][0] = p;
        rp[t][1] = q;
      }
      __syncthreads();
      // Columns: A <- A U, Q <- Q U with U = [[c, s e], [-s conj(e), c]].
      for (int w = threadIdx.x; w < pairs * m; w += blockDim.x) {
        const int t = w / m, i = w % m;
        const int p = rp[t][0], q = rp[t][1];
        if (q < 0) continue;
        const double c = rot[t][0], s = rot[t][1];
        const double2 e = make_double2(rot[t][2], rot[t][3]);
        const double2 aip = A[i + p * la], aiq = A[i + q * la];
        A[i + p * la] = zsub(zscale(c, aip), zscale(s, zmul(zconj(e), aiq)));
        A[i + q * la] = zadd(zscale(s, zmul(e, aip)), zscale(c, aiq));
        const double2 qip = W[i + p * la], qiq = W[i + q * la];
        W[i + p * la] = zsub(zscale(c, qip), zscale(s, zmul(zconj(e), qiq)));
        W[i + q * la] = zadd(zscale(s, zmul(e, qip)), zscale(c, qiq));
      }
      __syncthreads();
      // Rows: A <- U^H A; the pivot entries are set to their exact values.
      for (int w = threadIdx.x; w < pairs * m; w += blockDim.x) {
        const int t = w / m, j = w % m;
        const int p = rp[t][0], q = rp[t][1];
        if (q < 0) continue;
        const double c = rot[t][0], s = rot[t][1];
        const double2 e = make_double2(rot[t][2], rot[t][3]);
        const double2 apj = A[p + j * la], aqj = A[q + j * la];
        double2 np = zsub(zscale(c, apj), zscale(s, zmul(e, aqj)));
        double2 nq = zadd(zscale(s, zmul(zconj(e), apj)), zscale(c, aqj));
        if (j == p) {
          np.y = 0.0;
          nq = make_double2(0.0, 0.0);
        } else if (j == q) {
          np = make_double2(0.0, 0.0);
          nq.y = 0.0;
        }
        A[p + j * la] = np;
        A[q + j * la] = nq;
      }
      __syncthreads();
    }
    if (nrot == 0) break;
    __syncthreads();
  }
  // Ascending order (ties by index), the lowest nb.
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double di = A[i + i * la].x;
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      const double dj = A[j + j * la].x;
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    perm[rank] = i;
  }
  __syncthreads();
  int nf = 0;
  for (int r = threadIdx.x; r < nb; r += blockDim.x) {
    const double v = A[perm[r] + perm[r] * la].x;
    nf |= !isfinite(v);
    st.values[r] = v;
  }
  // C = L^-H Y (back substitution per wanted column).
  for (int r = threadIdx.x; r < nb; r += blockDim.x) {
    const int src = perm[r];
    double2* c = d.C + int64_t(r) * cap;
    for (int i = m - 1; i >= 0; --i) {
      double2 z = W[i + src * la];
      for (int p = i + 1; p < m; ++p) z = zsub(z, zcmul(Ls[p + i * la], c[p]));
      c[i] = zscale(1.0 / Ls[i + i * la].x, z);
      nf |= !(isfinite(c[i].x) && isfinite(c[i].y));
    }
  }
  if (nf) st.eig_bad = 1;
}

// Restart of contracting k points (workspace.rs:772-935, ideal_reset.rs):
// C = I; the column norms of the restarted block against 1 (the guard), the
// contraction count and the disabled flag choose the ideal reset (S = I,
// H = diag(theta), L = I) or a measured refresh (stat.refresh = 1).
__global__ __launch_bounds__(kKT) void k_kd_restart(const KDesc* __restrict__ D, KStat* __restrict__ St, int nb,
                                                    int cap) {
  __shared__ double red[kKT / 64];
  __shared__ int ok;
  const KDesc& d = D[blockIdx.x];
  if (!d.contract) return;
  KStat& st = St[blockIdx.x];
  const int contractions = st.contractions + 1;
  const bool periodic = contractions % kRefreshEvery == 0;
  const bool disabled = st.disabled != 0;
  if (threadIdx.x == 0) ok = 1;
  __syncthreads();
  if (!disabled && !periodic) {
    for (int j = 0; j < nb; ++j) {
      const double* v = d.V + j * d.n;
      double a = 0.0;
      for (int64_t i = threadIdx.x; i < d.n; i += blockDim.x) a += v[i] * v[i];
      const double s = kd_sum(a, red);
      if (threadIdx.x == 0 && !(fabs(s - 1.0) <= kOrthoLimit)) ok = 0;
    }
  }
  __syncthreads();
  const bool ideal = !disabled && !periodic && ok;
  for (int t = threadIdx.x; t < nb * nb; t += blockDim.x) {
    const int i = t % nb, j = t / nb;
    const double2 one = make_double2(i == j ? 1.0 : 0.0, 0.0);
    d.C[i + int64_t(j) * cap] = one;
    if (ideal) {
      d.S[i + int64_t(j) * cap] = one;
      d.L[i + int64_t(j) * cap] = one;
      d.H[i + int64_t(j) * cap] = make_double2(i == j ? st.values[i] : 0.0, 0.0);
    }
  }
  if (threadIdx.x == 0) {
    st.contractions = contractions;
    st.refresh = ideal ? 0 : 1;
    if (ideal) st.ideal_resets += 1;
    else if (disabled) {
    } else if (periodic) st.periodic += 1;
    else {
      st.disabled = 1;
      st.guard_fallbacks += 1;
    }
  }
}

// After a measured refresh that is still guarded: max |S - I| over the
// restarted block above 1e-12 (or not finite) disables the ideal reset.
__global__ __launch_bounds__(kKT) void k_kd_defect(const KDesc* __restrict__ D, KStat* __restrict__ St, int nb,
                                                   int cap) {
  __shared__ double red[kKT / 64];
  const KDesc& d = D[blockIdx.x];
  KStat& st = St[blockIdx.x];
  if (!d.contract || !st.refresh || st.disabled) return;
  double mx = 0.0;
  for (int t = threadIdx.x; t < nb * nb; t += blockDim.x) {
    const int i = t % nb, j = t / nb;
    double2 v = d.S[i + int64_t(j) * cap];
    if (i == j) v.x -= 1.0;
    const double a = sqrt(zabs2(v));
    mx = isfinite(a) ? fmax(mx, a) : INFINITY;
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < int(blockDim.x >> 6); ++w) r = fmax(r, red[w]);
    if (!(r <= kOrthoLimit)) {
      st.disabled = 1;
      st.guard_fallbacks += 1;
    }
  }
}

__global__ void k_kd_clear_flags(KStat* __restrict__ St, int count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) {
    St[k].kept = 0;
    St[k].factor_bad = 0;
    St[k].eig_bad = 0;
    St[k].nonfinite_h = 0;
    St[k].nonfinite_e = 0;
    St[k].refresh_bad = 0;
    St[k].repair = 0;
    St[k].repair_bad = 0;
    St[k].kept2 = 0;
  }
}

// Rank repair, first step (ref_davidson_complex.rs:826-850 with
// ComplexOverlapFactor::basis_transform_into, linalg.rs:342-360), one
// workgroup per k point whose border the factor rejected: the Ritz block's
// coordinates in the whitened basis V L^-H, C <- L^H C, and the whitening
// R = L^-H into S (the overlap is measured again after the rotation).
__global__ __launch_bounds__(kKT) void k_kd_repair_prep(const KDesc* __restrict__ D, KStat* __restrict__ St,
                                                        int nb, int cap) {
  extern __shared__ double2 rsh[];
  const KDesc& d = D[blockIdx.x];
  KStat& st = St[blockIdx.x];
  if (!d.stage2 || !st.factor_bad) return;
  const int m = d.m;
  double2* Ls = rsh;        // m x m
  double2* Cs = Ls + m * m;  // m x nb
  for (int t = threadIdx.x; t < m * m; t += blockDim.x) Ls[t] = d.L[t % m + int64_t(t / m) * cap];
  for (int t = threadIdx.x; t < m * nb; t += blockDim.x) Cs[t] = d.C[t % m + int64_t(t / m) * cap];
  __syncthreads();
  for (int t = threadIdx.x; t < m * nb; t += blockDim.x) {
    const int i = t % m, j = t / m;
    double2 z = make_double2(0.0, 0.0);
    for (int p = i; p < m; ++p) z = zadd(z, zcmul(Ls[p + i * m], Cs[p + j * m]));
    d.C[i + int64_t(j) * cap] = z;
  }
  // R = L^-H, upper triangular: one thread per column, back substitution.
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    double2* R = d.S + int64_t(j) * cap;
    for (int i = m - 1; i > j; --i) R[i] = make_double2(0.0, 0.0);
    R[j] = make_double2(1.0 / Ls[j + j * m].x, 0.0);
    for (int i = j - 1; i >= 0; --i) {
      double2 z = make_double2(0.0, 0.0);
      for (int p = i + 1; p <= j; ++p) z = zadd(z, zcmul(Ls[p + i * m], R[p]));
      R[i] = zscale(-1.0 / Ls[i + i * m].x, z);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    st.repair = 1;
    st.factor_bad = 0;
  }
}

// Rank repair, last step (ref_davidson_complex.rs:862-888), one workgroup per
// k point in repair: the projected corrections V[:, m:m + kept] in order,
// each made orthogonal to the ones kept before it (two classical
// Gram-Schmidt passes), dropped at norm <= 1e-12 or when the basis is full,
// else normalised into the next free column. stat.kept2 = the count kept.
__global__ __launch_bounds__(kKT) void k_kd_repair_gs(const KDesc* __restrict__ D, KStat* __restrict__ St,
                                                      int cap) {
  __shared__ double red[kKT / 64][2 * kKB];
  __shared__ double2 zs[kKB];
  __shared__ double nrm2;
  const KDesc& d = D[blockIdx.x];
  KStat& st = St[blockIdx.x];
  if (!d.stage2 || !st.repair) return;
  const int m = d.m, kin = st.kept;
  const int64_t nc = d.n / 2;
  double2* Q = reinterpret_cast<double2*>(d.V) + int64_t(m) * nc;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int kept = 0;
  for (int j = 0; j < kin; ++j) {
    double2* e = Q + int64_t(j) * nc;
    for (int pass = 0; pass < 2 && kept > 0; ++pass) {
      double2 acc[kKB];
#pragma unroll
      for (int a = 0; a < kKB; ++a) acc[a] = make_double2(0.0, 0.0);
      for (int64_t r = threadIdx.x; r < nc; r += blockDim.x) {
        const double2 er = e[r];
#pragma unroll
        for (int a = 0; a < kKB; ++a)
          if (a < kept) acc[a] = zadd(acc[a], zcmul(Q[r + a * nc], er));
      }
#pragma unroll
      for (int a = 0; a < kKB; ++a) {
        if (a < kept) {
          double x = acc[a].x, y = acc[a].y;
          for (int off = 32; off > 0; off >>= 1) {
            x += __shfl_down(x, off, 64);
            y += __shfl_down(y, off, 64);
          }
          if (lane == 0) {
            red[wave][2 * a] = x;
            red[wave][2 * a + 1] = y;
          }
        }
      }
      __syncthreads();
      if (threadIdx.x < kept) {
        double x = 0.0, y = 0.0;
        for (int w = 0; w < kKT / 64; ++w) {
          x += red[w][2 * threadIdx.x];
          y += red[w][2 * threadIdx.x + 1];
        }
        zs[threadIdx.x] = make_double2(x, y);
      }
      __syncthreads();
      for (int64_t r = threadIdx.x; r < nc; r += blockDim.x) {
        double2 v = e[r];
        for (int a = 0; a < kept; ++a) v = zsub(v, zmul(zs[a], Q[r + a * nc]));
        e[r] = v;
      }
      __syncthreads();
    }
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < nc; r += blockDim.x) s += zabs2(e[r]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) red[wave][0] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kKT / 64; ++w) t += red[w][0];
      nrm2 = t;
    }
    __syncthreads();
    const double nrm = sqrt(nrm2);
    if (!(nrm > 1e-12) || m + kept == cap) continue;  // DEPENDENCE_TOLERANCE, linalg.rs:12
    const double inv = 1.0 / nrm;
    double2* q = Q + int64_t(kept) * nc;
    for (int64_t r = threadIdx.x; r < nc; r += blockDim.x) q[r] = zscale(inv, e[r]);
    __syncthreads();
    ++kept;
  }
  if (threadIdx.x == 0) st.kept2 = kept;
}

size_t factor_lds(int cap) { return size_t(cap) * cap * 2 * sizeof(double2); }
size_t eigen_lds(int cap) { return size_t(3) * cap * (cap + 1) * sizeof(double2); }

// The two LDS-resident kernels may take their dynamic LDS at the largest
// capacity (their static LDS comes on top, within the 160 KB of a workgroup).
void set_lds_limits() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    err = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_kd_eigen), hipFuncAttributeMaxDynamicSharedMemorySize,
                              int(eigen_lds(kKCap)));
    if (err == hipSuccess)
      err = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_kd_factor),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(factor_lds(kKCap)));
    if (err == hipSuccess)
      err = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_kd_repair_prep),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(factor_lds(kKCap)));
    if (err != hipSuccess) (void)hipGetLastError();
  });
  if (err != hipSuccess) fail(BD_ERR_DEVICE, std::string("lock-step k Davidson LDS attribute: ") + hipGetErrorString(err));
}

struct Buf {
  double* p = nullptr;
  void alloc(size_t count) { BD_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(double))); }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};
}  // namespace

// Per-k host state of one lock-step solve.
struct KPoint {
  int m = 0, iterations = 0, nact = 0;
  bool done = false, ejected = false, contract = false;
  std::vector<int> active;
  std::vector<char> converged;
  std::vector<double> values, old;
  bd_davidson_record rec{};
};

struct KDavidson {
  std::vector<bd_gamma_op*> ops;
  int count, nb, cap;
  int64_t nmax = 0;
  hipStream_t main = nullptr;
  hipEvent_t ev = nullptr;
  Buf arena, norms;
  KStat* stat = nullptr;
  KDesc* desc = nullptr;
  std::vector<KStat> hstat;
  std::vector<KDesc> hdesc;
  DevBuf<KWork> works;
  std::vector<KWork> hworks;
  // Pinned staging for the uploads: a slot is reused only after the stream
  // has been synchronised since it was filled.
  static constexpr int kSlots = 8;
  char* pinned = nullptr;
  size_t slot_bytes = 0;
  int slots_used = 0;
  std::vector<KPoint> kp;
  FusedShape fshape;
  size_t flds = 0;

  KDavidson(const std::vector<bd_gamma_op*>& ops_, int nb_, int cap_)
      : ops(ops_), count(int(ops_.size())), nb(nb_), cap(cap_) {
    const GammaOp& o0 = ops[0]->op;
    fshape = o0.fshape;
    size_t per = 0;
    for (auto* h : ops) {
      const GammaOp& o = h->op;
      if (!o.kb || !o.fused) fail(BD_ERR_INVALID_PARAMETER, "lock-step k Davidson needs k-point ops on the fused path");
      if (o.sys != o0.sys || o.precision != o0.precision)
        fail(BD_ERR_INVALID_PARAMETER, "lock-step k Davidson ops must share the system and precision");
      if (!o.fused_valid(fshape)) fail(BD_ERR_INVALID_PARAMETER, "lock-step k Davidson ops need a common fused shape");
      if (nb > o.capacity) fail(BD_ERR_INVALID_PARAMETER, "bands exceed the operator capacity");
      if (2 * int64_t(cap) > o.n) fail(BD_ERR_INVALID_PARAMETER, "capacity exceeds the plane waves of a k point");
      flds = std::max(flds, o.fused_lds(fshape));
      nmax = std::max(nmax, o.n);
      per += size_t(o.n) * (4 * cap + nb) + size_t(2) * (3 * cap * cap + cap * nb);
    }
    arena.alloc(per);
    norms.alloc(size_t(count) * kKB);
    BD_HIP(hipMalloc(&stat, sizeof(KStat) * count));
    BD_HIP(hipMalloc(&desc, sizeof(KDesc) * count));
    hstat.resize(size_t(count));
    hdesc.resize(size_t(count));
    slot_bytes = (std::max(sizeof(KDesc), sizeof(KWork)) * count + 255) / 256 * 256;
    BD_HIP(hipHostMalloc(reinterpret_cast<void**>(&pinned), slot_bytes * kSlots, hipHostMallocDefault));
    works.alloc(size_t(count));
    hworks.resize(size_t(count));
    kp.resize(size_t(count));
    BD_HIP(hipStreamCreateWithFlags(&main, hipStreamNonBlocking));
    BD_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    std::memset(hdesc.data(), 0, sizeof(KDesc) * count);
    double* p = arena.p;
    for (int k = 0; k < count; ++k) {
      const GammaOp& o = ops[k]->op;
      KDesc& d = hdesc[k];
      d.n = o.n;
      d.V = p;
      p += o.n * cap;
      d.HV = p;
      p += o.n * cap;
      d.X = p;
      p += o.n * cap;
      d.HX = p;
      p += o.n * cap;
      d.E = p;
      p += o.n * nb;
      d.S = reinterpret_cast<double2*>(p);
      p += 2 * cap * cap;
      d.H = reinterpret_cast<double2*>(p);
      p += 2 * cap * cap;
      d.L = reinterpret_cast<double2*>(p);
      p += 2 * cap * cap;
      d.C = reinterpret_cast<double2*>(p);
      p += 2 * cap * nb;
      d.kinetic = o.d_kinetic.p;
    }
    set_lds_limits();
  }
  ~KDavidson() {
    if (ev) (void)hipEventDestroy(ev);
    if (main) (void)hipStreamDestroy(main);
    if (stat) (void)hipFree(stat);
    if (desc) (void)hipFree(desc);
    if (pinned) (void)hipHostFree(pinned);
  }

  void sync() {
    BD_HIP(hipStreamSynchronize(main));
    slots_used = 0;
  }
  void upload(void* dev, const void* host, size_t bytes) {
    if (slots_used == kSlots) sync();
    char* slot = pinned + slot_bytes * slots_used++;
    std::memcpy(slot, host, bytes);
    BD_HIP(hipMemcpyAsync(dev, slot, bytes, hipMemcpyHostToDevice, main));
  }
  void push_desc() { upload(desc, hdesc.data(), sizeof(KDesc) * count); }
  void pull_stat() {
    BD_HIP(hipMemcpyAsync(hstat.data(), stat, sizeof(KStat) * count, hipMemcpyDeviceToHost, main));
    sync();
  }
  unsigned row_blocks() const { return unsigned((nmax / 2 + kKT - 1) / kKT); }

  // H·V[:, m:m+kk] of every stage-2 k point in one fused launch.
  void apply(bool start) {
    int total = 0;
    for (int k = 0; k < count; ++k) {
      const KDesc& d = hdesc[k];
      const int cols = start ? (d.live ? nb : 0) : (d.stage2 ? d.kk : 0);
      const int at = start ? 0 : d.m;
      hworks[k] = ops[k]->op.kwork(d.V + int64_t(at) * d.n, d.n, cols, d.HV + int64_t(at) * d.n, d.n, total);
      total += cols;
      if (cols) {
        kp[k].rec.h_calls += 1;
        kp[k].rec.h_columns += cols;
      }
    }
    if (total == 0) return;
    upload(works.p, hworks.data(), sizeof(KWork) * count);
    const KWork none{};
    fused_launch(ops[0]->op, fshape, flds, works.p, none, count, total, main);
  }

  void border(int which, int mode) {
    hipLaunchKernelGGL(k_kd_border, dim3((cap + 15) / 16, (cap + 15) / 16, count), dim3(kKT), 0, main, desc, stat,
                       which, mode, nb, cap);
  }
  void factor(int mode) {
    hipLaunchKernelGGL(k_kd_factor, dim3(count), dim3(kKT), factor_lds(cap), main, desc, stat, mode, nb, cap);
  }
  void eigen(int mode) {
    hipLaunchKernelGGL(k_kd_eigen, dim3(count), dim3(kEigT), eigen_lds(cap), main, desc, stat, mode, nb, cap);
  }

  double band_threshold(const bd_davidson_options& o, int i) const {
    if (o.first_step || i < o.occupied) return o.threshold;
    return std::max(5.0 * o.threshold, 5e-6);
  }

  // Solve every k point; per-k results into the caller's arrays. Returns the
  // index of a failed k point and its message, or -1.
  int solve(const double* const* x0, const int64_t* ldx0, const double* const* hx0, const double* const* diag,
            const bd_davidson_options& o, double* eig, double* const* xout, double* const* hxout,
            const int64_t* ldx, int32_t* conv, bd_davidson_record* records, bd_status* statuses, std::string* msg) {
    const auto t0 = std::chrono::steady_clock::now();
    for (auto* h : ops) {
      BD_HIP(hipEventRecord(ev, h->op.stream));
      BD_HIP(hipStreamWaitEvent(main, ev, 0));
    }
    const int mode = o.preconditioner == BD_PRECOND_QE ? 0 : 1;
    for (int k = 0; k < count; ++k) {
      KDesc& d = hdesc[k];
      d.diag = diag[k];
      d.in_x = x0[k];
      d.in_hx = hx0 ? hx0[k] : nullptr;
      d.ldi = ldx0[k];
      d.m = nb;
      d.kk = 0;
      d.live = 1;
      d.stage2 = 0;
      d.contract = 0;
      d.nact = nb;
      for (int j = 0; j < nb; ++j) d.act[j] = j;
      KPoint& p = kp[k];
      p = KPoint();
      p.m = nb;
      p.active.resize(nb);
      for (int j = 0; j < nb; ++j) p.active[j] = j;
      p.converged.assign(nb, 0);
      p.values.assign(nb, 0.0);
      p.old.assign(nb, 0.0);
      p.rec.mlp_fallback = -1;
      statuses[k] = BD_OK;
    }
    BD_HIP(hipMemsetAsync(stat, 0, sizeof(KStat) * count, main));
    push_desc();
    hipLaunchKernelGGL(k_kd_copy_in, dim3(64, count), dim3(256), 0, main, desc, nb);
    if (!hx0) apply(true);
    border(0, 2);
    factor(2);
    border(1, 2);
    eigen(1);
    pull_stat();
    double t_dense = 0.0;
    for (int k = 0; k < count; ++k) {
      const KStat& st = hstat[k];
      if (st.factor_bad) {
        *msg = "complex Davidson start is numerically dependent";
        statuses[k] = BD_ERR_INVALID_PARAMETER;
        return k;
      }
      if (st.nonfinite_h || st.eig_bad) {
        *msg = st.nonfinite_h ? "complex H application failed (non-finite output)" : "complex reduced eigensolve failed";
        statuses[k] = BD_ERR_APPLICATION;
        return k;
      }
      std::copy(st.values, st.values + nb, kp[k].values.begin());
    }
    std::vector<int> ejected;
    for (int step = 0; step < o.max_iterations; ++step) {
      int live = 0, contracting = 0;
      for (int k = 0; k < count; ++k) {
        KPoint& p = kp[k];
        KDesc& d = hdesc[k];
        d.live = !p.done && !p.ejected;
        d.stage2 = 0;
        d.kk = 0;
        d.m = p.m;
        d.contract = d.live && p.contract;
        if (d.contract) p.rec.restarts += 1;
        p.contract = false;
        d.nact = int(p.active.size());
        for (int j = 0; j < d.nact; ++j) d.act[j] = p.active[j];
        live += d.live;
        contracting += d.contract;
      }
      if (!live) break;
      hipLaunchKernelGGL(k_kd_clear_flags, dim3((count + 255) / 256), dim3(256), 0, main, stat, count);
      // Contraction (ref_davidson_complex.rs:1004-1082): X = V C over the
      // whole basis, V = X, then the ideal reset or a measured refresh of S,
      // H and the factor.
      if (contracting) {
        push_desc();
        hipLaunchKernelGGL(k_kd_rotate, dim3(row_blocks(), (nb + 7) / 8, count), dim3(kKT), 0, main, desc, stat, 1,
                           nb, cap);
        hipLaunchKernelGGL(k_kd_take, dim3(64, count), dim3(256), 0, main, desc, stat, 0, nb);
        hipLaunchKernelGGL(k_kd_restart, dim3(count), dim3(kKT), 0, main, desc, stat, nb, cap);
        border(0, 1);
        border(1, 1);
        factor(1);
        hipLaunchKernelGGL(k_kd_defect, dim3(count), dim3(kKT), 0, main, desc, stat, nb, cap);
        for (int k = 0; k < count; ++k)
          if (hdesc[k].contract) {
            hdesc[k].m = nb;
            kp[k].m = nb;
          }
      }
      // The Ritz columns of the active bands, their corrections (dropped and
      // normalised on the device, appended at m), the border of S and the
      // bordered factor; then one read-back for every k point.
      push_desc();
      hipLaunchKernelGGL(k_kd_rotate, dim3(row_blocks(), (nb + 7) / 8, count), dim3(kKT), 0, main, desc, stat, 0, nb,
                         cap);
      hipLaunchKernelGGL(k_kd_corrections, dim3(nb, count), dim3(kKT), 0, main, desc, stat, mode, norms.p);
      hipLaunchKernelGGL(k_kd_normalize, dim3(count), dim3(kKT), 0, main, desc, stat, norms.p, cap);
      border(0, 0);
      factor(0);
      pull_stat();
      int stage2 = 0, repairs = 0;
      for (int k = 0; k < count; ++k) {
        KDesc& d = hdesc[k];
        if (!d.live) continue;
        const KStat& st = hstat[k];
        KPoint& p = kp[k];
        if (d.contract && (st.refresh_bad || st.nonfinite_h)) {
          *msg = "complex restart metric is dependent";
          statuses[k] = BD_ERR_APPLICATION;
          return k;
        }
        if (st.nonfinite_e) {
          *msg = "nonfinite complex correction";
          statuses[k] = BD_ERR_APPLICATION;
          return k;
        }
        d.kk = st.kept;
        if (d.kk > 0 && p.m + d.kk > cap) {
          // Not staged (the capacity rule keeps m + active <= cap, so this
          // does not happen): this k point continues in the single-k
          // complex Davidson from its Ritz block.
          p.ejected = true;
          ejected.push_back(k);
          d.kk = 0;
          continue;
        }
        d.stage2 = d.kk > 0;
        stage2 += d.stage2;
        repairs += d.stage2 && st.factor_bad;
      }
      auto td = std::chrono::steady_clock::now();
      if (stage2) {
        push_desc();
        if (repairs) {
          // Rank repair of the k points whose border the factor rejected
          // (ref_davidson_complex.rs:822-900): whiten the old basis (V, HV
          // <- V L^-H, HV L^-H, C <- L^H C), measure S and H of it and
          // factor, project the corrections out twice, orthonormalise them
          // among themselves, border S and extend the factor.
          hipLaunchKernelGGL(k_kd_repair_prep, dim3(count), dim3(kKT), factor_lds(cap), main, desc, stat, nb, cap);
          hipLaunchKernelGGL(k_kd_rotate, dim3(row_blocks(), (cap + 7) / 8, count), dim3(kKT), 0, main, desc, stat, 3,
                             nb, cap);
          hipLaunchKernelGGL(k_kd_take, dim3(64, count), dim3(256), 0, main, desc, stat, 1, nb);
          border(0, 3);
          border(1, 3);
          factor(3);
          for (int pass = 0; pass < 2; ++pass) {
            border(2, 0);
            hipLaunchKernelGGL(k_kd_rotate, dim3(row_blocks(), (nb + 7) / 8, count), dim3(kKT), 0, main, desc, stat,
                               4, nb, cap);
          }
          hipLaunchKernelGGL(k_kd_repair_gs, dim3(count), dim3(kKT), 0, main, desc, stat, cap);
          border(0, 4);
          factor(4);
        }
        apply(false);
        border(1, 5);
        eigen(0);
      }
      pull_stat();
      t_dense += std::chrono::duration<double>(std::chrono::steady_clock::now() - td).count();
      for (int k = 0; k < count; ++k) {
        KDesc& d = hdesc[k];
        KPoint& p = kp[k];
        if (!d.live || p.ejected) continue;
        const KStat& st = hstat[k];
        std::copy(p.values.begin(), p.values.end(), p.old.begin());
        if (d.stage2) {
          if (st.nonfinite_h) {
            *msg = "complex H application failed (non-finite output)";
            statuses[k] = BD_ERR_APPLICATION;
            return k;
          }
          if (st.repair_bad) {
            *msg = "complex overlap remains dependent after repair";
            statuses[k] = BD_ERR_APPLICATION;
            return k;
          }
          if (st.eig_bad) {
            *msg = "complex reduced eigensolve failed";
            statuses[k] = BD_ERR_APPLICATION;
            return k;
          }
          int added = d.kk;
          if (st.repair) {
            // H was applied to every correction; the reference applies it to
            // the kept ones only, and so counts them.
            added = st.kept2;
            p.rec.repairs += 1;
            p.rec.h_columns -= d.kk - added;
          }
          p.m += added;
          std::copy(st.values, st.values + nb, p.values.begin());
        }
        p.active.clear();
        for (int i = 0; i < nb; ++i) {
          p.converged[i] = std::fabs(p.values[i] - p.old[i]) < band_threshold(o, i);
          if (!p.converged[i]) p.active.push_back(i);
        }
        p.iterations = step + 1;
        if (p.active.empty() || p.iterations == o.max_iterations) p.done = true;
        else if (p.m + int(p.active.size()) > cap) p.contract = true;
      }
    }
    // Ritz blocks of the k points solved here.
    for (int k = 0; k < count; ++k) {
      KDesc& d = hdesc[k];
      d.live = !kp[k].ejected;
      d.stage2 = 0;
      d.contract = 0;
      d.m = kp[k].m;
      d.out_x = xout[k];
      d.out_hx = hxout ? hxout[k] : nullptr;
      d.ldo = ldx[k];
    }
    push_desc();
    hipLaunchKernelGGL(k_kd_rotate, dim3(row_blocks(), (nb + 7) / 8, count), dim3(kKT), 0, main, desc, stat, 2, nb,
                       cap);
    pull_stat();
    const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int failed = -1;
    for (int k = 0; k < count; ++k) {
      KPoint& p = kp[k];
      const KStat& st = hstat[k];
      bd_davidson_record& r = p.rec;
      if (p.ejected) {
        // The single-k complex Davidson from this k point's Ritz block, with
        // the iterations left.
        KDesc& d = hdesc[k];
        d.live = 1;
        d.m = p.m;
        d.out_x = xout[k];
        d.out_hx = nullptr;
        for (int q = 0; q < count; ++q)
          if (q != k) hdesc[q].live = 0;
        push_desc();
        hipLaunchKernelGGL(k_kd_rotate, dim3(row_blocks(), (nb + 7) / 8, count), dim3(kKT), 0, main, desc, stat, 2,
                           nb, cap);
        sync();
        bd_davidson_options o2 = o;
        o2.complex_columns = 1;
        o2.max_iterations = std::max(1, o.max_iterations - p.iterations);
        o2.profile = 0;
        o2.return_hx = hxout && hxout[k] ? 1 : 0;
        bd_davidson_record r2{};
        const bd_status s2 = bd_davidson_solve_dev(ops[k], xout[k], ldx[k], nullptr, diag[k], &o2, eig + int64_t(k) * nb,
                                                   xout[k], o2.return_hx ? hxout[k] : nullptr, ldx[k],
                                                   conv ? conv + int64_t(k) * nb : nullptr, &r2);
        statuses[k] = s2;
        r2.iterations += p.iterations;
        r2.h_calls += r.h_calls;
        r2.h_columns += r.h_columns;
        r2.restarts += r.restarts;
        r2.repairs += 1;
        r2.ideal_resets += st.ideal_resets;
        r2.periodic_refreshes += st.periodic;
        r2.guard_fallbacks += st.guard_fallbacks;
        if (records) records[k] = r2;
        if (s2 != BD_OK && s2 != BD_NOT_CONVERGED && failed < 0) {
          failed = k;
          *msg = "single-k continuation failed";
        }
        continue;
      }
      bool all = true;
      for (int i = 0; i < nb; ++i) {
        eig[int64_t(k) * nb + i] = p.values[i];
        if (conv) conv[int64_t(k) * nb + i] = p.converged[i];
        all = all && p.converged[i];
      }
      r.iterations = p.iterations;
      r.all_converged = all ? 1 : 0;
      r.ideal_resets = st.ideal_resets;
      r.periodic_refreshes = st.periodic;
      r.guard_fallbacks = st.guard_fallbacks;
      r.seconds_total = total;
      r.seconds_dense = t_dense;
      if (records) records[k] = r;
      statuses[k] = all ? BD_OK : BD_NOT_CONVERGED;
    }
    return failed;
  }
};

}  // namespace bd

struct bd_k_davidson {
  std::unique_ptr<bd::KDavidson> kd;
};

extern "C" bd_status bd_k_davidson_create(bd_gamma_op* const* ops, int32_t count, int32_t bands, int32_t capacity,
                                          bd_k_davidson** out) {
  BD_GUARD({
    if (!ops || !out || count <= 0) bd::fail(BD_ERR_INVALID_PARAMETER, "empty k-point set");
    const int cap = capacity == 0 ? 2 * bands : capacity;
    if (bands <= 0 || bands > bd::kKB || cap < 2 * bands || cap > bd::kKCap)
      bd::fail(BD_ERR_INVALID_PARAMETER, "lock-step k Davidson: bands 1..24, capacity 2 bands..48");
    std::vector<bd_gamma_op*> v(ops, ops + count);
    for (int i = 0; i < count; ++i) {
      if (!ops[i]) bd::fail(BD_ERR_INVALID_PARAMETER, "null op");
      for (int j = 0; j < i; ++j)
        if (ops[i] == ops[j]) bd::fail(BD_ERR_INVALID_PARAMETER, "an op appears twice");
    }
    auto h = std::make_unique<bd_k_davidson>();
    h->kd = std::make_unique<bd::KDavidson>(v, bands, cap);
    *out = h.release();
  });
}

extern "C" void bd_k_davidson_destroy(bd_k_davidson* h) { delete h; }

extern "C" bd_status bd_k_davidson_solve_dev(bd_k_davidson* h, const double* const* x0, const int64_t* ldx0,
                                             const double* const* hx0, const double* const* diag,
                                             const bd_davidson_options* options, double* eigenvalues,
                                             double* const* x, double* const* hx, const int64_t* ldx,
                                             int32_t* converged, bd_davidson_record* records, bd_status* statuses) {
  bd::clear_last_error();
  try {
    if (!h || !x0 || !ldx0 || !diag || !options || !eigenvalues || !x || !ldx || !statuses)
      bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    bd::KDavidson& kd = *h->kd;
    const bd_davidson_options& o = *options;
    if (o.bands != kd.nb || o.occupied > o.bands || o.max_iterations <= 0 || !(o.threshold > 0.0) ||
        !std::isfinite(o.threshold) || (o.return_hx && !hx))
      bd::fail(BD_ERR_INVALID_PARAMETER, "lock-step k Davidson options");
    if (o.capacity != 0 && o.capacity != kd.cap)
      bd::fail(BD_ERR_INVALID_PARAMETER, "capacity differs from the one the solver was created with");
    if (o.preconditioner != BD_PRECOND_QE && o.preconditioner != BD_PRECOND_TPA)
      bd::fail(BD_ERR_INVALID_PARAMETER, "complex Davidson preconditioner (QE or TPA)");
    for (int k = 0; k < kd.count; ++k) {
      const bd::GammaOp& op = kd.ops[k]->op;
      if (!x0[k] || !diag[k] || !x[k]) bd::fail(BD_ERR_INVALID_PARAMETER, "null block");
      if (ldx0[k] < op.n || ldx[k] < op.n || (ldx0[k] & 1) || (ldx[k] & 1) ||
          ((reinterpret_cast<uintptr_t>(x[k]) | reinterpret_cast<uintptr_t>(x0[k])) & 15) ||
          (o.return_hx && (!hx[k] || (reinterpret_cast<uintptr_t>(hx[k]) & 15))))
        bd::fail(BD_ERR_INVALID_PARAMETER, "k-point blocks must be 16-byte aligned with even ld >= 2 nk");
      if (!op.has_potential) bd::fail(BD_ERR_INVALID_PARAMETER, "no potential installed");
    }
    std::string msg;
    const int bad = kd.solve(x0, ldx0, hx0, diag, o, eigenvalues, x, o.return_hx ? hx : nullptr, ldx, converged,
                             records, statuses, &msg);
    if (bad >= 0) {
      bd::set_last_error("k " + std::to_string(bad) + ": " + msg);
      return statuses[bad];
    }
    bd_status result = BD_OK;
    for (int k = 0; k < kd.count; ++k)
      if (statuses[k] == BD_NOT_CONVERGED) result = BD_NOT_CONVERGED;
    return result;
  } catch (const bd::Error& e) {
    bd::set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    bd::set_last_error(e.what());
    return BD_ERR_APPLICATION;
  }
}
