// Small dense FP64 routines of the Davidson subspace step (host side).
#include "dense.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <numeric>
#include <vector>

#include <omp.h>

namespace bd {

namespace {
constexpr double kEps = std::numeric_limits<double>::epsilon();
constexpr double kPivotRatioMin = kEps / 1e-12;  // lean_davidson.rs:62-64
}  // namespace

bool extend_inverse_cholesky(const double* S, int64_t ld, int from, int to, double* T, double* min_ratio) {
  const int k = to - from;
  *min_ratio = std::numeric_limits<double>::infinity();
  if (k <= 0) return true;
  auto s = [&](int i, int j) { return S[i + int64_t(j) * ld]; };
  auto t = [&](int i, int j) -> double& { return T[i + int64_t(j) * ld]; };
  // W = T11 S12 (from x k), G = S22 - W^T W.
  std::vector<double> W(size_t(from) * k), G(size_t(k) * k), L(size_t(k) * k, 0.0);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < from; ++i) {
      double v = 0.0;
      for (int l = 0; l <= i; ++l) v += t(i, l) * s(l, from + j);
      W[i + size_t(j) * from] = v;
    }
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < k; ++i) {
      double v = 0.0;
      for (int l = 0; l < from; ++l) v += W[l + size_t(i) * from] * W[l + size_t(j) * from];
      G[i + size_t(j) * k] = s(from + i, from + j) - v;
    }
  // Cholesky with the pivot-ratio test.
  for (int j = 0; j < k; ++j) {
    double d = G[j + size_t(j) * k];
    for (int l = 0; l < j; ++l) d -= L[j + size_t(l) * k] * L[j + size_t(l) * k];
    const double ratio = d / s(from + j, from + j);
    if (std::isnan(ratio)) {
      *min_ratio = 0.0;
      return false;
    }
    *min_ratio = std::min(*min_ratio, ratio);
    if (ratio < kPivotRatioMin) return false;
    const double ljj = std::sqrt(d);
    L[j + size_t(j) * k] = ljj;
    for (int i = j + 1; i < k; ++i) {
      double x = G[i + size_t(j) * k];
      for (int l = 0; l < j; ++l) x -= L[i + size_t(l) * k] * L[j + size_t(l) * k];
      L[i + size_t(j) * k] = x / ljj;
    }
  }
  // T22 = L^{-1} by forward substitution.
  std::vector<double> T22(size_t(k) * k, 0.0);
  for (int c = 0; c < k; ++c)
    for (int i = c; i < k; ++i) {
      double x = (i == c) ? 1.0 : 0.0;
      for (int l = c; l < i; ++l) x -= L[i + size_t(l) * k] * T22[l + size_t(c) * k];
      T22[i + size_t(c) * k] = x / L[i + size_t(i) * k];
    }
  // U = W^T T11 (k x from); T21 = -T22 U.
  std::vector<double> U(size_t(k) * from);
  for (int c = 0; c < from; ++c)
    for (int i = 0; i < k; ++i) {
      double v = 0.0;
      for (int l = c; l < from; ++l) v += W[l + size_t(i) * from] * t(l, c);
      U[i + size_t(c) * k] = v;
    }
  // Write the new rows/columns of T (T is lower triangular).
  for (int c = 0; c < k; ++c) {
    for (int i = 0; i < from + c; ++i) t(i, from + c) = 0.0;
    for (int i = 0; i < k; ++i) t(from + i, from + c) = T22[i + size_t(c) * k];
  }
  for (int c = 0; c < from; ++c)
    for (int i = 0; i < k; ++i) {
      double v = 0.0;
      for (int r = 0; r <= i; ++r) v += T22[i + size_t(r) * k] * U[r + size_t(c) * k];
      t(from + i, c) = -v;
    }
  // Finite values and the row-growth guard (lean_davidson.rs:290-326).
  for (int i = from; i < to; ++i) {
    double growth = 0.0;
    for (int j = 0; j <= i; ++j) {
      const double v = t(i, j);
      if (!std::isfinite(v)) return false;
      growth += std::fabs(v) * std::sqrt(s(j, j));
    }
    if (!std::isfinite(growth) || kEps * growth * growth > std::sqrt(kEps)) return false;
  }
  return true;
}

bool symmetric_eigen(const double* A, int64_t lda, int n, double* values, double* vectors) {
  if (n <= 0) return true;
  // Column-major working copy of the symmetric input.
  std::vector<double> a(size_t(n) * n), d(n, 0.0), e(n, 0.0), tau(n, 0.0), p(n), w(n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[i + size_t(j) * n] = A[i + int64_t(j) * lda];
  auto at = [&](int i, int j) -> double& { return a[i + size_t(j) * n]; };
  // Householder tridiagonalisation: step k reflects A[k+1:, k] onto e_1;
  // the reflector v is stored in place of that column, tau[k] beside it.
  for (int k = 0; k + 2 < n; ++k) {
    const int L = n - k - 1;
    double* x = &at(k + 1, k);
    double sigma = 0.0;
    for (int i = 0; i < L; ++i) sigma += x[i] * x[i];
    sigma = std::sqrt(sigma);
    d[k] = at(k, k);
    if (sigma == 0.0) {
      e[k] = 0.0;
      tau[k] = 0.0;
      continue;
    }
    const double alpha = x[0] >= 0.0 ? -sigma : sigma;
    x[0] -= alpha;
    double vv = 0.0;
    for (int i = 0; i < L; ++i) vv += x[i] * x[i];
    const double t = 2.0 / vv;
    tau[k] = t;
    e[k] = alpha;
    // p = t B v over the trailing block B = A[k+1:, k+1:].
    const bool par = int64_t(L) * L >= 16384;
#pragma omp parallel for schedule(static) if (par)
    for (int i = 0; i < L; ++i) {
      // B is symmetric: row i of B is column i (contiguous).
      const double* col = &at(k + 1, k + 1 + i);
      double s = 0.0;
      for (int j = 0; j < L; ++j) s += col[j] * x[j];
      p[i] = t * s;
    }
    double pv = 0.0;
    for (int i = 0; i < L; ++i) pv += p[i] * x[i];
    const double K = 0.5 * t * pv;
    for (int i = 0; i < L; ++i) w[i] = p[i] - K * x[i];
    // B -= v w^T + w v^T (columns in parallel).
#pragma omp parallel for schedule(static) if (par)
    for (int j = 0; j < L; ++j) {
      double* col = &at(k + 1, k + 1 + j);
      const double vj = x[j], wj = w[j];
      for (int i = 0; i < L; ++i) col[i] -= x[i] * wj + w[i] * vj;
    }
  }
  if (n >= 2) {
    d[n - 2] = at(n - 2, n - 2);
    e[n - 2] = at(n - 1, n - 2);
  }
  d[n - 1] = at(n - 1, n - 1);
  // Q = H_0 H_1 ... accumulated backwards (column-major, contiguous columns).
  // Columns are padded to whole 64-byte lines so that threads working on
  // disjoint row blocks never share a cache line.
  const int ldq = (n + 7) / 8 * 8;
  double* qc = static_cast<double*>(std::aligned_alloc(64, size_t(ldq) * n * sizeof(double)));
  if (!qc) return false;
  struct Free {
    double* p;
    ~Free() { std::free(p); }
  } qc_guard{qc};
  std::fill(qc, qc + size_t(ldq) * n, 0.0);
  for (int i = 0; i < n; ++i) qc[size_t(i) * ldq + i] = 1.0;
  for (int k = n - 3; k >= 0; --k) {
    if (tau[k] == 0.0) continue;
    const int L = n - k - 1;
    const double* v = &at(k + 1, k);
    const double t = tau[k];
#pragma omp parallel for schedule(static) if (int64_t(L) * L >= 16384)
    for (int c = k + 1; c < n; ++c) {
      double* col = &qc[size_t(c) * ldq + k + 1];
      double s = 0.0;
      for (int i = 0; i < L; ++i) s += v[i] * col[i];
      s *= t;
      for (int i = 0; i < L; ++i) col[i] -= s * v[i];
    }
  }
  // Implicit-shift QL on (d, e): e[i] couples d[i], d[i+1]. Every rotation
  // (column pair i, i+1; c, s) is recorded in order; rows of Q are then
  // independent, so the whole sequence is applied in one parallel pass.
  struct Rot {
    int i;
    double c, s;
  };
  std::vector<Rot> rots;
  rots.reserve(size_t(n) * n);
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= kEps * dd) break;
      }
      if (m != l) {
        if (iter++ == 90) return false;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, pp = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i];
          const double b = c * e[i];
          e[i + 1] = (r = std::hypot(f, g));
          if (r == 0.0) {
            d[i + 1] -= pp;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - pp;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (pp = s * r);
          g = c * r - b;
          rots.push_back({i, c, s});
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= pp;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  {
    const int threads = std::max(1, std::min(omp_get_max_threads(), 16));
    const int nblk = (rots.size() * size_t(n) >= (1u << 16)) ? threads : 1;
    const int kBlock = ((n + nblk - 1) / nblk + 7) / 8 * 8;
#pragma omp parallel for schedule(static) num_threads(nblk) if (nblk > 1)
    for (int bk = 0; bk < nblk; ++bk) {
      const int r0 = std::min(n, bk * kBlock), r1 = std::min(n, r0 + kBlock);
      for (const Rot& t : rots) {
        double* zi = &qc[size_t(t.i) * ldq];
        double* zj = &qc[size_t(t.i + 1) * ldq];
        for (int row = r0; row < r1; ++row) {
          const double f = zj[row];
          zj[row] = t.s * zi[row] + t.c * f;
          zi[row] = t.c * zi[row] - t.s * f;
        }
      }
    }
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; ++c) {
    values[c] = d[order[c]];
    const int src = order[c];
    std::copy(&qc[size_t(src) * ldq], &qc[size_t(src) * ldq] + n, vectors + size_t(c) * n);
  }
  return true;
}

}  // namespace bd
