// Small dense FP64 routines of the Davidson subspace step (host side).
#include "dense.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>

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
  // Row-major working copy a[i*n+j]; symmetric input.
  std::vector<double> a(size_t(n) * n), d(n), e(n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[size_t(i) * n + j] = A[i + int64_t(j) * lda];
  auto at = [&](int i, int j) -> double& { return a[size_t(i) * n + j]; };
  // Householder reduction to tridiagonal form, accumulating the transform.
  for (int i = n - 1; i > 0; --i) {
    const int l = i - 1;
    double h = 0.0;
    if (l > 0) {
      double scale = 0.0;
      for (int k = 0; k <= l; ++k) scale += std::fabs(at(i, k));
      if (scale == 0.0) {
        e[i] = at(i, l);
      } else {
        for (int k = 0; k <= l; ++k) {
          at(i, k) /= scale;
          h += at(i, k) * at(i, k);
        }
        double f = at(i, l);
        double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        at(i, l) = f - g;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          at(j, i) = at(i, j) / h;
          g = 0.0;
          for (int k = 0; k <= j; ++k) g += at(j, k) * at(i, k);
          for (int k = j + 1; k <= l; ++k) g += at(k, j) * at(i, k);
          e[j] = g / h;
          f += e[j] * at(i, j);
        }
        const double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          f = at(i, j);
          e[j] = g = e[j] - hh * f;
          for (int k = 0; k <= j; ++k) at(j, k) -= (f * e[k] + g * at(i, k));
        }
      }
    } else {
      e[i] = at(i, l);
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < n; ++i) {
    const int l = i - 1;
    if (d[i] != 0.0) {
      for (int j = 0; j <= l; ++j) {
        double g = 0.0;
        for (int k = 0; k <= l; ++k) g += at(i, k) * at(k, j);
        for (int k = 0; k <= l; ++k) at(k, j) -= g * at(k, i);
      }
    }
    d[i] = at(i, i);
    at(i, i) = 1.0;
    for (int j = 0; j <= l; ++j) at(j, i) = at(i, j) = 0.0;
  }
  // z[i][k] = component k of vector i (rows contiguous for the rotations).
  std::vector<double> z(size_t(n) * n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) z[size_t(i) * n + k] = at(k, i);
  // Implicit-shift QL on the tridiagonal (d, e).
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
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
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i];
          const double b = c * e[i];
          e[i + 1] = (r = std::hypot(f, g));
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (p = s * r);
          g = c * r - b;
          double* zi = &z[size_t(i) * n];
          double* zj = &z[size_t(i + 1) * n];
          for (int k = 0; k < n; ++k) {
            f = zj[k];
            zj[k] = s * zi[k] + c * f;
            zi[k] = c * zi[k] - s * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
  for (int c = 0; c < n; ++c) {
    values[c] = d[order[c]];
    const double* src = &z[size_t(order[c]) * n];
    std::copy(src, src + n, vectors + size_t(c) * n);
  }
  return true;
}

}  // namespace bd
