// CPU FP64 restatement of the Gamma H·X local part, for the CPU baseline.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_build/libbd_cpu_hx.so and
// loaded by oracle/cpu_hx.py, which bench.py's cpu_baseline leg and tests/
// use. The product (libbaandi_amd.so) never loads it.
//
// Follows the reference's CPU path of one H·X block:
//  * two ORC columns packed into one Hermitian-mixed complex vector
//    (packed_value/pack_pair, crates/dft-physics/src/distributed_real_gamma.rs:638-682);
//  * the pruned 3D transform of crates/dft-solvers/src/distributed_fft.rs:837-1288:
//    inverse 1D FFTs along axis 0 on the occupied sticks only, along axis 1 on
//    the occupied axis-2 planes only, then full lines along axis 2; V(r); the
//    forward transforms in reverse order on the same pruned sets; gather, 1/N;
//  * the even/odd unpack (mirror_back_and_unpack, :1270-1352) and + T_G x.
// The nonlocal products P D P^T X are dense GEMMs; oracle/cpu_hx.py does them
// with NumPy's BLAS (the reference uses faer or MKL dgemm for them).
// Pairs are independent: OpenMP over pairs, one grid-sized scratch per thread.
//
// The 1D transforms are an iterative Stockham autosort FFT over the factors
// 4, 2, 3, 5 (any other prime factor by a direct DFT pass), eight lines at a
// time, with twiddles
// from long-double sines and cosines: e^{+2 pi i/n} for the inverse, e^{-}
// for the forward direction, both unscaled (crates/dft-numerics/src/lib.rs:46-56).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <vector>

namespace {
using cd = std::complex<double>;

struct Plan1D {
  int n = 0;
  std::vector<int> radix;          // factors in pass order
  std::vector<cd> tw_fwd, tw_inv;  // per pass: w_{nc}^(p k), p < nc/R, k < R
  std::vector<cd> rt_fwd, rt_inv;  // per pass: the R-th roots of unity
  std::vector<size_t> tw_off, rt_off;
};

cd unit_root(long double num, long double den, bool inverse) {
  const long double ang = (inverse ? 2.0L : -2.0L) * 3.14159265358979323846264338327950288L * num / den;
  return cd(double(cosl(ang)), double(sinl(ang)));
}

Plan1D make_plan(int n) {
  Plan1D p;
  p.n = n;
  int r = n;
  for (int f : {4, 2, 3, 5})
    while (r % f == 0) {
      p.radix.push_back(f);
      r /= f;
    }
  for (int f = 7; r > 1; f += 2)
    while (r % f == 0) {
      p.radix.push_back(f);
      r /= f;
    }
  int nc = n;  // length of the sub-transforms a pass works on
  for (int R : p.radix) {
    const int m = nc / R;
    p.tw_off.push_back(p.tw_fwd.size());
    for (int q = 0; q < m; ++q)
      for (int k = 0; k < R; ++k) {
        p.tw_fwd.push_back(unit_root((long double)q * k, nc, false));
        p.tw_inv.push_back(unit_root((long double)q * k, nc, true));
      }
    p.rt_off.push_back(p.rt_fwd.size());
    for (int e = 0; e < R; ++e) {
      p.rt_fwd.push_back(unit_root(e, R, false));
      p.rt_inv.push_back(unit_root(e, R, true));
    }
    nc = m;
  }
  return p;
}

// Transforms of B lines at once, held line-interleaved: x[i * B + b] is
// element i of line b. Self-sorting Stockham passes, decimation in frequency:
// with stride s (the product of the earlier radices) and m = nc / R,
//   y[j + s (R q + k)] = w_nc^(q k) sum_r x[j + s (q + m r)] e^{-+2 pi i r k / R}
// for every line; the innermost loop runs over the B lines (contiguous).
// Result in x, y is scratch of the same size.
constexpr int kB = 8;  // lines per batch

template <int R>
void pass(const cd* __restrict__ src, cd* __restrict__ dst, int s, int m, const cd* tw, const cd* rt) {
  for (int q = 0; q < m; ++q)
    for (int j = 0; j < s; ++j) {
      const cd* in = src + size_t(j + size_t(s) * q) * kB;
      const size_t istride = size_t(s) * m * kB;
      cd* out = dst + size_t(j + size_t(s) * size_t(R) * q) * kB;
      const size_t ostride = size_t(s) * kB;
      cd w[R];
      for (int k = 0; k < R; ++k) w[k] = tw[q * R + k];
      for (int b = 0; b < kB; ++b) {
        cd a[R];
        for (int r = 0; r < R; ++r) a[r] = in[r * istride + b];
        cd o[R];
        if constexpr (R == 2) {
          o[0] = a[0] + a[1];
          o[1] = a[0] - a[1];
        } else if constexpr (R == 4) {
          const cd t0 = a[0] + a[2], t1 = a[0] - a[2], t2 = a[1] + a[3], t3 = (a[1] - a[3]) * rt[1];
          o[0] = t0 + t2;
          o[1] = t1 + t3;
          o[2] = t0 - t2;
          o[3] = t1 - t3;
        } else {
          for (int k = 0; k < R; ++k) {
            cd acc = a[0];
            for (int r = 1; r < R; ++r) acc += a[r] * rt[(r * k) % R];
            o[k] = acc;
          }
        }
        for (int k = 0; k < R; ++k) out[k * ostride + b] = o[k] * w[k];
      }
    }
}

void pass_generic(int R, const cd* src, cd* dst, int s, int m, const cd* tw, const cd* rt) {
  for (int q = 0; q < m; ++q)
    for (int j = 0; j < s; ++j)
      for (int b = 0; b < kB; ++b)
        for (int k = 0; k < R; ++k) {
          cd acc = 0.0;
          for (int r = 0; r < R; ++r) acc += src[size_t(j + size_t(s) * (q + size_t(m) * r)) * kB + b] * rt[(r * k) % R];
          dst[size_t(j + size_t(s) * (size_t(R) * q + k)) * kB + b] = acc * tw[q * R + k];
        }
}

void fft_lines(const Plan1D& p, cd* x, cd* y, bool inverse) {
  const int n = p.n;
  if (n == 1) return;
  cd* src = x;
  cd* dst = y;
  int s = 1, nc = n;
  for (size_t ps = 0; ps < p.radix.size(); ++ps) {
    const int R = p.radix[ps];
    const int m = nc / R;
    const cd* tw = (inverse ? p.tw_inv.data() : p.tw_fwd.data()) + p.tw_off[ps];
    const cd* rt = (inverse ? p.rt_inv.data() : p.rt_fwd.data()) + p.rt_off[ps];
    switch (R) {
      case 2: pass<2>(src, dst, s, m, tw, rt); break;
      case 3: pass<3>(src, dst, s, m, tw, rt); break;
      case 4: pass<4>(src, dst, s, m, tw, rt); break;
      case 5: pass<5>(src, dst, s, m, tw, rt); break;
      default: pass_generic(R, src, dst, s, m, tw, rt);
    }
    std::swap(src, dst);
    s *= R;
    nc = m;
  }
  if (src != x) std::memcpy(x, src, sizeof(cd) * size_t(n) * kB);
}

struct Op {
  int n0, n1, n2;
  int64_t n, nrep, N;
  std::vector<int64_t> plus, minus;  // flat index of +G / -G per representative
  std::vector<double> kinetic, potential;
  Plan1D p0, p1, p2;
  std::vector<std::pair<int, int>> sticks;  // occupied (i1, i2) along axis 0
  std::vector<int> planes;                  // occupied i2 values
  int threads = 1;
};

inline void decompose(const Op& o, int64_t f, int& i0, int& i1, int& i2) {
  i2 = int(f % o.n2);
  const int64_t t = f / o.n2;
  i1 = int(t % o.n1);
  i0 = int(t / o.n1);
}

// Inverse (to real space) or forward transform of the full grid g, pruned to
// the occupied sticks and planes (the grid holds zeros outside them on the
// inverse direction; on the forward direction only those entries are read).
// Lines go through fft_lines kB at a time (a short last batch is padded with
// copies of its first line, whose results are not stored).
void transform(const Op& o, cd* g, cd* buf, cd* scratch, bool inverse) {
  const int n0 = o.n0, n1 = o.n1, n2 = o.n2;
  auto at = [&](int i0, int i1, int i2) -> cd& { return g[(int64_t(i0) * n1 + i1) * n2 + i2]; };
  auto axis0 = [&]() {
    const int S = int(o.sticks.size());
    for (int b0 = 0; b0 < S; b0 += kB) {
      const int nb = std::min(kB, S - b0);
      for (int i0 = 0; i0 < n0; ++i0)
        for (int b = 0; b < kB; ++b) {
          const auto& st = o.sticks[b0 + (b < nb ? b : 0)];
          buf[size_t(i0) * kB + b] = at(i0, st.first, st.second);
        }
      fft_lines(o.p0, buf, scratch, inverse);
      for (int i0 = 0; i0 < n0; ++i0)
        for (int b = 0; b < nb; ++b) {
          const auto& st = o.sticks[b0 + b];
          at(i0, st.first, st.second) = buf[size_t(i0) * kB + b];
        }
    }
  };
  auto axis1 = [&]() {
    const int Z = int(o.planes.size());
    for (int i0 = 0; i0 < n0; ++i0)
      for (int b0 = 0; b0 < Z; b0 += kB) {
        const int nb = std::min(kB, Z - b0);
        for (int i1 = 0; i1 < n1; ++i1)
          for (int b = 0; b < kB; ++b) buf[size_t(i1) * kB + b] = at(i0, i1, o.planes[b0 + (b < nb ? b : 0)]);
        fft_lines(o.p1, buf, scratch, inverse);
        for (int i1 = 0; i1 < n1; ++i1)
          for (int b = 0; b < nb; ++b) at(i0, i1, o.planes[b0 + b]) = buf[size_t(i1) * kB + b];
      }
  };
  auto axis2 = [&]() {
    for (int i0 = 0; i0 < n0; ++i0)
      for (int b0 = 0; b0 < n1; b0 += kB) {
        const int nb = std::min(kB, n1 - b0);
        for (int i2 = 0; i2 < n2; ++i2)
          for (int b = 0; b < kB; ++b) buf[size_t(i2) * kB + b] = at(i0, b0 + (b < nb ? b : 0), i2);
        fft_lines(o.p2, buf, scratch, inverse);
        for (int i2 = 0; i2 < n2; ++i2)
          for (int b = 0; b < nb; ++b) at(i0, b0 + b, i2) = buf[size_t(i2) * kB + b];
      }
  };
  if (inverse) {
    axis0();
    axis1();
    axis2();
  } else {
    axis2();
    axis1();
    axis0();
  }
}
}  // namespace

extern "C" {

void* bd_cpu_hx_create(int n0, int n1, int n2, int64_t n, int64_t nrep, const int64_t* plus, const int64_t* minus,
                       const double* kinetic) {
  auto* o = new Op;
  o->n0 = n0;
  o->n1 = n1;
  o->n2 = n2;
  o->n = n;
  o->nrep = nrep;
  o->N = int64_t(n0) * n1 * n2;
  o->plus.assign(plus, plus + nrep);
  o->minus.assign(minus, minus + nrep);
  o->kinetic.assign(kinetic, kinetic + n);
  o->potential.assign(size_t(o->N), 0.0);
  o->p0 = make_plan(n0);
  o->p1 = make_plan(n1);
  o->p2 = make_plan(n2);
  std::set<std::pair<int, int>> st;
  std::set<int> pl;
  for (int64_t i = 0; i < nrep; ++i)
    for (int64_t f : {plus[i], minus[i]}) {
      int i0, i1, i2;
      decompose(*o, f, i0, i1, i2);
      st.insert({i1, i2});
      pl.insert(i2);
    }
  o->sticks.assign(st.begin(), st.end());
  o->planes.assign(pl.begin(), pl.end());
  o->threads = omp_get_max_threads();
  return o;
}

void bd_cpu_hx_destroy(void* h) { delete static_cast<Op*>(h); }

void bd_cpu_hx_set_potential(void* h, const double* v) {
  auto* o = static_cast<Op*>(h);
  std::memcpy(o->potential.data(), v, sizeof(double) * o->N);
}

void bd_cpu_hx_set_threads(void* h, int threads) { static_cast<Op*>(h)->threads = threads > 0 ? threads : 1; }

// Sticks and planes of the pruned plan (for tests): S, Z.
void bd_cpu_hx_plan(void* h, int64_t* out2) {
  auto* o = static_cast<Op*>(h);
  out2[0] = int64_t(o->sticks.size());
  out2[1] = int64_t(o->planes.size());
}

// Y = (local + kinetic) X for m ORC columns (column-major, leading dims);
// the nonlocal term is added by the caller.
void bd_cpu_hx_apply_local(void* h, const double* x, int64_t ldx, int m, double* y, int64_t ldy) {
  auto* o = static_cast<Op*>(h);
  const double s2 = std::sqrt(2.0), is2 = 1.0 / std::sqrt(2.0);
  const int pairs = (m + 1) / 2;
  const double invN = 1.0 / double(o->N);
  const int nmax = std::max(o->n0, std::max(o->n1, o->n2));
#pragma omp parallel num_threads(o->threads)
  {
    std::vector<cd> grid(size_t(o->N)), line(size_t(nmax) * kB), scratch(size_t(nmax) * kB);
#pragma omp for schedule(dynamic, 1)
    for (int t = 0; t < pairs; ++t) {
      const int ca = 2 * t, cb = 2 * t + 1;
      const double* xa = x + int64_t(ca) * ldx;
      const double* xb = cb < m ? x + int64_t(cb) * ldx : nullptr;
      std::fill(grid.begin(), grid.end(), cd(0.0, 0.0));
      cd* g = grid.data();
      g[o->plus[0]] = cd(xa[0], xb ? xb[0] : 0.0);
      for (int64_t i = 1; i < o->nrep; ++i) {
        const cd a(xa[2 * i - 1] * is2, xa[2 * i] * is2);
        const cd b = xb ? cd(xb[2 * i - 1] * is2, xb[2 * i] * is2) : cd(0.0, 0.0);
        const cd ib(-b.imag(), b.real());
        g[o->plus[i]] = a + ib;                              // a + i b
        g[o->minus[i]] = std::conj(a) + cd(b.imag(), b.real());  // conj(a) + i conj(b)
      }
      transform(*o, g, line.data(), scratch.data(), true);
      for (int64_t f = 0; f < o->N; ++f) g[f] *= o->potential[f];
      transform(*o, g, line.data(), scratch.data(), false);
      double* ya = y + int64_t(ca) * ldy;
      double* yb = cb < m ? y + int64_t(cb) * ldy : nullptr;
      const cd h0 = g[o->plus[0]] * invN;
      ya[0] = h0.real() + o->kinetic[0] * xa[0];
      if (yb) yb[0] = h0.imag() + o->kinetic[0] * xb[0];
      for (int64_t i = 1; i < o->nrep; ++i) {
        const cd hp = g[o->plus[i]] * invN, hm = g[o->minus[i]] * invN;
        const cd even = (hp + std::conj(hm)) * 0.5, odd = (hp - std::conj(hm)) * 0.5;
        const int64_t r = 2 * i - 1;
        ya[r] = even.real() * s2 + o->kinetic[r] * xa[r];
        ya[r + 1] = even.imag() * s2 + o->kinetic[r + 1] * xa[r + 1];
        if (yb) {
          yb[r] = odd.imag() * s2 + o->kinetic[r] * xb[r];
          yb[r + 1] = -odd.real() * s2 + o->kinetic[r + 1] * xb[r + 1];
        }
      }
    }
  }
}
}
