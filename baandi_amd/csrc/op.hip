// Gamma H·X on MI355X: band-pair-batched pruned 3D FFT (sticks along x, LDS
// y-z planes with the local-potential multiply), kinetic term and the
// nonlocal P·D·PᵀX products.
//
// Data path per apply (m ORC columns, P = ceil(m/2) complex pairs):
//   k_stick_inverse  X (ORC) -> pack two columns (distributed_real_gamma.rs:651-682)
//                    -> scatter to sticks -> inverse x FFT -> stick buffer [pair][x][stick]
//   k_plane          per (x, pair): sticks -> LDS plane (z-compact) -> inverse y FFT ->
//                    per y-row chunk: inverse z FFT, * V(r), forward z FFT -> forward y FFT
//                    -> sticks (distributed_fft.rs:956-1080 fused in one kernel)
//   k_stick_forward  forward x FFT -> gather +G/-G (both in one tile) -> even/odd unpack
//                    (distributed_real_gamma.rs:1316-1339), 1/N, + T_G x -> Y
//   projections      Q = P^T X (split-K GEMM), W = D Q (per-atom blocks), Y += P W.
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <set>
#include <string>
#include <type_traits>

#include "common.hpp"
#include "gemm.hpp"
#include "op.hpp"

namespace bd {

namespace {
constexpr int kThreads = 256;
constexpr int kDensityGroups = 8;  // band-density pair groups per x plane
constexpr double kSqrt2 = 1.41421356237309504880168872420969808;
constexpr double kInvSqrt2 = 0.70710678118654752440084436210484904;

bool pass_fits(int R, int nlines, int n, int threads) {
  return int64_t(nlines) * (n / R) <= int64_t(pass_mb(R, threads)) * threads;
}
}  // namespace

// The kernels are built for workgroups of 256 and 512 threads.
bool fft_fits_threads(int nlines, const RadixPlan& rp, int threads) {
  if (threads != 256 && threads != 512) return false;
  for (int p = 0; p < rp.count; ++p)
    if (!pass_fits(rp.radix[p], nlines, rp.n, threads)) return false;
  return true;
}

namespace {
bool fft_fits(int nlines, const RadixPlan& rp) { return fft_fits_threads(nlines, rp, kThreads); }

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void stage_roots(c2<T>* dst, const c2<double>* src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = {T(src[i].x), T(src[i].y)};
}

// Pack ORC column pair t into its sticks and inverse-transform along x. The
// stick buffer and the transform use F (double, or float in the mixed tier);
// the ORC columns are always double.
template <typename F>
__global__ __launch_bounds__(kThreads) void k_stick_inverse(
    const double* __restrict__ x, int64_t ldx, int m, int n0, int T, int Spad,
    const int* __restrict__ tile_off, const int4* __restrict__ entries,
    const c2<double>* __restrict__ roots_x, RadixPlan rx, c2<F>* __restrict__ sticks) {
  extern __shared__ char smem[];
  c2<F>* L = reinterpret_cast<c2<F>*>(smem);
  c2<F>* rt = L + T * n0;
  const int tile = blockIdx.x, t = blockIdx.y;
  const int ca = 2 * t, cb = 2 * t + 1;
  const double* xa = x + int64_t(ca) * ldx;
  const double* xb = cb < m ? x + int64_t(cb) * ldx : nullptr;
  for (int i = threadIdx.x; i < T * n0; i += blockDim.x) L[i] = {F(0), F(0)};
  stage_roots(rt, roots_x, n0);
  __syncthreads();
  const int e0 = tile_off[tile], e1 = tile_off[tile + 1];
  for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int4 en = entries[e];
    const int i = en.x;
    if (i == 0) {
      L[en.y] = {F(xa[0]), F(xb ? xb[0] : 0.0)};
    } else {
      const double ar = xa[2 * i - 1] * kInvSqrt2, ai = xa[2 * i] * kInvSqrt2;
      const double br = xb ? xb[2 * i - 1] * kInvSqrt2 : 0.0, bi = xb ? xb[2 * i] * kInvSqrt2 : 0.0;
      L[en.y] = {F(ar - bi), F(ai + br)};   // a + i b
      L[en.z] = {F(ar + bi), F(br - ai)};   // conj(a) + i conj(b)
    }
  }
  __syncthreads();
  lds_fft<true>(L, T, n0, 1, rx, rt);
  c2<F>* out = sticks + int64_t(t) * n0 * Spad + int64_t(tile) * T;
  for (int e = threadIdx.x; e < T * n0; e += blockDim.x) {
    const int xi = e / T, slot = e - xi * T;
    out[int64_t(xi) * Spad + slot] = L[slot * n0 + xi];
  }
}

// Forward x FFT of a tile, gather +G/-G, unpack to ORC (in double), add T_G x.
template <typename F>
__global__ __launch_bounds__(kThreads) void k_stick_forward(
    const double* __restrict__ x, int64_t ldx, double* __restrict__ y, int64_t ldy, int m,
    int n0, int T, int Spad, const int* __restrict__ tile_off, const int4* __restrict__ entries,
    const c2<double>* __restrict__ roots_x, RadixPlan rx, const c2<F>* __restrict__ sticks,
    const double* __restrict__ kinetic, double inv_n, int accumulate) {
  extern __shared__ char smem[];
  c2<F>* L = reinterpret_cast<c2<F>*>(smem);
  c2<F>* rt = L + T * n0;
  const int tile = blockIdx.x, t = blockIdx.y;
  const c2<F>* in = sticks + int64_t(t) * n0 * Spad + int64_t(tile) * T;
  for (int e = threadIdx.x; e < T * n0; e += blockDim.x) {
    const int xi = e / T, slot = e - xi * T;
    L[slot * n0 + xi] = in[int64_t(xi) * Spad + slot];
  }
  stage_roots(rt, roots_x, n0);
  __syncthreads();
  lds_fft<false>(L, T, n0, 1, rx, rt);
  const int ca = 2 * t, cb = 2 * t + 1;
  const bool hb = cb < m;
  const double* xa = x + int64_t(ca) * ldx;
  const double* xb = x + int64_t(hb ? cb : ca) * ldx;
  double* ya = y + int64_t(ca) * ldy;
  double* yb = y + int64_t(hb ? cb : ca) * ldy;
  const int e0 = tile_off[tile], e1 = tile_off[tile + 1];
  for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int4 en = entries[e];
    const int i = en.x;
    const c2<F> lp = L[en.y];
    const c2<double> hp = {double(lp.x) * inv_n, double(lp.y) * inv_n};
    // accumulate: y += the local part alone (the warm start's dV X);
    // otherwise y = local part + T_G x.
    if (i == 0) {
      if (accumulate) {
        ya[0] += hp.x;
        if (hb) yb[0] += hp.y;
      } else {
        const double k0 = kinetic[0];
        ya[0] = hp.x + k0 * xa[0];
        if (hb) yb[0] = hp.y + k0 * xb[0];
      }
    } else {
      const c2<F> lm = L[en.z];
      const c2<double> hm = {double(lm.x) * inv_n, double(lm.y) * inv_n};
      // even = (hp + conj(hm))/2, odd = (hp - conj(hm))/2, rows x sqrt2.
      const double er = 0.5 * (hp.x + hm.x), ei = 0.5 * (hp.y - hm.y);
      const double orr = 0.5 * (hp.x - hm.x), oi = 0.5 * (hp.y + hm.y);
      const int r1 = 2 * i - 1, r2 = 2 * i;
      if (accumulate) {
        ya[r1] += er * kSqrt2;
        ya[r2] += ei * kSqrt2;
        if (hb) {
          yb[r1] += oi * kSqrt2;
          yb[r2] += -orr * kSqrt2;
        }
      } else {
        const double k = kinetic[r1];
        ya[r1] = er * kSqrt2 + k * xa[r1];
        ya[r2] = ei * kSqrt2 + k * xa[r2];
        if (hb) {
          yb[r1] = oi * kSqrt2 + k * xb[r1];
          yb[r2] = -orr * kSqrt2 + k * xb[r2];
        }
      }
    }
  }
}

// k point: scatter complex column t (rows 2g, 2g+1 = Re, Im of coefficient g)
// into its sticks and inverse-transform along x; one complex FFT per column
// (distributed_complex.rs:340-386).
template <typename F>
__global__ __launch_bounds__(kThreads) void k_kstick_inverse(
    const double* __restrict__ x, int64_t ldx, int n0, int T, int Spad, const int* __restrict__ tile_off,
    const int4* __restrict__ entries, const c2<double>* __restrict__ roots_x, RadixPlan rx,
    c2<F>* __restrict__ sticks) {
  extern __shared__ char smem[];
  c2<F>* L = reinterpret_cast<c2<F>*>(smem);
  c2<F>* rt = L + T * n0;
  const int tile = blockIdx.x, t = blockIdx.y;
  const double* xc = x + int64_t(t) * ldx;
  for (int i = threadIdx.x; i < T * n0; i += blockDim.x) L[i] = {F(0), F(0)};
  stage_roots(rt, roots_x, n0);
  __syncthreads();
  const int e0 = tile_off[tile], e1 = tile_off[tile + 1];
  for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int4 en = entries[e];
    const double2 v = *reinterpret_cast<const double2*>(xc + 2 * int64_t(en.x));
    L[en.y] = {F(v.x), F(v.y)};
  }
  __syncthreads();
  lds_fft<true>(L, T, n0, 1, rx, rt);
  c2<F>* out = sticks + int64_t(t) * n0 * Spad + int64_t(tile) * T;
  for (int e = threadIdx.x; e < T * n0; e += blockDim.x) {
    const int xi = e / T, slot = e - xi * T;
    out[int64_t(xi) * Spad + slot] = L[slot * n0 + xi];
  }
}

// k point: forward x FFT of a tile, gather the sphere, 1/N, + |k+G|^2/2 x.
template <typename F>
__global__ __launch_bounds__(kThreads) void k_kstick_forward(
    const double* __restrict__ x, int64_t ldx, double* __restrict__ y, int64_t ldy, int n0, int T, int Spad,
    const int* __restrict__ tile_off, const int4* __restrict__ entries, const c2<double>* __restrict__ roots_x,
    RadixPlan rx, const c2<F>* __restrict__ sticks, const double* __restrict__ kinetic, double inv_n) {
  extern __shared__ char smem[];
  c2<F>* L = reinterpret_cast<c2<F>*>(smem);
  c2<F>* rt = L + T * n0;
  const int tile = blockIdx.x, t = blockIdx.y;
  const c2<F>* in = sticks + int64_t(t) * n0 * Spad + int64_t(tile) * T;
  for (int e = threadIdx.x; e < T * n0; e += blockDim.x) {
    const int xi = e / T, slot = e - xi * T;
    L[slot * n0 + xi] = in[int64_t(xi) * Spad + slot];
  }
  stage_roots(rt, roots_x, n0);
  __syncthreads();
  lds_fft<false>(L, T, n0, 1, rx, rt);
  const double* xc = x + int64_t(t) * ldx;
  double* yc = y + int64_t(t) * ldy;
  const int e0 = tile_off[tile], e1 = tile_off[tile + 1];
  for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int4 en = entries[e];
    const int64_t r = 2 * int64_t(en.x);
    const c2<F> h = L[en.y];
    const double2 v = *reinterpret_cast<const double2*>(xc + r);
    const double k = kinetic[r];
    *reinterpret_cast<double2*>(yc + r) = make_double2(double(h.x) * inv_n + k * v.x, double(h.y) * inv_n + k * v.y);
  }
}

// Plane pass: inverse y, per-row-chunk inverse z * V * forward z, forward y.
// mode 0: H apply (writes sticks back); mode 1: band density (accumulates
// w_a psi_a^2 + w_b psi_b^2 over all pairs into rho of plane x).
template <int MODE, typename F, int NT>
__global__ __launch_bounds__(NT) void k_plane(
    c2<F>* __restrict__ sticks, int n0, int n1, int n2, int Spad, int Z, int chunk,
    const int2* __restrict__ stick_yz, const int* __restrict__ z_of_zc,
    const c2<double>* __restrict__ roots, RadixPlan ry, RadixPlan rz,
    const F* __restrict__ potential, int pairs, const double* __restrict__ weights,
    double inv_volume, double* __restrict__ rho, c2<F>* __restrict__ keep) {
  extern __shared__ char smem[];
  c2<F>* P = reinterpret_cast<c2<F>*>(smem);  // [Z][n1]
  c2<F>* Rb = P + Z * n1;                     // [chunk][n2]
  c2<F>* rty = Rb + chunk * n2;
  c2<F>* rtz = rty + n1;
  const int xi = MODE == 0 ? blockIdx.y : blockIdx.x;
  stage_roots(rty, roots + n0, n1);
  stage_roots(rtz, roots + n0 + n1, n2);
  // MODE 1: group blockIdx.y accumulates pairs y, y+G, y+2G, ... into its own
  // partial density rho + y*N (summed in fixed group order afterwards).
  const int groups = MODE == 0 ? 1 : gridDim.y;
  const int tbeg = MODE == 0 ? blockIdx.x : blockIdx.y;
  const int tend = MODE == 0 ? tbeg + 1 : pairs;
  if (MODE == 1) rho += int64_t(blockIdx.y) * n0 * n1 * n2;
  for (int t = tbeg; t < tend; t += groups) {
    c2<F>* col = sticks + int64_t(t) * n0 * Spad + int64_t(xi) * Spad;
    for (int i = threadIdx.x; i < Z * n1; i += blockDim.x) P[i] = {F(0), F(0)};
    __syncthreads();
    for (int s = threadIdx.x; s < Spad; s += blockDim.x) {
      const int2 yz = stick_yz[s];
      if (yz.x >= 0) P[yz.y * n1 + yz.x] = col[s];
    }
    __syncthreads();
    lds_fft<true, F, NT>(P, Z, n1, 1, ry, rty);
    double wa = 0.0, wb = 0.0;
    if (MODE == 1) {
      wa = weights[2 * t] * inv_volume;
      wb = (2 * t + 1 < 2 * pairs) ? weights[2 * t + 1] * inv_volume : 0.0;
    }
    for (int y0 = 0; y0 < n1; y0 += chunk) {
      const int c = min(chunk, n1 - y0);
      for (int i = threadIdx.x; i < c * n2; i += blockDim.x) Rb[i] = {F(0), F(0)};
      __syncthreads();
      for (int i = threadIdx.x; i < c * Z; i += blockDim.x) {
        const int r = i / Z, zc = i - r * Z;
        Rb[r * n2 + z_of_zc[zc]] = P[zc * n1 + y0 + r];
      }
      __syncthreads();
      lds_fft<true, F, NT>(Rb, c, n2, 1, rz, rtz);
      if (MODE == 0) {
        const F* v = potential + (int64_t(xi) * n1 + y0) * n2;
        for (int i = threadIdx.x; i < c * n2; i += blockDim.x) Rb[i] = cscale(Rb[i], v[i]);
        __syncthreads();
        lds_fft<false, F, NT>(Rb, c, n2, 1, rz, rtz);
        for (int i = threadIdx.x; i < c * Z; i += blockDim.x) {
          const int r = i / Z, zc = i - r * Z;
          P[zc * n1 + y0 + r] = Rb[r * n2 + z_of_zc[zc]];
        }
        __syncthreads();
      } else {
        double* out = rho + (int64_t(xi) * n1 + y0) * n2;
        // With `keep`: the pair's real-space field too, [pair][x][y][z].
        c2<F>* kp = keep ? keep + int64_t(t) * n0 * n1 * n2 + (int64_t(xi) * n1 + y0) * n2 : nullptr;
        for (int i = threadIdx.x; i < c * n2; i += blockDim.x) {
          const double re = double(Rb[i].x), im = double(Rb[i].y);
          const double d = wa * re * re + wb * im * im;
          out[i] = t == tbeg ? d : out[i] + d;
          if (kp) kp[i] = Rb[i];
        }
        __syncthreads();
      }
    }
    if (MODE == 0) {
      lds_fft<false, F, NT>(P, Z, n1, 1, ry, rty);
      for (int s = threadIdx.x; s < Spad; s += blockDim.x) {
        const int2 yz = stick_yz[s];
        if (yz.x >= 0) col[s] = P[yz.y * n1 + yz.x];
      }
    }
  }
}

// Kept-field pass of the warm start (apply_local_change,
// distributed_real_gamma.rs:1370-1438): from a pair's kept real-space field
// psi(r) (written by the band density of the same columns), dV psi, forward z
// per row chunk, forward y, back to the sticks. Only the r -> G half runs.
template <typename F, int NT>
__global__ __launch_bounds__(NT) void k_plane_kept(
    c2<F>* __restrict__ sticks, int n0, int n1, int n2, int Spad, int Z, int chunk,
    const int2* __restrict__ stick_yz, const int* __restrict__ z_of_zc, const c2<double>* __restrict__ roots,
    RadixPlan ry, RadixPlan rz, const F* __restrict__ dv, const c2<F>* __restrict__ kept) {
  extern __shared__ char smem[];
  c2<F>* P = reinterpret_cast<c2<F>*>(smem);  // [Z][n1]
  c2<F>* Rb = P + Z * n1;                     // [chunk][n2]
  c2<F>* rty = Rb + chunk * n2;
  c2<F>* rtz = rty + n1;
  const int t = blockIdx.x, xi = blockIdx.y;
  stage_roots(rty, roots + n0, n1);
  stage_roots(rtz, roots + n0 + n1, n2);
  const c2<F>* kp = kept + int64_t(t) * n0 * n1 * n2 + int64_t(xi) * n1 * n2;
  for (int y0 = 0; y0 < n1; y0 += chunk) {
    const int c = min(chunk, n1 - y0);
    const F* v = dv + (int64_t(xi) * n1 + y0) * n2;
    __syncthreads();
    for (int i = threadIdx.x; i < c * n2; i += blockDim.x) Rb[i] = cscale(kp[int64_t(y0) * n2 + i], v[i]);
    __syncthreads();
    lds_fft<false, F, NT>(Rb, c, n2, 1, rz, rtz);
    for (int i = threadIdx.x; i < c * Z; i += blockDim.x) {
      const int r = i / Z, zc = i - r * Z;
      P[zc * n1 + y0 + r] = Rb[r * n2 + z_of_zc[zc]];
    }
  }
  __syncthreads();
  lds_fft<false, F, NT>(P, Z, n1, 1, ry, rty);
  c2<F>* col = sticks + int64_t(t) * n0 * Spad + int64_t(xi) * Spad;
  for (int s = threadIdx.x; s < Spad; s += blockDim.x) {
    const int2 yz = stick_yz[s];
    if (yz.x >= 0) col[s] = P[yz.y * n1 + yz.x];
  }
}

// Split plane pass, for planes whose Z x n1 values do not fit one workgroup's
// LDS (the Si512 grid's 144-point planes in FP64): the same transforms as
// k_plane through a plane buffer [pair][x][Z][n1] in HBM. k_plane_y does the
// y transforms of zb z-lines at a time (inverse: sticks -> buffer; forward:
// buffer -> sticks); k_plane_z the z transforms of a chunk of y rows (MODE 0:
// inverse, * V(r), forward, back into the buffer; MODE 1: inverse and the
// density sum). `pair0` is the first pair of the batch the buffer holds.
template <bool INV, typename F, int NT>
__global__ __launch_bounds__(NT) void k_plane_y(c2<F>* __restrict__ sticks, int n0, int n1, int Spad, int Z, int zb,
                                                const int2* __restrict__ stick_yz, const c2<double>* __restrict__ roots,
                                                RadixPlan ry, int pair0, c2<F>* __restrict__ planes) {
  extern __shared__ char smem[];
  c2<F>* P = reinterpret_cast<c2<F>*>(smem);  // [zb][n1]
  c2<F>* rty = P + zb * n1;
  const int tl = blockIdx.x, xi = blockIdx.y, z0 = blockIdx.z * zb;
  const int nz = min(zb, Z - z0);
  stage_roots(rty, roots + n0, n1);
  c2<F>* col = sticks + (int64_t(pair0 + tl) * n0 + xi) * Spad;
  c2<F>* pl = planes + (int64_t(tl) * n0 + xi) * Z * n1 + int64_t(z0) * n1;
  if (INV) {
    for (int i = threadIdx.x; i < nz * n1; i += blockDim.x) P[i] = {F(0), F(0)};
    __syncthreads();
    for (int s = threadIdx.x; s < Spad; s += blockDim.x) {
      const int2 yz = stick_yz[s];
      if (yz.x >= 0 && yz.y >= z0 && yz.y < z0 + nz) P[(yz.y - z0) * n1 + yz.x] = col[s];
    }
    __syncthreads();
    lds_fft<true, F, NT>(P, nz, n1, 1, ry, rty);
    for (int i = threadIdx.x; i < nz * n1; i += blockDim.x) pl[i] = P[i];
  } else {
    for (int i = threadIdx.x; i < nz * n1; i += blockDim.x) P[i] = pl[i];
    __syncthreads();
    lds_fft<false, F, NT>(P, nz, n1, 1, ry, rty);
    for (int s = threadIdx.x; s < Spad; s += blockDim.x) {
      const int2 yz = stick_yz[s];
      if (yz.x >= 0 && yz.y >= z0 && yz.y < z0 + nz) col[s] = P[(yz.y - z0) * n1 + yz.x];
    }
  }
}

// MODE 0: grid (pairs of the batch, n0, y chunks). MODE 1: grid (n0, y chunks,
// groups); group g sums pairs g, g + groups, ... of the batch into its partial
// density rho + g N, after the earlier batches (`first`: the batch that
// writes instead of adding).
template <int MODE, typename F, int NT>
__global__ __launch_bounds__(NT) void k_plane_z(c2<F>* __restrict__ planes, int n0, int n1, int n2, int Z, int chunk,
                                                const int* __restrict__ z_of_zc, const c2<double>* __restrict__ roots,
                                                RadixPlan rz, const F* __restrict__ potential, int batch_pairs,
                                                int pair0, int total_pairs, const double* __restrict__ weights,
                                                double inv_volume, int first, double* __restrict__ rho) {
  extern __shared__ char smem[];
  c2<F>* Rb = reinterpret_cast<c2<F>*>(smem);  // [chunk][n2]
  c2<F>* rtz = Rb + chunk * n2;
  const int xi = MODE == 0 ? blockIdx.y : blockIdx.x;
  const int y0 = (MODE == 0 ? blockIdx.z : blockIdx.y) * chunk;
  const int c = min(chunk, n1 - y0);
  stage_roots(rtz, roots + n0 + n1, n2);
  const int groups = MODE == 0 ? 1 : gridDim.z;
  const int tbeg = MODE == 0 ? blockIdx.x : blockIdx.z;
  const int tend = MODE == 0 ? tbeg + 1 : batch_pairs;
  if (MODE == 1) rho += int64_t(blockIdx.z) * n0 * n1 * n2;
  for (int tl = tbeg; tl < tend; tl += groups) {
    c2<F>* pl = planes + (int64_t(tl) * n0 + xi) * Z * n1;
    for (int i = threadIdx.x; i < c * n2; i += blockDim.x) Rb[i] = {F(0), F(0)};
    __syncthreads();
    for (int i = threadIdx.x; i < c * Z; i += blockDim.x) {
      const int zc = i / c, r = i - zc * c;
      Rb[r * n2 + z_of_zc[zc]] = pl[zc * n1 + y0 + r];
    }
    __syncthreads();
    lds_fft<true, F, NT>(Rb, c, n2, 1, rz, rtz);
    if (MODE == 0) {
      const F* v = potential + (int64_t(xi) * n1 + y0) * n2;
      for (int i = threadIdx.x; i < c * n2; i += blockDim.x) Rb[i] = cscale(Rb[i], v[i]);
      __syncthreads();
      lds_fft<false, F, NT>(Rb, c, n2, 1, rz, rtz);
      for (int i = threadIdx.x; i < c * Z; i += blockDim.x) {
        const int zc = i / c, r = i - zc * c;
        pl[zc * n1 + y0 + r] = Rb[r * n2 + z_of_zc[zc]];
      }
      __syncthreads();
    } else {
      const int t = pair0 + tl;
      const double wa = weights[2 * t] * inv_volume;
      const double wb = (2 * t + 1 < 2 * total_pairs) ? weights[2 * t + 1] * inv_volume : 0.0;
      double* out = rho + (int64_t(xi) * n1 + y0) * n2;
      const bool write = first && tl == tbeg;
      for (int i = threadIdx.x; i < c * n2; i += blockDim.x) {
        const double re = double(Rb[i].x), im = double(Rb[i].y);
        const double d = wa * re * re + wb * im * im;
        out[i] = write ? d : out[i] + d;
      }
      __syncthreads();
    }
  }
}

__global__ void k_to_float(const double* __restrict__ src, int64_t n, float* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = float(src[i]);
}

// rho = sum over groups of the partial densities, in group order.
// rho = sum of the group partials (accumulate: rho += that sum).
__global__ void k_sum_groups(const double* __restrict__ part, int64_t N, int groups, double* __restrict__ rho,
                             int accumulate) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double s = 0.0;
  for (int g = 0; g < groups; ++g) s += part[int64_t(g) * N + i];
  rho[i] = accumulate ? rho[i] + s : s;
}

// W = D Q per atom block (distributed_real_gamma.rs:1183-1195).
__global__ void k_coupling(const double* __restrict__ q, double* __restrict__ w, int64_t nproj,
                           int m, int nc, const double* __restrict__ blockd) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nproj * m) return;
  const int64_t row = e % nproj, col = e / nproj;
  const int64_t start = (row / nc) * nc;
  const int a = int(row - start);
  const double* qc = q + col * nproj + start;
  double s = 0.0;
  for (int b = 0; b < nc; ++b) s += blockd[a * nc + b] * qc[b];
  w[e] = s;
}
}  // namespace

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------
// Passes of radix 9 and 8 first when `big` (fewer LDS round trips), then 4,
// 2, 3, 5.
RadixPlan radix_plan(int n, bool big) {
  RadixPlan rp{};
  rp.n = n;
  int r = n;
  auto take = [&](int f) {
    while (r % f == 0) {
      if (rp.count >= 12) fail(BD_ERR_INVALID_PARAMETER, "FFT length has too many factors");
      rp.radix[rp.count++] = f;
      r /= f;
    }
  };
  if (big) {
    take(9);
    take(8);
  }
  take(4);
  take(2);
  take(3);
  take(5);
  if (r != 1) fail(BD_ERR_INVALID_PARAMETER, "FFT length must be 2,3,5-smooth");
  return rp;
}

StickPlan build_stick_plan(const System& s) {
  StickPlan p{};
  p.n0 = s.dims[0];
  p.n1 = s.dims[1];
  p.n2 = s.dims[2];
  p.rx = radix_plan(p.n0, true);
  p.ry = radix_plan(p.n1, true);
  p.rz = radix_plan(p.n2, true);
  auto split = [&](int64_t f, int& i0, int& i1, int& i2) {
    i2 = int(f % p.n2);
    i1 = int((f / p.n2) % p.n1);
    i0 = int(f / (int64_t(p.n1) * p.n2));
  };
  // Occupied sticks keyed (z, y), as the reference keys sticks (distributed_fft.rs:296-312).
  std::set<std::pair<int, int>> keys;
  for (int64_t r = 0; r < s.nrep; ++r)
    for (int64_t f : {s.plus_flat[r], s.minus_flat[r]}) {
      int a, b, c;
      split(f, a, b, c);
      keys.insert({c, b});
    }
  std::vector<std::pair<int, int>> sticks(keys.begin(), keys.end());
  p.S = int(sticks.size());
  std::map<std::pair<int, int>, int> index;
  for (int i = 0; i < p.S; ++i) index[sticks[i]] = i;
  std::set<int> zs;
  for (auto& k : sticks) zs.insert(k.first);
  p.z_of_zc.assign(zs.begin(), zs.end());
  p.Z = int(p.z_of_zc.size());
  std::map<int, int> zc_of;
  for (int i = 0; i < p.Z; ++i) zc_of[p.z_of_zc[i]] = i;
  // Tile size: the largest power of two <= 32 whose x transform fits a workgroup.
  // Radix-9/8 passes only where the batches fit as with radices 2-5.
  if (!fft_fits(32, p.rx)) p.rx = radix_plan(p.n0, false);
  if (!fft_fits(p.Z, p.ry)) p.ry = radix_plan(p.n1, false);
  if (!fft_fits(std::min(16, p.n1), p.rz)) p.rz = radix_plan(p.n2, false);
  p.T = 32;
  while (p.T > 1 && !fft_fits(p.T, p.rx)) p.T /= 2;
  if (!fft_fits(p.T, p.rx)) fail(BD_ERR_INVALID_PARAMETER, "x FFT does not fit a workgroup");
  p.chunk = 16;
  while (p.chunk > 1 && !fft_fits(p.chunk, p.rz)) p.chunk /= 2;
  p.chunk = std::min(p.chunk, p.n1);
  // Mirror units (stick + its inversion partner) packed whole into tiles, so
  // +G and -G of every representative share a tile.
  std::vector<int> tile_of(p.S, -1), slot_of(p.S, -1);
  int tile = 0, used = 0;
  for (int i = 0; i < p.S; ++i) {
    if (tile_of[i] >= 0) continue;
    auto [z, y] = sticks[i];
    std::pair<int, int> mk{(p.n2 - z) % p.n2, (p.n1 - y) % p.n1};
    auto it = index.find(mk);
    if (it == index.end()) fail(BD_ERR_INVALID_PARAMETER, "stick set is not inversion symmetric");
    int j = it->second;
    int need = (j == i) ? 1 : 2;
    if (need > p.T) fail(BD_ERR_INVALID_PARAMETER, "tile too small for a mirror unit");
    if (used + need > p.T) {
      ++tile;
      used = 0;
    }
    tile_of[i] = tile;
    slot_of[i] = used++;
    if (j != i) {
      tile_of[j] = tile;
      slot_of[j] = used++;
    }
  }
  p.ntiles = tile + 1;
  p.Spad = p.ntiles * p.T;
  p.stick_yz.assign(p.Spad, make_int2(-1, -1));
  for (int i = 0; i < p.S; ++i)
    p.stick_yz[tile_of[i] * p.T + slot_of[i]] = make_int2(sticks[i].second, zc_of[sticks[i].first]);
  // Representative entries grouped by tile.
  std::vector<std::vector<int4>> per(p.ntiles);
  for (int64_t r = 0; r < s.nrep; ++r) {
    int a0, b0, c0, a1, b1, c1;
    split(s.plus_flat[r], a0, b0, c0);
    split(s.minus_flat[r], a1, b1, c1);
    int sp = index[{c0, b0}], sm = index[{c1, b1}];
    if (tile_of[sp] != tile_of[sm]) fail(BD_ERR_APPLICATION, "mirror pair split across tiles");
    per[tile_of[sp]].push_back(
        make_int4(int(r), slot_of[sp] * p.n0 + a0, slot_of[sm] * p.n0 + a1, 0));
  }
  p.tile_off.assign(1, 0);
  for (auto& v : per) {
    p.entries.insert(p.entries.end(), v.begin(), v.end());
    p.tile_off.push_back(int(p.entries.size()));
  }
  return p;
}

// The k sphere has no inversion partner to keep in a tile: occupied sticks
// (keyed (z, y) as at Gamma) fill tiles of T in key order.
StickPlan build_k_stick_plan(const System& s, const KBasis& kb) {
  StickPlan p{};
  p.n0 = s.dims[0];
  p.n1 = s.dims[1];
  p.n2 = s.dims[2];
  p.rx = radix_plan(p.n0, true);
  p.ry = radix_plan(p.n1, true);
  p.rz = radix_plan(p.n2, true);
  auto split = [&](int64_t f, int& i0, int& i1, int& i2) {
    i2 = int(f % p.n2);
    i1 = int((f / p.n2) % p.n1);
    i0 = int(f / (int64_t(p.n1) * p.n2));
  };
  std::set<std::pair<int, int>> keys;
  for (int64_t f : kb.flat) {
    int a, b, c;
    split(f, a, b, c);
    keys.insert({c, b});
  }
  std::vector<std::pair<int, int>> sticks(keys.begin(), keys.end());
  p.S = int(sticks.size());
  std::map<std::pair<int, int>, int> index;
  for (int i = 0; i < p.S; ++i) index[sticks[i]] = i;
  std::set<int> zs;
  for (auto& k : sticks) zs.insert(k.first);
  p.z_of_zc.assign(zs.begin(), zs.end());
  p.Z = int(p.z_of_zc.size());
  std::map<int, int> zc_of;
  for (int i = 0; i < p.Z; ++i) zc_of[p.z_of_zc[i]] = i;
  // Radix-9/8 passes only where the batches fit as with radices 2-5.
  if (!fft_fits(32, p.rx)) p.rx = radix_plan(p.n0, false);
  if (!fft_fits(p.Z, p.ry)) p.ry = radix_plan(p.n1, false);
  if (!fft_fits(std::min(16, p.n1), p.rz)) p.rz = radix_plan(p.n2, false);
  p.T = 32;
  while (p.T > 1 && !fft_fits(p.T, p.rx)) p.T /= 2;
  if (!fft_fits(p.T, p.rx)) fail(BD_ERR_INVALID_PARAMETER, "x FFT does not fit a workgroup");
  p.chunk = 16;
  while (p.chunk > 1 && !fft_fits(p.chunk, p.rz)) p.chunk /= 2;
  p.chunk = std::min(p.chunk, p.n1);
  p.ntiles = (p.S + p.T - 1) / p.T;
  p.Spad = p.ntiles * p.T;
  p.stick_yz.assign(p.Spad, make_int2(-1, -1));
  for (int i = 0; i < p.S; ++i) p.stick_yz[i] = make_int2(sticks[i].second, zc_of[sticks[i].first]);
  std::vector<std::vector<int4>> per(p.ntiles);
  for (int64_t g = 0; g < kb.nk; ++g) {
    int a, b, c;
    split(kb.flat[g], a, b, c);
    const int st = index[{c, b}];
    per[st / p.T].push_back(make_int4(int(g), (st % p.T) * p.n0 + a, -1, 0));
    p.fpos.push_back(st * p.n0 + a);
  }
  p.tile_off.assign(1, 0);
  for (auto& v : per) {
    p.entries.insert(p.entries.end(), v.begin(), v.end());
    p.tile_off.push_back(int(p.entries.size()));
  }
  return p;
}

// ---------------------------------------------------------------------------
// Operator
// ---------------------------------------------------------------------------
size_t GammaOp::plane_lds_bytes(int chunk) const {
  const size_t e = fft32() ? sizeof(c2<float>) : sizeof(c2<double>);
  return size_t(plan.Z * plan.n1 + chunk * plan.n2 + plan.n1 + plan.n2) * e;
}

size_t GammaOp::elem_bytes() const { return fft32() ? sizeof(c2<float>) : sizeof(c2<double>); }

// A (threads, chunk) shape of the plane pass: the fused k_plane holds the
// whole Z x n1 plane; the split pass holds zb z-lines (y transforms) or a
// chunk of y rows (z transforms) at a time.
bool GammaOp::plane_shape_valid(int threads, int chunk) const {
  const StickPlan& p = plan;
  const size_t e = elem_bytes(), lim = 160 * 1024;
  if ((threads != 256 && threads != 512) || chunk < 1 || chunk > p.n1) return false;
  if (!fft_fits_threads(chunk, p.rz, threads)) return false;
  if (p.zb > 0)
    return fft_fits_threads(p.zb, p.ry, threads) && size_t(p.zb * p.n1 + p.n1) * e <= lim &&
           size_t(chunk * p.n2 + p.n2) * e <= lim;
  return fft_fits_threads(p.Z, p.ry, threads) && plane_lds_bytes(chunk) <= lim;
}

// Switch to the split plane pass with zb z-lines per y batch (0: the largest
// that fits), and size its plane buffer: up to ~2 GiB of pairs per batch, at
// least kDensityGroups pairs (or all).
void GammaOp::enable_split(int zb) {
  const StickPlan& p = plan;
  const size_t e = elem_bytes(), lim = 160 * 1024;
  auto fits = [&](int z) {
    return z >= 1 && z <= p.Z && fft_fits_threads(z, p.ry, 256) && size_t(z * p.n1 + p.n1) * e <= lim;
  };
  if (zb == 0) {
    zb = p.Z;
    while (zb > 1 && !fits(zb)) --zb;
  }
  if (!fits(zb)) fail(BD_ERR_INVALID_PARAMETER, "no y-line batch of the split plane pass fits a workgroup");
  const size_t plane = size_t(p.n0) * p.Z * p.n1 * e;
  const int pairs = slots(capacity);
  int batch = int(std::max<size_t>(1, (size_t(2) << 30) / plane));
  batch = std::min(pairs, std::max(batch, kDensityGroups));
  if (d_planes.n < (size_t(batch) * plane + 15) / 16) {
    DevBuf<c2<double>> fresh;
    fresh.alloc((size_t(batch) * plane + 15) / 16);
    std::swap(d_planes.p, fresh.p);
    std::swap(d_planes.n, fresh.n);
  }
  plan.zb = zb;
  plane_batch = batch;
  plane_threads = 256;
  int c = std::min(plan.chunk, p.n1);
  while (c > 1 && !plane_shape_valid(256, c)) --c;
  plan.chunk = c;
}

// The k kernels move a coefficient's (Re, Im) as one 16-byte access.
static bool complex_aligned(const double* p, int64_t ld) {
  return (reinterpret_cast<uintptr_t>(p) & 15) == 0 && (ld & 1) == 0;
}

static size_t stick_lds(const StickPlan& p, size_t elem) { return size_t(p.T * p.n0 + p.n0) * elem; }
static size_t plane_lds(const StickPlan& p, size_t elem) {
  return size_t(p.Z * p.n1 + p.chunk * p.n2 + p.n1 + p.n2) * elem;
}

template <typename F>
static void allow_lds(const StickPlan& plan) {
  const size_t e = sizeof(c2<F>);
  // Every plane instantiation may take up to the whole LDS (the chunk is tuned).
  for (const void* f : {reinterpret_cast<const void*>(&k_plane<0, F, 256>),
                        reinterpret_cast<const void*>(&k_plane<1, F, 256>),
                        reinterpret_cast<const void*>(&k_plane<0, F, 512>),
                        reinterpret_cast<const void*>(&k_plane<1, F, 512>),
                        reinterpret_cast<const void*>(&k_plane_y<true, F, 256>),
                        reinterpret_cast<const void*>(&k_plane_y<false, F, 256>),
                        reinterpret_cast<const void*>(&k_plane_y<true, F, 512>),
                        reinterpret_cast<const void*>(&k_plane_y<false, F, 512>),
                        reinterpret_cast<const void*>(&k_plane_z<0, F, 256>),
                        reinterpret_cast<const void*>(&k_plane_z<1, F, 256>),
                        reinterpret_cast<const void*>(&k_plane_z<0, F, 512>),
                        reinterpret_cast<const void*>(&k_plane_z<1, F, 512>),
                        reinterpret_cast<const void*>(&k_plane_kept<F, 256>),
                        reinterpret_cast<const void*>(&k_plane_kept<F, 512>)})
    BD_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
  // The same largest value for every op: the attribute is per kernel and
  // process-wide, so a per-plan size set by one op could undercut another's.
  if (stick_lds(plan, e) > 65536)
    for (const void* f : {reinterpret_cast<const void*>(&k_stick_inverse<F>),
                          reinterpret_cast<const void*>(&k_stick_forward<F>),
                          reinterpret_cast<const void*>(&k_kstick_inverse<F>),
                          reinterpret_cast<const void*>(&k_kstick_forward<F>)})
      BD_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
}

GammaOp::GammaOp(const System& s, const KBasis* k, int cap, int prec) : sys(&s), kb(k) {
  if (prec != BD_FP64 && prec != BD_MIXED && prec != BD_FP32_HX)
    fail(BD_ERR_INVALID_PARAMETER, "unknown precision mode");
  if (cap <= 0) fail(BD_ERR_INVALID_PARAMETER, "capacity must be positive");
  int dev = 0, count = 0;
  BD_HIP(hipGetDeviceCount(&count));
  if (count <= 0) fail(BD_ERR_DEVICE, "no HIP device");
  BD_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  BD_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
    fail(BD_ERR_DEVICE, std::string("the library is compiled for gfx950, device is ") + prop.gcnArchName);
  n = kb ? 2 * kb->nk : s.n;
  nrep = kb ? kb->nk : s.nrep;
  for (int i = 0; i < 3; ++i) dims[i] = s.dims[i];
  N = int64_t(dims[0]) * dims[1] * dims[2];
  capacity = cap;
  precision = prec;
  gemm.bf16x3 = prec == BD_MIXED;
  nproj = kb ? 2 * s.nproj : s.nproj;
  plan = kb ? build_k_stick_plan(s, *kb) : build_stick_plan(s);
  size_t lds_max = size_t(prop.sharedMemPerBlock);
  const size_t elem = fft32() ? sizeof(c2<float>) : sizeof(c2<double>);
  if (stick_lds(plan, elem) > lds_max) fail(BD_ERR_INVALID_PARAMETER, "x FFT tile exceeds the device LDS");
  if (fft32()) allow_lds<float>(plan);
  else allow_lds<double>(plan);
  BD_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  // Roots e^{-2 pi i k/n} per axis, in long double.
  std::vector<c2<double>> roots;
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < dims[a]; ++k) {
      long double ang = -2.0L * 3.14159265358979323846264338327950288L * k / dims[a];
      roots.push_back({double(cosl(ang)), double(sinl(ang))});
    }
  d_roots.alloc(roots.size());
  d_roots.upload(roots.data(), roots.size(), stream);
  d_tile_off.alloc(plan.tile_off.size());
  d_tile_off.upload(plan.tile_off.data(), plan.tile_off.size(), stream);
  d_entries.alloc(plan.entries.size());
  d_entries.upload(plan.entries.data(), plan.entries.size(), stream);
  d_stick_yz.alloc(plan.stick_yz.size());
  d_stick_yz.upload(plan.stick_yz.data(), plan.stick_yz.size(), stream);
  d_zc.alloc(plan.z_of_zc.size());
  d_zc.upload(plan.z_of_zc.data(), plan.z_of_zc.size(), stream);
  d_kinetic.alloc(n);
  d_kinetic.upload(kb ? kb->kinetic2.data() : s.kinetic_orc.data(), n, stream);
  d_potential.alloc(N);
  if (fft32()) d_potential32.alloc(N);
  if (nproj > 0) {
    d_proj.alloc(size_t(n) * nproj);
    d_proj.upload(kb ? kb->proj2.data() : s.projectors_orc.data(), size_t(n) * nproj, stream);
    nc = s.channel_count;
    nblocks = int(s.blocks.size());
    d_blockd.alloc(s.block_d.size());
    d_blockd.upload(s.block_d.data(), s.block_d.size(), stream);
    d_q.alloc(size_t(nproj) * cap);
    d_w.alloc(size_t(nproj) * cap);
    d_work.alloc(size_t(Gemm::tn_work(nproj, cap, n)));
  }
  const int pairs = slots(cap);
  d_sticks.alloc(size_t(pairs) * plan.n0 * plan.Spad);
  // Planes that do not fit one workgroup go through the split pass.
  if (!fft_fits(plan.Z, plan.ry) || plane_lds(plan, elem) > lds_max) enable_split(0);
  // Host-pointer entry points stage through these (no allocation after create).
  d_x.alloc(size_t(n) * cap);
  d_y.alloc(size_t(n) * cap);
  d_weights.alloc(size_t(2 * pairs));
  d_rho_part.alloc(size_t(std::min(pairs, kDensityGroups)) * N);
  if (kb) {
    d_fpos.alloc(plan.fpos.size());
    d_fpos.upload(plan.fpos.data(), plan.fpos.size(), stream);
    // Default shape: 256 threads, every stick in one x batch and the whole
    // plane in one z batch when they fit, halved until they do.
    FusedShape f;
    f.xlines = plan.S;
    while (f.xlines > 1 && !fft_fits_threads(f.xlines, plan.rx, f.threads)) f.xlines = (f.xlines + 1) / 2;
    f.zrows = plan.n1;
    while (f.zrows > 1 && !fft_fits_threads(f.zrows, plan.rz, f.threads)) f.zrows = (f.zrows + 1) / 2;
    // Planes per pass: as many as keep two workgroups per CU (80 KB each).
    for (int xp = 8; xp >= 1; --xp) {
      f.xplanes = xp;
      if (fused_valid(f) && fused_lds(f) <= 80 * 1024) break;
    }
    if (!fused_valid(f)) f.xplanes = 1;
    fused = fused_valid(f) && fused_lds(f) <= lds_max;
    if (fused) fshape = f;
  }
  for (auto& e : ev) BD_HIP(hipEventCreate(&e));
  BD_HIP(hipStreamSynchronize(stream));
}

GammaOp::~GammaOp() {
  destroy_davidson_work(dav);
  destroy_complex_work(cdav);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

void GammaOp::set_potential_dev(const double* v, uint64_t ep) {
  if (has_potential && ep <= epoch) fail(BD_ERR_INVALID_PARAMETER, "potential epoch must increase");
  BD_HIP(hipMemcpyAsync(d_potential.p, v, size_t(N) * sizeof(double), hipMemcpyDeviceToDevice, stream));
  epoch = ep;
  has_potential = true;
  refresh_potential32();
}

void GammaOp::refresh_potential32() {
  if (!fft32()) return;
  hipLaunchKernelGGL(k_to_float, dim3(unsigned((N + 255) / 256)), dim3(256), 0, stream, d_potential.p, N,
                     d_potential32.p);
  BD_HIP(hipGetLastError());
}

// The split plane pass over all `pairs` FFT slots, plane_batch at a time:
// MODE 0 transforms the sticks in place with V(r); MODE 1 sums the band
// density into kDensityGroups-style group partials at rho_part.
template <int MODE, typename F, int NT>
void GammaOp::split_pass(int pairs, const F* pot, const double* w, double* rho_part) {
  const StickPlan& p = plan;
  const size_t e = sizeof(c2<F>);
  c2<F>* sticks = reinterpret_cast<c2<F>*>(d_sticks.p);
  c2<F>* planes = reinterpret_cast<c2<F>*>(d_planes.p);
  const size_t ly = size_t(p.zb * p.n1 + p.n1) * e, lz = size_t(p.chunk * p.n2 + p.n2) * e;
  const unsigned zbatches = unsigned((p.Z + p.zb - 1) / p.zb), ychunks = unsigned((p.n1 + p.chunk - 1) / p.chunk);
  const int groups = std::min(pairs, kDensityGroups);
  for (int t0 = 0; t0 < pairs; t0 += plane_batch) {
    const int bp = std::min(plane_batch, pairs - t0);
    hipLaunchKernelGGL((k_plane_y<true, F, NT>), dim3(bp, p.n0, zbatches), dim3(NT), ly, stream, sticks, p.n0, p.n1,
                       p.Spad, p.Z, p.zb, d_stick_yz.p, d_roots.p, p.ry, t0, planes);
    if (MODE == 0) {
      hipLaunchKernelGGL((k_plane_z<0, F, NT>), dim3(bp, p.n0, ychunks), dim3(NT), lz, stream, planes, p.n0, p.n1,
                         p.n2, p.Z, p.chunk, d_zc.p, d_roots.p, p.rz, pot, bp, t0, pairs, nullptr, 0.0, 0, nullptr);
      hipLaunchKernelGGL((k_plane_y<false, F, NT>), dim3(bp, p.n0, zbatches), dim3(NT), ly, stream, sticks, p.n0,
                         p.n1, p.Spad, p.Z, p.zb, d_stick_yz.p, d_roots.p, p.ry, t0, planes);
    } else {
      hipLaunchKernelGGL((k_plane_z<1, F, NT>), dim3(p.n0, ychunks, groups), dim3(NT), lz, stream, planes, p.n0,
                         p.n1, p.n2, p.Z, p.chunk, d_zc.p, d_roots.p, p.rz, nullptr, bp, t0, pairs, w,
                         1.0 / sys->volume, t0 == 0 ? 1 : 0, rho_part);
    }
  }
}

template <typename F>
void GammaOp::local_apply(const double* x, int64_t ldx, int m, double* y, int64_t ldy) {
  const StickPlan& p = plan;
  const int pairs = slots(m);
  const size_t e = sizeof(c2<F>);
  c2<F>* sticks = reinterpret_cast<c2<F>*>(d_sticks.p);
  const F* pot;
  if constexpr (std::is_same_v<F, float>) pot = d_potential32.p;
  else pot = d_potential.p;
  if (profiling) BD_HIP(hipEventRecord(ev[0], stream));
  if (kb)
    hipLaunchKernelGGL(k_kstick_inverse<F>, dim3(p.ntiles, pairs), dim3(kThreads), stick_lds(p, e), stream, x,
                       ldx, p.n0, p.T, p.Spad, d_tile_off.p, d_entries.p, d_roots.p, p.rx, sticks);
  else
    hipLaunchKernelGGL(k_stick_inverse<F>, dim3(p.ntiles, pairs), dim3(kThreads), stick_lds(p, e), stream, x,
                       ldx, m, p.n0, p.T, p.Spad, d_tile_off.p, d_entries.p, d_roots.p, p.rx, sticks);
  if (profiling) BD_HIP(hipEventRecord(ev[1], stream));
  if (p.zb > 0) {
    if (plane_threads == 512) split_pass<0, F, 512>(pairs, pot, nullptr, nullptr);
    else split_pass<0, F, 256>(pairs, pot, nullptr, nullptr);
  } else if (plane_threads == 512) {
    hipLaunchKernelGGL((k_plane<0, F, 512>), dim3(pairs, p.n0), dim3(512), plane_lds(p, e), stream, sticks, p.n0,
                       p.n1, p.n2, p.Spad, p.Z, p.chunk, d_stick_yz.p, d_zc.p, d_roots.p, p.ry, p.rz, pot, pairs,
                       nullptr, 0.0, nullptr, nullptr);
  } else {
    hipLaunchKernelGGL((k_plane<0, F, 256>), dim3(pairs, p.n0), dim3(256), plane_lds(p, e), stream, sticks, p.n0,
                       p.n1, p.n2, p.Spad, p.Z, p.chunk, d_stick_yz.p, d_zc.p, d_roots.p, p.ry, p.rz, pot, pairs,
                       nullptr, 0.0, nullptr, nullptr);
  }
  if (profiling) BD_HIP(hipEventRecord(ev[2], stream));
  if (kb)
    hipLaunchKernelGGL(k_kstick_forward<F>, dim3(p.ntiles, pairs), dim3(kThreads), stick_lds(p, e), stream, x,
                       ldx, y, ldy, p.n0, p.T, p.Spad, d_tile_off.p, d_entries.p, d_roots.p, p.rx, sticks,
                       d_kinetic.p, 1.0 / double(N));
  else
    hipLaunchKernelGGL(k_stick_forward<F>, dim3(p.ntiles, pairs), dim3(kThreads), stick_lds(p, e), stream, x,
                       ldx, y, ldy, m, p.n0, p.T, p.Spad, d_tile_off.p, d_entries.p, d_roots.p, p.rx,
                       sticks, d_kinetic.p, 1.0 / double(N), 0);
  if (profiling) BD_HIP(hipEventRecord(ev[3], stream));
}

template <typename F>
void GammaOp::density(const double* x, int64_t ldx, int m, const double* w, double* rho, bool accumulate) {
  const StickPlan& p = plan;
  const int pairs = slots(m);
  const size_t e = sizeof(c2<F>);
  c2<F>* sticks = reinterpret_cast<c2<F>*>(d_sticks.p);
  if (kb)
    hipLaunchKernelGGL(k_kstick_inverse<F>, dim3(p.ntiles, pairs), dim3(kThreads), stick_lds(p, e), stream, x,
                       ldx, p.n0, p.T, p.Spad, d_tile_off.p, d_entries.p, d_roots.p, p.rx, sticks);
  else
    hipLaunchKernelGGL(k_stick_inverse<F>, dim3(p.ntiles, pairs), dim3(kThreads), stick_lds(p, e), stream, x,
                       ldx, m, p.n0, p.T, p.Spad, d_tile_off.p, d_entries.p, d_roots.p, p.rx, sticks);
  const int groups = std::min(pairs, kDensityGroups);
  // Warm start: keep the pairs' real-space fields (Gamma, fused plane pass).
  c2<F>* keep = kept_on ? reinterpret_cast<c2<F>*>(d_kept.p) : nullptr;
  kept_valid = keep != nullptr;
  kept_x = x;
  kept_ldx = ldx;
  kept_m = m;
  if (p.zb > 0) {
    if (plane_threads == 512) split_pass<1, F, 512>(pairs, nullptr, w, d_rho_part.p);
    else split_pass<1, F, 256>(pairs, nullptr, w, d_rho_part.p);
  } else if (plane_threads == 512)
    hipLaunchKernelGGL((k_plane<1, F, 512>), dim3(p.n0, groups), dim3(512), plane_lds(p, e), stream, sticks, p.n0,
                       p.n1, p.n2, p.Spad, p.Z, p.chunk, d_stick_yz.p, d_zc.p, d_roots.p, p.ry, p.rz, nullptr,
                       pairs, w, 1.0 / sys->volume, d_rho_part.p, keep);
  else
    hipLaunchKernelGGL((k_plane<1, F, 256>), dim3(p.n0, groups), dim3(256), plane_lds(p, e), stream, sticks, p.n0,
                       p.n1, p.n2, p.Spad, p.Z, p.chunk, d_stick_yz.p, d_zc.p, d_roots.p, p.ry, p.rz, nullptr,
                       pairs, w, 1.0 / sys->volume, d_rho_part.p, keep);
  hipLaunchKernelGGL(k_sum_groups, dim3(unsigned((N + 255) / 256)), dim3(256), 0, stream, d_rho_part.p, N,
                     groups, rho, accumulate ? 1 : 0);
}

// Keep (on = true) the real-space pair fields of every later band density,
// for apply_kept: pairs_cap x N values of the op's precision.
void GammaOp::enable_kept(bool on) {
  if (on && (kb || plan.zb > 0))
    fail(BD_ERR_INVALID_PARAMETER, "kept fields: Gamma ops on the one-workgroup plane pass only");
  if (on && !d_kept.p) d_kept.alloc((size_t(slots(capacity)) * N * elem_bytes() + 15) / 16);
  kept_on = on;
  kept_valid = false;
}

// y += dV X, from the fields kept by the last band density, which must have
// been of these same columns (x, ldx, m): the r -> G half of the local apply
// (apply_local_change, distributed_real_gamma.rs:1370-1438).
template <typename F>
void GammaOp::kept_apply(const double* x, int64_t ldx, int m, const double* dv, double* y, int64_t ldy) {
  const StickPlan& p = plan;
  const int pairs = slots(m);
  const size_t e = sizeof(c2<F>);
  c2<F>* sticks = reinterpret_cast<c2<F>*>(d_sticks.p);
  const F* dvf;
  if constexpr (std::is_same_v<F, float>) {
    hipLaunchKernelGGL(k_to_float, dim3(unsigned((N + 255) / 256)), dim3(256), 0, stream, dv, N, d_potential32.p);
    dvf = d_potential32.p;
  } else {
    dvf = dv;
  }
  const c2<F>* kept = reinterpret_cast<const c2<F>*>(d_kept.p);
  if (plane_threads == 512)
    hipLaunchKernelGGL((k_plane_kept<F, 512>), dim3(pairs, p.n0), dim3(512), plane_lds(p, e), stream, sticks, p.n0,
                       p.n1, p.n2, p.Spad, p.Z, p.chunk, d_stick_yz.p, d_zc.p, d_roots.p, p.ry, p.rz, dvf, kept);
  else
    hipLaunchKernelGGL((k_plane_kept<F, 256>), dim3(pairs, p.n0), dim3(256), plane_lds(p, e), stream, sticks, p.n0,
                       p.n1, p.n2, p.Spad, p.Z, p.chunk, d_stick_yz.p, d_zc.p, d_roots.p, p.ry, p.rz, dvf, kept);
  hipLaunchKernelGGL(k_stick_forward<F>, dim3(p.ntiles, pairs), dim3(kThreads), stick_lds(p, e), stream, x, ldx, y,
                     ldy, m, p.n0, p.T, p.Spad, d_tile_off.p, d_entries.p, d_roots.p, p.rx, sticks, d_kinetic.p,
                     1.0 / double(N), 1);
  if constexpr (std::is_same_v<F, float>) refresh_potential32();  // the float copy is V(r) again
}

void GammaOp::apply_kept(const double* x, int64_t ldx, int m, const double* dv, double* y, int64_t ldy) {
  if (!kept_on) fail(BD_ERR_INVALID_PARAMETER, "kept fields are off (bd_gamma_op_keep_fields)");
  if (!kept_valid || x != kept_x || ldx != kept_ldx || m != kept_m)
    fail(BD_ERR_INVALID_PARAMETER, "no kept fields of these columns (band density of the same X first)");
  if (ldy < n) fail(BD_ERR_INVALID_PARAMETER, "leading dimension smaller than rows");
  if (m == 0) return;
  if (fft32()) kept_apply<float>(x, ldx, m, dv, y, ldy);
  else kept_apply<double>(x, ldx, m, dv, y, ldy);
  BD_HIP(hipGetLastError());
}

void GammaOp::projections(const double* x, int64_t ldx, int m) {
  gemm.tn(stream, nproj, m, n, 1.0, d_proj.p, n, x, ldx, 0.0, d_q.p, nproj, d_work.p);
}

void GammaOp::apply(const double* x, int64_t ldx, int m, double* y, int64_t ldy) {
  if (m < 0 || m > capacity) fail(BD_ERR_INVALID_PARAMETER, "column count exceeds capacity");
  if (ldx < n || ldy < n) fail(BD_ERR_INVALID_PARAMETER, "leading dimension smaller than rows");
  if (!has_potential) fail(BD_ERR_INVALID_PARAMETER, "no potential installed");
  if (kb && !complex_aligned(x, ldx)) fail(BD_ERR_INVALID_PARAMETER, "k-point X must be 16-byte aligned, even ldx");
  if (kb && !complex_aligned(y, ldy)) fail(BD_ERR_INVALID_PARAMETER, "k-point Y must be 16-byte aligned, even ldy");
  if (m == 0) return;
  const int pairs = slots(m);
  if (fused) {
    if (profiling) BD_HIP(hipEventRecord(ev[0], stream));
    fused_apply(x, ldx, m, y, ldy);
    if (profiling) {
      BD_HIP(hipEventRecord(ev[5], stream));
      BD_HIP(hipEventSynchronize(ev[5]));
      float ms = 0.f;
      BD_HIP(hipEventElapsedTime(&ms, ev[0], ev[5]));
      stage_ms[1] = ms;
      stage_total_ms[1] += ms;
      applies += 1.0;
      pairs_total += pairs;
      columns_total += m;
    }
    BD_HIP(hipGetLastError());
    return;
  }
  if (fft32()) local_apply<float>(x, ldx, m, y, ldy);
  else local_apply<double>(x, ldx, m, y, ldy);
  if (nproj > 0) {
    projections(x, ldx, m);
    const int64_t tot = nproj * m;
    hipLaunchKernelGGL(k_coupling, dim3(unsigned((tot + 255) / 256)), dim3(256), 0, stream, d_q.p,
                       d_w.p, nproj, m, nc, d_blockd.p);
    if (profiling) BD_HIP(hipEventRecord(ev[4], stream));
    gemm.nn(stream, n, m, nproj, 1.0, d_proj.p, n, d_w.p, nproj, 1.0, y, ldy);
  } else if (profiling) {
    BD_HIP(hipEventRecord(ev[4], stream));
  }
  if (profiling) {
    BD_HIP(hipEventRecord(ev[5], stream));
    BD_HIP(hipEventSynchronize(ev[5]));
    for (int i = 0; i < 5; ++i) {
      float ms = 0.f;
      BD_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      stage_ms[i] = ms;
      stage_total_ms[i] += ms;
    }
    applies += 1.0;
    pairs_total += pairs;
    columns_total += m;
  }
  BD_HIP(hipGetLastError());
}

void GammaOp::band_density(const double* x, int64_t ldx, int m, const double* w, double* rho, bool accumulate) {
  if (m <= 0 || m > capacity) fail(BD_ERR_INVALID_PARAMETER, "column count");
  if (ldx < n) fail(BD_ERR_INVALID_PARAMETER, "leading dimension smaller than rows");
  if (kb && !complex_aligned(x, ldx)) fail(BD_ERR_INVALID_PARAMETER, "k-point X must be 16-byte aligned, even ldx");
  if (fft32()) density<float>(x, ldx, m, w, rho, accumulate);
  else density<double>(x, ldx, m, w, rho, accumulate);
  BD_HIP(hipGetLastError());
}

}  // namespace bd

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct bd_system;
namespace bd {
const System& system_of(const bd_system* s);
}

extern "C" bd_status bd_gamma_op_create(const bd_system* sys, int32_t capacity, int32_t precision,
                                        bd_gamma_op** out) {
  BD_GUARD({
    if (!sys || !out) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    *out = new bd_gamma_op(bd::system_of(sys), capacity, precision);
  });
}

extern "C" bd_status bd_k_op_create(const bd_system* sys, const double* k_frac, int32_t capacity,
                                    int32_t precision, bd_gamma_op** out) {
  BD_GUARD({
    if (!sys || !k_frac || !out) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    const bd::System& s = bd::system_of(sys);
    auto kb = std::make_unique<bd::KBasis>(s, k_frac);
    *out = new bd_gamma_op(s, std::move(kb), capacity, precision);
  });
}

extern "C" bd_status bd_k_op_info(const bd_gamma_op* op, int64_t* nk, double* k_frac3, double* k_cart3) {
  BD_GUARD({
    if (!op || !nk) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    if (!op->kbasis) bd::fail(BD_ERR_INVALID_PARAMETER, "not a k-point operator");
    const auto& kb = *op->kbasis;
    *nk = kb.nk;
    for (int i = 0; i < 3; ++i) {
      if (k_frac3) k_frac3[i] = kb.k_frac[i];
      if (k_cart3) k_cart3[i] = kb.k_cart[i];
    }
  });
}

extern "C" bd_status bd_k_op_get(const bd_gamma_op* op, const char* name, void* out, int64_t count) {
  BD_GUARD({
    if (!op || !name || !out) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    if (!op->kbasis) bd::fail(BD_ERR_INVALID_PARAMETER, "not a k-point operator");
    const auto& kb = *op->kbasis;
    std::string k(name);
    auto put = [&](const void* src, size_t elems, size_t bytes_each) {
      if (int64_t(elems) != count) bd::fail(BD_ERR_INVALID_PARAMETER, "count mismatch for " + k);
      std::memcpy(out, src, elems * bytes_each);
    };
    if (k == "miller") put(kb.miller.data(), kb.miller.size() * 3, 4);
    else if (k == "g2") put(kb.g2.data(), kb.g2.size(), 8);
    else if (k == "q2") put(kb.q2.data(), kb.q2.size(), 8);
    else if (k == "fft_index") put(kb.flat.data(), kb.flat.size(), 8);
    else if (k == "kinetic") put(kb.kinetic2.data(), kb.kinetic2.size(), 8);
    else if (k == "diagonal_static") put(kb.diagonal2.data(), kb.diagonal2.size(), 8);
    else if (k == "projectors") put(kb.proj2.data(), kb.proj2.size(), 8);
    else bd::fail(BD_ERR_INVALID_PARAMETER, "unknown array " + k);
  });
}

extern "C" void bd_gamma_op_destroy(bd_gamma_op* op) { delete op; }

extern "C" void* bd_gamma_op_stream(bd_gamma_op* op) { return op ? op->op.stream : nullptr; }

extern "C" bd_status bd_gamma_op_set_potential_dev(bd_gamma_op* op, const double* v, uint64_t epoch) {
  BD_GUARD({
    if (!op || !v) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    op->op.set_potential_dev(v, epoch);
  });
}

extern "C" bd_status bd_gamma_op_set_potential(bd_gamma_op* op, const double* v, uint64_t epoch) {
  BD_GUARD({
    if (!op || !v) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& o = op->op;
    if (o.has_potential && epoch <= o.epoch) bd::fail(BD_ERR_INVALID_PARAMETER, "potential epoch must increase");
    BD_HIP(hipMemcpyAsync(o.d_potential.p, v, size_t(o.N) * sizeof(double), hipMemcpyHostToDevice, o.stream));
    o.refresh_potential32();
    BD_HIP(hipStreamSynchronize(o.stream));
    o.epoch = epoch;
    o.has_potential = true;
  });
}

extern "C" bd_status bd_gamma_op_apply_dev(bd_gamma_op* op, const double* x, int64_t ldx, int32_t m,
                                           double* y, int64_t ldy) {
  BD_GUARD({
    if (!op || (m > 0 && (!x || !y))) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    op->op.apply(x, ldx, m, y, ldy);
  });
}

extern "C" bd_status bd_gamma_op_apply(bd_gamma_op* op, const double* x, int64_t ldx, int32_t m,
                                       double* y, int64_t ldy) {
  BD_GUARD({
    if (!op || (m > 0 && (!x || !y))) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& o = op->op;
    if (m < 0 || m > o.capacity) bd::fail(BD_ERR_INVALID_PARAMETER, "column count exceeds capacity");
    if (ldx < o.n || ldy < o.n) bd::fail(BD_ERR_INVALID_PARAMETER, "leading dimension smaller than rows");
    if (m == 0) return BD_OK;
    for (int j = 0; j < m; ++j)
      BD_HIP(hipMemcpyAsync(o.d_x.p + size_t(j) * o.n, x + size_t(j) * ldx, size_t(o.n) * 8,
                            hipMemcpyHostToDevice, o.stream));
    o.apply(o.d_x.p, o.n, m, o.d_y.p, o.n);
    for (int j = 0; j < m; ++j)
      BD_HIP(hipMemcpyAsync(y + size_t(j) * ldy, o.d_y.p + size_t(j) * o.n, size_t(o.n) * 8,
                            hipMemcpyDeviceToHost, o.stream));
    BD_HIP(hipStreamSynchronize(o.stream));
  });
}

extern "C" bd_status bd_gamma_op_band_density_dev(bd_gamma_op* op, const double* x, int64_t ldx, int32_t m,
                                                  const double* weights, double* rho) {
  BD_GUARD({
    if (!op || !x || !weights || !rho) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& o = op->op;
    if (m <= 0 || m > o.capacity) bd::fail(BD_ERR_INVALID_PARAMETER, "column count");
    // Per FFT slot (w_a, w_b): a column pair at Gamma; at k one complex
    // column, whose |psi|^2 = Re^2 + Im^2 takes its weight twice.
    std::vector<double> hw(size_t(2 * o.slots(m)), 0.0);
    if (o.kb)
      for (int j = 0; j < m; ++j) hw[2 * j] = hw[2 * j + 1] = weights[j];
    else
      std::copy(weights, weights + m, hw.begin());
    o.d_weights.upload(hw.data(), hw.size(), o.stream);
    o.band_density(x, ldx, m, o.d_weights.p, rho);
    BD_HIP(hipStreamSynchronize(o.stream));
  });
}

namespace bd {
namespace {
// rho[i] = sum over k in order of part[k N + i].
__global__ void k_sum_k(const double* __restrict__ part, int64_t N, int count, double* __restrict__ rho) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double s = part[i];
  for (int k = 1; k < count; ++k) s += part[int64_t(k) * N + i];
  rho[i] = s;
}

// Device scratch of the k-point batch calls, grown on demand (one per process;
// the calls are synchronous and not meant to overlap).
struct KScratch {
  std::mutex mu;
  double* p = nullptr;
  size_t n = 0;
  double* get(size_t count) {
    if (count > n) {
      if (p) BD_HIP(hipFree(p));
      p = nullptr;
      n = 0;
      BD_HIP(hipMalloc(&p, count * sizeof(double)));
      n = count;
    }
    return p;
  }
};
KScratch& k_scratch() {
  static KScratch s;
  return s;
}

// Order every op's stream after `caller` (null: no caller stream to wait for).
void k_after(bd_gamma_op* const* ops, int count, hipStream_t caller, hipEvent_t ev) {
  if (!caller) return;
  BD_HIP(hipEventRecord(ev, caller));
  for (int k = 0; k < count; ++k) BD_HIP(hipStreamWaitEvent(ops[k]->op.stream, ev, 0));
}
}  // namespace
}  // namespace bd

// The band densities of every k point of a map summed into one grid, and the
// potential of every k point set from one grid: one call each instead of one
// per k point. The k points' densities run side by side on their ops'
// streams into separate grids, summed in k order at the end (the bits of the
// per-k densities added in k order).
extern "C" bd_status bd_k_band_density_dev(bd_gamma_op* const* ops, int32_t count, const double* const* x,
                                           const int64_t* ldx, const int32_t* m, const double* const* weights,
                                           double* rho, void* stream) {
  BD_GUARD({
    if (!ops || !x || !ldx || !m || !weights || !rho || count <= 0) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    int mmax = 0;
    for (int k = 0; k < count; ++k) {
      if (!ops[k] || !x[k] || !weights[k]) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
      const auto& o = ops[k]->op;
      if (!o.kb) bd::fail(BD_ERR_INVALID_PARAMETER, "k-point ops only");
      if (o.N != ops[0]->op.N) bd::fail(BD_ERR_INVALID_PARAMETER, "k-point ops must share the grid");
      if (m[k] <= 0 || m[k] > o.capacity) bd::fail(BD_ERR_INVALID_PARAMETER, "column count");
      mmax = std::max(mmax, int(m[k]));
    }
    const int64_t N = ops[0]->op.N;
    bd::KScratch& ks = bd::k_scratch();
    std::lock_guard<std::mutex> lock(ks.mu);
    // Per k: the 2 m weights (each complex column's |psi|^2 = Re^2 + Im^2 takes
    // its weight twice), then count density grids.
    const size_t wstride = size_t(2 * mmax);
    double* dev = ks.get(wstride * count + size_t(N) * count);
    double* wdev = dev;
    double* part = dev + wstride * count;
    std::vector<double> hw(wstride * count, 0.0);
    for (int k = 0; k < count; ++k)
      for (int j = 0; j < m[k]; ++j) hw[k * wstride + 2 * j] = hw[k * wstride + 2 * j + 1] = weights[k][j];
    hipStream_t s0 = ops[0]->op.stream;
    hipEvent_t ev = nullptr;
    BD_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    try {
      bd::k_after(ops, count, static_cast<hipStream_t>(stream), ev);
      BD_HIP(hipMemcpyAsync(wdev, hw.data(), hw.size() * sizeof(double), hipMemcpyHostToDevice, s0));
      BD_HIP(hipEventRecord(ev, s0));
      for (int k = 0; k < count; ++k) {
        auto& o = ops[k]->op;
        if (k > 0) BD_HIP(hipStreamWaitEvent(o.stream, ev, 0));
        o.band_density(x[k], ldx[k], m[k], wdev + k * wstride, part + int64_t(k) * N, false);
      }
      for (int k = 1; k < count; ++k) {
        BD_HIP(hipEventRecord(ev, ops[k]->op.stream));
        BD_HIP(hipStreamWaitEvent(s0, ev, 0));
      }
      hipLaunchKernelGGL(bd::k_sum_k, dim3(unsigned((N + 255) / 256)), dim3(256), 0, s0, part, N, count, rho);
      BD_HIP(hipGetLastError());
      BD_HIP(hipStreamSynchronize(s0));
    } catch (...) {
      (void)hipEventDestroy(ev);
      throw;
    }
    BD_HIP(hipEventDestroy(ev));
  });
}

extern "C" bd_status bd_k_set_potential_dev(bd_gamma_op* const* ops, int32_t count, const double* v,
                                            uint64_t epoch, void* stream) {
  BD_GUARD({
    if (!ops || !v || count <= 0) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    for (int k = 0; k < count; ++k)
      if (!ops[k]) bd::fail(BD_ERR_INVALID_PARAMETER, "null op");
    hipEvent_t ev = nullptr;
    BD_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    try {
      bd::k_after(ops, count, static_cast<hipStream_t>(stream), ev);
      for (int k = 0; k < count; ++k) ops[k]->op.set_potential_dev(v, epoch);
      // The caller may overwrite v once every copy has been made.
      if (stream)
        for (int k = 0; k < count; ++k) {
          BD_HIP(hipEventRecord(ev, ops[k]->op.stream));
          BD_HIP(hipStreamWaitEvent(static_cast<hipStream_t>(stream), ev, 0));
        }
    } catch (...) {
      (void)hipEventDestroy(ev);
      throw;
    }
    BD_HIP(hipEventDestroy(ev));
  });
}

extern "C" bd_status bd_gamma_op_keep_fields(bd_gamma_op* op, int32_t on) {
  BD_GUARD({
    if (!op) bd::fail(BD_ERR_INVALID_PARAMETER, "null op");
    op->op.enable_kept(on != 0);
  });
}

extern "C" bd_status bd_gamma_op_apply_kept_dev(bd_gamma_op* op, const double* x, int64_t ldx, int32_t m,
                                                const double* delta_v_dev, double* y, int64_t ldy) {
  BD_GUARD({
    if (!op || (m > 0 && (!x || !y || !delta_v_dev))) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    op->op.apply_kept(x, ldx, m, delta_v_dev, y, ldy);
  });
}

extern "C" bd_status bd_gamma_op_nonlocal_energies_dev(bd_gamma_op* op, const double* x, int64_t ldx,
                                                       int32_t m, double* out) {
  BD_GUARD({
    if (!op || !x || !out) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& o = op->op;
    if (m <= 0 || m > o.capacity) bd::fail(BD_ERR_INVALID_PARAMETER, "column count");
    std::fill(out, out + m, 0.0);
    if (o.nproj == 0) return BD_OK;
    o.projections(x, ldx, m);
    std::vector<double> q(size_t(o.nproj) * m);
    BD_HIP(hipMemcpyAsync(q.data(), o.d_q.p, q.size() * 8, hipMemcpyDeviceToHost, o.stream));
    BD_HIP(hipStreamSynchronize(o.stream));
    const auto& s = *o.sys;
    // At k, rows [0, p) of q hold Re(P^H x) and rows [p, 2p) Im(P^H x); D is
    // real, so <x|P D P^H|x> = Re^T D Re + Im^T D Im.
    const int halves = o.kb ? 2 : 1;
    for (int b = 0; b < m; ++b) {
      double sum = 0.0;
      for (int h = 0; h < halves; ++h)
        for (auto& blk : s.blocks) {
          const double* qb = &q[size_t(b) * o.nproj + size_t(h) * s.nproj + blk.start];
          for (int j = 0; j < blk.count; ++j) {
            double v = 0.0;
            for (int k = 0; k < blk.count; ++k) v += s.block_d[j * blk.count + k] * qb[k];
            sum += qb[j] * v;
          }
        }
      out[b] = sum;
    }
  });
}

extern "C" bd_status bd_gamma_op_plan_info(bd_gamma_op* op, int64_t* out8) {
  BD_GUARD({
    if (!op || !out8) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    const auto& p = op->op.plan;
    const int64_t v[8] = {p.n0, p.n1, p.n2, p.S, p.Spad, p.Z, p.T, p.ntiles};
    for (int i = 0; i < 8; ++i) out8[i] = v[i];
  });
}

extern "C" bd_status bd_gamma_op_set_profiling(bd_gamma_op* op, int32_t on) {
  BD_GUARD({
    if (!op) bd::fail(BD_ERR_INVALID_PARAMETER, "null op");
    auto& o = op->op;
    o.profiling = on != 0;
    for (double& t : o.stage_total_ms) t = 0.0;
    o.applies = o.pairs_total = o.columns_total = 0.0;
  });
}

extern "C" bd_status bd_gamma_op_stage_times(bd_gamma_op* op, double* out8) {
  BD_GUARD({
    if (!op || !out8) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    const auto& o = op->op;
    for (int i = 0; i < 5; ++i) out8[i] = o.stage_total_ms[i];
    out8[5] = o.applies;
    out8[6] = o.pairs_total;
    out8[7] = o.columns_total;
  });
}
