// The SCF field step on the GPU: the density-sphere (4 Ecut) coefficients
// and their grid transforms, the Hartree and LDA-PZ potentials, the
// half-sphere Hartree metric and the difference-history mixing updates.
// Reference: crates/dft-engine/src/refscf/map.rs:41-198 (potential of the
// input density), refscf/metric.rs (dr2), refscf/broyden.rs:106-113, 317-445
// (Anderson update, Thomas-Fermi screening), crates/dft-physics/src/xc.rs:48-73
// (Slater exchange + Perdew-Zunger 81 correlation), hartree.rs.
//
// Grid transforms are full 3D FFTs of the density grid: one launch per axis,
// a workgroup transforming a batch of lines in LDS (the lines of a batch are
// adjacent in memory along axis 2, so every load and store is a run of the
// batch). Reductions are two-stage in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "fft_device.hpp"
#include "op.hpp"

namespace bd {
bool fft_fits_threads(int nlines, const RadixPlan& rp, int threads);

namespace {
constexpr int kFT = 256;         // threads per workgroup
constexpr int kRedGroups = 256;  // partial sums of the reductions

template <typename T>
__device__ __forceinline__ void stage(c2<T>* dst, const c2<double>* src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = {T(src[i].x), T(src[i].y)};
}

// Lines along `axis` of the n0 x n1 x n2 grid (row-major, axis 2 fastest),
// B per workgroup. Axis 2: B consecutive rows; axes 0, 1: B consecutive
// axis-2 positions of one line family, so each line element is a run of B.
template <bool INV>
__global__ __launch_bounds__(kFT) void k_grid_fft(c2<double>* __restrict__ g, int n0, int n1, int n2, int axis, int B,
                                                  RadixPlan rp, const c2<double>* __restrict__ roots) {
  extern __shared__ char smem[];
  const int n = rp.n;
  c2<double>* L = reinterpret_cast<c2<double>*>(smem);  // [B][n]
  c2<double>* rt = L + B * n;
  stage(rt, roots, n);
  int64_t base, step;  // element (b, i) at g[base + b * bstride + i * step]
  int64_t bstride;
  int nb;
  if (axis == 2) {
    const int64_t rows = int64_t(n0) * n1;
    const int64_t r0 = int64_t(blockIdx.x) * B;
    nb = int(min<int64_t>(B, rows - r0));
    base = r0 * n2;
    bstride = n2;
    step = 1;
  } else {
    const int per = (n2 + B - 1) / B;  // batches along axis 2
    const int q = blockIdx.x / per, z0 = (blockIdx.x - q * per) * B;
    nb = min(B, n2 - z0);
    bstride = 1;
    if (axis == 1) {  // q = i0
      base = int64_t(q) * n1 * n2 + z0;
      step = n2;
    } else {  // q = i1
      base = int64_t(q) * n2 + z0;
      step = int64_t(n1) * n2;
    }
  }
  for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
    const int i = e / B, b = e - i * B;  // b fastest: runs of B in memory
    L[b * n + i] = b < nb ? g[base + b * bstride + i * step] : c2<double>{0.0, 0.0};
  }
  __syncthreads();
  lds_fft<INV, double, kFT>(L, B, n, 1, rp, rt);
  for (int e = threadIdx.x; e < nb * n; e += blockDim.x) {
    const int i = e / nb, b = e - i * nb;
    g[base + b * bstride + i * step] = L[b * n + i];
  }
}

// grid = 0, then grid[flat[i]] = c[i] (c complex, interleaved).
__global__ void k_zero_c(c2<double>* __restrict__ g, int64_t N) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < N) g[i] = {0.0, 0.0};
}
__global__ void k_scatter(const c2<double>* __restrict__ c, const int64_t* __restrict__ flat, int64_t nd,
                          c2<double>* __restrict__ g) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < nd) g[flat[i]] = c[i];
}
// c[i] = grid[flat[i]] * scale.
__global__ void k_gather(const c2<double>* __restrict__ g, const int64_t* __restrict__ flat, int64_t nd, double scale,
                         c2<double>* __restrict__ c) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < nd) c[i] = cscale(g[flat[i]], scale);
}
// complex grid = real field.
__global__ void k_real_to_complex(const double* __restrict__ r, int64_t N, c2<double>* __restrict__ g) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < N) g[i] = {r[i], 0.0};
}

// LDA-PZ per point (xc.rs:48-73): eps_xc and v_xc, zero at or below 1e-10.
__device__ __forceinline__ void lda_pz(double n, double& eps, double& vxc) {
  if (!(n > 1e-10)) {
    eps = 0.0;
    vxc = 0.0;
    return;
  }
  const double cd = cbrt(n);
  const double rs = cbrt(3.0 / (4.0 * M_PI)) / cd;
  const double ex = -0.75 * cbrt(3.0 / M_PI) * cd;
  double ec, dec;
  if (rs >= 1.0) {
    const double sr = sqrt(rs);
    const double den = 1.0 + 1.0529 * sr + 0.3334 * rs;
    ec = -0.1423 / den;
    dec = 0.1423 * (1.0529 / (2.0 * sr) + 0.3334) / (den * den);
  } else {
    const double lg = log(rs);
    ec = 0.0311 * lg - 0.048 + 0.0020 * rs * lg - 0.0116 * rs;
    dec = 0.0311 / rs + 0.0020 * (lg + 1.0) - 0.0116;
  }
  eps = ex + ec;
  vxc = 4.0 * ex / 3.0 + ec - rs * dec / 3.0;
}

// Workgroup sum to thread 0 (fixed tree).
__device__ __forceinline__ double wg_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < int(blockDim.x >> 6); ++w) s += red[w];
  return s;
}

// The grid g holds rho_in + i V_H (both real fields). V = vion + V_H +
// v_xc(|rho_in|); partial sums of V per group.
__global__ __launch_bounds__(kFT) void k_potential(const c2<double>* __restrict__ g, const double* __restrict__ vion,
                                                   int64_t N, double* __restrict__ v, double* __restrict__ part) {
  __shared__ double red[kFT / 64];
  double acc = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += int64_t(gridDim.x) * blockDim.x) {
    double eps, vxc;
    lda_pz(fabs(g[i].x), eps, vxc);
    const double val = vion[i] + g[i].y + vxc;
    v[i] = val;
    acc += val;
  }
  const double s = wg_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[j] = sum over groups of part[j * groups + g], in group order.
__global__ void k_finish(const double* __restrict__ part, int groups, int count, double* __restrict__ out) {
  const int j = threadIdx.x;
  if (j >= count) return;
  double s = 0.0;
  for (int gi = 0; gi < groups; ++gi) s += part[int64_t(j) * groups + gi];
  out[j] = s;
}

// y = coefficients of the output density (y[zero] = electrons / volume), f =
// y - x, and the partial metric sums of |f|^2.
__global__ __launch_bounds__(kFT) void k_residual(const c2<double>* __restrict__ y_in, const c2<double>* __restrict__ x,
                                                  const double* __restrict__ metric, int64_t nd, int64_t zero,
                                                  double rho0, c2<double>* __restrict__ f, double* __restrict__ part) {
  __shared__ double red[kFT / 64];
  double acc = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nd; i += int64_t(gridDim.x) * blockDim.x) {
    c2<double> y = y_in[i];
    if (i == zero) y = {rho0, 0.0};
    const c2<double> d = {y.x - x[i].x, y.y - x[i].y};
    f[i] = d;
    acc += metric[i] * (d.x * d.x + d.y * d.y);
  }
  const double s = wg_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Metric products of p history rows R (p x nd complex) with each other and
// with f: part[(i p + j) groups + g] and part[(p p + i) groups + g].
__global__ __launch_bounds__(kFT) void k_gram(const c2<double>* __restrict__ R, int p, const c2<double>* __restrict__ f,
                                              const double* __restrict__ metric, int64_t nd,
                                              double* __restrict__ part) {
  __shared__ double red[kFT / 64];
  const int groups = gridDim.x;
  for (int i = 0; i < p; ++i)
    for (int j = 0; j <= p; ++j) {
      if (j < p && j < i) continue;  // symmetric: the upper triangle
      const c2<double>* a = R + int64_t(i) * nd;
      const c2<double>* b = j < p ? R + int64_t(j) * nd : f;
      double acc = 0.0;
      for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < nd; t += int64_t(groups) * blockDim.x)
        acc += metric[t] * (a[t].x * b[t].x + a[t].y * b[t].y);
      const double s = wg_sum(acc, red);
      if (threadIdx.x == 0) {
        const int slot = j < p ? i * p + j : p * p + i;
        part[int64_t(slot) * groups + blockIdx.x] = s;
        if (j < p && j != i) part[int64_t(j * p + i) * groups + blockIdx.x] = s;
      }
    }
}

// out = (x - sum_i gamma_i dX_i) + beta (f - sum_i gamma_i dF_i) * screen
// (screen = 1 where null).
__global__ void k_mix(const c2<double>* __restrict__ x, const c2<double>* __restrict__ f,
                      const c2<double>* __restrict__ dX, const c2<double>* __restrict__ dF,
                      const double* __restrict__ gamma, int p, double beta, const double* __restrict__ screen,
                      int64_t nd, c2<double>* __restrict__ out) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nd) return;
  c2<double> xx = x[t], ff = f[t];
  for (int i = 0; i < p; ++i) {
    const double gm = gamma[i];
    const c2<double> a = dX[int64_t(i) * nd + t], b = dF[int64_t(i) * nd + t];
    xx = {xx.x - gm * a.x, xx.y - gm * a.y};
    ff = {ff.x - gm * b.x, ff.y - gm * b.y};
  }
  const double sc = screen ? screen[t] : 1.0;
  out[t] = {xx.x + beta * ff.x * sc, xx.y + beta * ff.y * sc};
}

// Energy sums of a density grid rho with its Hartree potential (the real
// part of vh): per group, rho V_H, |rho| eps_xc(|rho|), rho V_ion.
__global__ __launch_bounds__(kFT) void k_energies(const double* __restrict__ rho, const c2<double>* __restrict__ vh,
                                                  const double* __restrict__ vion, int64_t N,
                                                  double* __restrict__ part) {
  __shared__ double red[kFT / 64];
  double h = 0.0, x = 0.0, l = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += int64_t(gridDim.x) * blockDim.x) {
    const double r = rho[i];
    double eps, vxc;
    lda_pz(fabs(r), eps, vxc);
    h += r * vh[i].x;
    x += fabs(r) * eps;
    l += r * vion[i];
  }
  const int groups = gridDim.x;
  double s = wg_sum(h, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  s = wg_sum(x, red);
  if (threadIdx.x == 0) part[groups + blockIdx.x] = s;
  s = wg_sum(l, red);
  if (threadIdx.x == 0) part[2 * groups + blockIdx.x] = s;
}

// c[i] *= w[i] (complex by real), into out.
__global__ void k_weight(const c2<double>* __restrict__ c, const double* __restrict__ w, int64_t nd,
                         c2<double>* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < nd) out[i] = cscale(c[i], w[i]);
}
// out = a + i b for complex a, b (so that a real-field pair shares one transform).
__global__ void k_pack_i(const c2<double>* __restrict__ a, const c2<double>* __restrict__ b, int64_t nd,
                         c2<double>* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < nd) out[i] = {a[i].x - b[i].y, a[i].y + b[i].x};
}
}  // namespace

// Device tables and scratch of the field step of one system.
struct Field {
  const System* sys;
  int dims[3];
  int64_t N, nd, zero = -1;
  hipStream_t stream = nullptr;
  RadixPlan rp[3];
  int batch[3];
  DevBuf<c2<double>> d_roots[3], d_grid, d_c, d_c2;
  DevBuf<int64_t> d_flat;
  DevBuf<double> d_metric, d_coulomb, d_screen, d_part, d_out;
  double volume, electrons;

  explicit Field(const System& s) : sys(&s) {
    for (int a = 0; a < 3; ++a) dims[a] = s.dims[a];
    N = int64_t(dims[0]) * dims[1] * dims[2];
    nd = int64_t(s.dens_flat.size());
    volume = s.volume;
    electrons = s.electrons;
    BD_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (int a = 0; a < 3; ++a) {
      rp[a] = radix_plan(dims[a], true);
      int b = 32;
      auto ok = [&](int bb) {
        return fft_fits_threads(bb, rp[a], kFT) && size_t(bb * dims[a] + dims[a]) * 16 <= 64 * 1024;
      };
      while (b > 1 && !ok(b)) b /= 2;
      if (!ok(b)) {
        rp[a] = radix_plan(dims[a], false);
        b = 32;
        while (b > 1 && !ok(b)) b /= 2;
      }
      if (!ok(b)) fail(BD_ERR_INVALID_PARAMETER, "density-grid FFT line does not fit a workgroup");
      batch[a] = b;
      std::vector<c2<double>> roots(dims[a]);
      for (int k = 0; k < dims[a]; ++k) {
        const long double ang = -2.0L * 3.14159265358979323846264338327950288L * k / dims[a];
        roots[k] = {double(cosl(ang)), double(sinl(ang))};
      }
      d_roots[a].alloc(dims[a]);
      d_roots[a].upload(roots.data(), dims[a], stream);
    }
    // Half-sphere Hartree metric 8 pi Omega / G^2 (Ry) on positive
    // representatives (refscf/metric.rs), Coulomb 4 pi / G^2, Thomas-Fermi
    // screening G^2 / (G^2 + a), a = (12/pi)^(2/3) / r_s (broyden.rs:106-113).
    std::vector<double> metric(nd, 0.0), coulomb(nd, 0.0), screen(nd, 0.0);
    const double rs = std::cbrt(3.0 * volume / (4.0 * M_PI * std::max(electrons, 1e-300)));
    const double a_tf = std::pow(12.0 / M_PI, 2.0 / 3.0) / rs;
    for (int64_t i = 0; i < nd; ++i) {
      const auto& m = s.dens_miller[i];
      const double g2 = s.dens_g2[i];
      const int first = m[0] != 0 ? m[0] : (m[1] != 0 ? m[1] : m[2]);
      const bool is_zero = m[0] == 0 && m[1] == 0 && m[2] == 0;
      if (is_zero) zero = i;
      if (first > 0) metric[i] = 8.0 * M_PI * volume / g2;
      if (!is_zero) coulomb[i] = 4.0 * M_PI / g2;
      screen[i] = g2 / (g2 + a_tf);
    }
    d_flat.alloc(nd);
    d_flat.upload(s.dens_flat.data(), nd, stream);
    d_metric.alloc(nd);
    d_metric.upload(metric.data(), nd, stream);
    d_coulomb.alloc(nd);
    d_coulomb.upload(coulomb.data(), nd, stream);
    d_screen.alloc(nd);
    d_screen.upload(screen.data(), nd, stream);
    d_grid.alloc(N);
    d_c.alloc(nd);
    d_c2.alloc(nd);
    d_part.alloc(size_t(kRedGroups) * 100);
    d_out.alloc(100);
    BD_HIP(hipStreamSynchronize(stream));
  }
  ~Field() {
    if (stream) (void)hipStreamDestroy(stream);
  }

  unsigned blocks(int64_t count) const { return unsigned((count + kFT - 1) / kFT); }

  // In place on d_grid: all three axes.
  void fft(bool inverse) {
    for (int a = 2; a >= 0; --a) {
      const int B = batch[a];
      const int64_t launches = a == 2 ? (int64_t(dims[0]) * dims[1] + B - 1) / B
                                      : int64_t(a == 1 ? dims[0] : dims[1]) * ((dims[2] + B - 1) / B);
      const size_t lds = size_t(B * dims[a] + dims[a]) * 16;
      if (inverse)
        hipLaunchKernelGGL(k_grid_fft<true>, dim3(unsigned(launches)), dim3(kFT), lds, stream, d_grid.p, dims[0],
                           dims[1], dims[2], a, B, rp[a], d_roots[a].p);
      else
        hipLaunchKernelGGL(k_grid_fft<false>, dim3(unsigned(launches)), dim3(kFT), lds, stream, d_grid.p, dims[0],
                           dims[1], dims[2], a, B, rp[a], d_roots[a].p);
    }
  }
  void to_grid(const c2<double>* c) {
    hipLaunchKernelGGL(k_zero_c, dim3(blocks(N)), dim3(kFT), 0, stream, d_grid.p, N);
    hipLaunchKernelGGL(k_scatter, dim3(blocks(nd)), dim3(kFT), 0, stream, c, d_flat.p, nd, d_grid.p);
    fft(true);
  }
  void from_real(const double* r, c2<double>* c) {
    hipLaunchKernelGGL(k_real_to_complex, dim3(blocks(N)), dim3(kFT), 0, stream, r, N, d_grid.p);
    fft(false);
    hipLaunchKernelGGL(k_gather, dim3(blocks(nd)), dim3(kFT), 0, stream, d_grid.p, d_flat.p, nd, 1.0 / double(N), c);
  }
  // Host values of `count` reductions whose kRedGroups partials are in d_part.
  void finish(int count, double* host) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(128), 0, stream, d_part.p, kRedGroups, count, d_out.p);
    BD_HIP(hipMemcpyAsync(host, d_out.p, size_t(count) * 8, hipMemcpyDeviceToHost, stream));
    BD_HIP(hipStreamSynchronize(stream));
  }
};

}  // namespace bd

struct bd_system;
namespace bd {
const System& system_of(const bd_system* s);
}

struct bd_field {
  bd::Field f;
  explicit bd_field(const bd::System& s) : f(s) {}
};

using bd::c2;

extern "C" bd_status bd_field_create(const bd_system* sys, bd_field** out) {
  BD_GUARD({
    if (!sys || !out) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    *out = new bd_field(bd::system_of(sys));
  });
}

extern "C" void bd_field_destroy(bd_field* f) { delete f; }

extern "C" void* bd_field_stream(bd_field* f) { return f ? f->f.stream : nullptr; }

extern "C" bd_status bd_field_to_grid_dev(bd_field* h, const double* coeff, double* grid) {
  BD_GUARD({
    if (!h || !coeff || !grid) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& f = h->f;
    f.to_grid(reinterpret_cast<const c2<double>*>(coeff));
    BD_HIP(hipMemcpyAsync(grid, f.d_grid.p, size_t(f.N) * 16, hipMemcpyDeviceToDevice, f.stream));
    BD_HIP(hipStreamSynchronize(f.stream));
  });
}

extern "C" bd_status bd_field_from_grid_dev(bd_field* h, const double* real_grid, double* coeff) {
  BD_GUARD({
    if (!h || !real_grid || !coeff) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& f = h->f;
    f.from_real(real_grid, reinterpret_cast<c2<double>*>(coeff));
    BD_HIP(hipStreamSynchronize(f.stream));
  });
}

extern "C" bd_status bd_field_potential_dev(bd_field* h, const double* x, const double* vion, double* v,
                                            double* vmean) {
  BD_GUARD({
    if (!h || !x || !vion || !v || !vmean) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& f = h->f;
    const auto* xc = reinterpret_cast<const c2<double>*>(x);
    // rho_in + i V_H in one transform: V_H(G) = 4 pi rho(G) / G^2 (hartree.rs).
    hipLaunchKernelGGL(bd::k_weight, dim3(f.blocks(f.nd)), dim3(bd::kFT), 0, f.stream, xc, f.d_coulomb.p, f.nd,
                       f.d_c.p);
    hipLaunchKernelGGL(bd::k_pack_i, dim3(f.blocks(f.nd)), dim3(bd::kFT), 0, f.stream, xc, f.d_c.p, f.nd, f.d_c2.p);
    f.to_grid(f.d_c2.p);
    hipLaunchKernelGGL(bd::k_potential, dim3(bd::kRedGroups), dim3(bd::kFT), 0, f.stream, f.d_grid.p, vion, f.N, v,
                       f.d_part.p);
    double s = 0.0;
    f.finish(1, &s);
    *vmean = s / double(f.N);
  });
}

extern "C" bd_status bd_field_residual_dev(bd_field* h, const double* rho_out, const double* x, double* fres,
                                           double* dr2) {
  BD_GUARD({
    if (!h || !rho_out || !x || !fres || !dr2) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& f = h->f;
    f.from_real(rho_out, f.d_c.p);
    hipLaunchKernelGGL(bd::k_residual, dim3(bd::kRedGroups), dim3(bd::kFT), 0, f.stream, f.d_c.p,
                       reinterpret_cast<const c2<double>*>(x), f.d_metric.p, f.nd, f.zero, f.electrons / f.volume,
                       reinterpret_cast<c2<double>*>(fres), f.d_part.p);
    f.finish(1, dr2);
  });
}

extern "C" bd_status bd_field_gram_dev(bd_field* h, int32_t p, const double* rows, const double* fres, double* a,
                                       double* c) {
  BD_GUARD({
    if (!h || !rows || !fres || !a || !c || p <= 0 || p > 9) bd::fail(BD_ERR_INVALID_PARAMETER, "gram arguments");
    auto& f = h->f;
    hipLaunchKernelGGL(bd::k_gram, dim3(bd::kRedGroups), dim3(bd::kFT), 0, f.stream,
                       reinterpret_cast<const c2<double>*>(rows), p, reinterpret_cast<const c2<double>*>(fres),
                       f.d_metric.p, f.nd, f.d_part.p);
    std::vector<double> out(size_t(p) * p + p);
    f.finish(p * p + p, out.data());
    std::copy(out.begin(), out.begin() + p * p, a);
    std::copy(out.begin() + p * p, out.end(), c);
  });
}

extern "C" bd_status bd_field_mix_dev(bd_field* h, int32_t p, const double* dX, const double* dF,
                                      const double* gamma_host, const double* x, const double* fres, double beta,
                                      int32_t screening, double* out) {
  BD_GUARD({
    if (!h || !x || !fres || !out || p < 0 || p > 9 || (p > 0 && (!dX || !dF || !gamma_host)))
      bd::fail(BD_ERR_INVALID_PARAMETER, "mix arguments");
    auto& f = h->f;
    double* g = f.d_out.p;  // gamma on the device (d_out is free between reductions)
    if (p > 0) BD_HIP(hipMemcpyAsync(g, gamma_host, size_t(p) * 8, hipMemcpyHostToDevice, f.stream));
    hipLaunchKernelGGL(bd::k_mix, dim3(f.blocks(f.nd)), dim3(bd::kFT), 0, f.stream,
                       reinterpret_cast<const c2<double>*>(x), reinterpret_cast<const c2<double>*>(fres),
                       reinterpret_cast<const c2<double>*>(dX), reinterpret_cast<const c2<double>*>(dF), g, p, beta,
                       screening ? f.d_screen.p : nullptr, f.nd, reinterpret_cast<c2<double>*>(out));
    BD_HIP(hipStreamSynchronize(f.stream));
  });
}

extern "C" bd_status bd_field_energies_dev(bd_field* h, const double* rho, const double* vion, double* out3) {
  BD_GUARD({
    if (!h || !rho || !vion || !out3) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    auto& f = h->f;
    // V_H(G) = 4 pi rho(G) / G^2 of the density's sphere coefficients, V_H(r)
    // on the grid; the sums use the grid density itself (map.rs:466-540).
    f.from_real(rho, f.d_c.p);
    hipLaunchKernelGGL(bd::k_weight, dim3(f.blocks(f.nd)), dim3(bd::kFT), 0, f.stream, f.d_c.p, f.d_coulomb.p, f.nd,
                       f.d_c2.p);
    f.to_grid(f.d_c2.p);
    hipLaunchKernelGGL(bd::k_energies, dim3(bd::kRedGroups), dim3(bd::kFT), 0, f.stream, rho, f.d_grid.p, vion, f.N,
                       f.d_part.p);
    double sums[3];
    f.finish(3, sums);
    const double dv = f.volume / double(f.N);
    out3[0] = 0.5 * dv * sums[0];  // Hartree
    out3[1] = dv * sums[1];        // XC
    out3[2] = dv * sums[2];        // local
  });
}
