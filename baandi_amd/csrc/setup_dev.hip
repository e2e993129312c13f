// The atom loops of the set-up on the GPU (System with on_device): the
// projector columns' structure-factor phases and their ORC packing, the
// Davidson diagonal, and the ionic potential's structure factors. The host
// keeps the atom-independent parts (radial transforms per |G| shell, real
// Y_lm); these loops are O(rows x atoms) and dominate the host set-up of a
// large cell (the Si512 cell: 192k rows and 1.5M density G against 512 atoms).
// Reference: crates/dft-physics/src/hamiltonian.rs:885-1035 (projector table,
// D blocks), 211-340 (V_ion on the density sphere); ORC rows
// distributed_real_gamma.rs:363-402; diagonal crates/dft-engine/src/refstart.rs:605-650.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "common.hpp"
#include "fft_device.hpp"
#include "op.hpp"
#include "system.hpp"

namespace bd {
namespace {
constexpr double kSqrt2 = 1.41421356237309504880168872420969808;

// Projector rows in ORC order: thread (representative r, atom a) writes the
// nc columns a*nc + c. Row 0 (G = 0) holds Re f, representative r >= 1 rows
// 2r-1 and 2r hold sqrt2 (Re f, Im f) at +G; f = table_c(G) e^{-i G.tau_a}.
__global__ void k_dev_projectors(int64_t n, int64_t nrep, int nc, int natoms, const int64_t* __restrict__ plus,
                                 const double* __restrict__ cart, const double* __restrict__ atoms,
                                 const c2<double>* __restrict__ table, double* __restrict__ orc) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (r >= nrep) return;
  const int64_t g = plus[r];
  const double ang = cart[3 * g] * atoms[3 * a] + cart[3 * g + 1] * atoms[3 * a + 1] + cart[3 * g + 2] * atoms[3 * a + 2];
  double sn, cs;
  sincos(-ang, &sn, &cs);
  const int64_t nproj = int64_t(nc) * natoms;
  for (int c = 0; c < nc; ++c) {
    const c2<double> t = table[int64_t(c) * n + g];
    const double fr = t.x * cs - t.y * sn, fi = t.x * sn + t.y * cs;
    double* col = orc + (int64_t(a) * nc + c) * n;
    if (r == 0) {
      col[0] = fr;
    } else {
      col[2 * r - 1] = kSqrt2 * fr;
      col[2 * r] = kSqrt2 * fi;
    }
  }
  (void)nproj;
}

// Davidson diagonal per representative: T + sum_a sum_ij D_ij Re(b_i b_j*)
// with b_c = table_c(G) e^{-i G.tau_a}, shared by the representative's rows.
__global__ void k_dev_diagonal(int64_t n, int64_t nrep, int nc, int natoms, const int64_t* __restrict__ plus,
                               const double* __restrict__ cart, const double* __restrict__ g2,
                               const double* __restrict__ atoms, const c2<double>* __restrict__ table,
                               const double* __restrict__ block_d, double* __restrict__ diag) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  const int64_t g = plus[r];
  double v = 0.5 * g2[g];
  for (int a = 0; a < natoms; ++a) {
    const double ang = cart[3 * g] * atoms[3 * a] + cart[3 * g + 1] * atoms[3 * a + 1] + cart[3 * g + 2] * atoms[3 * a + 2];
    double sn, cs;
    sincos(-ang, &sn, &cs);
    c2<double> b[16];
    for (int c = 0; c < nc; ++c) {
      const c2<double> t = table[int64_t(c) * n + g];
      b[c] = {t.x * cs - t.y * sn, t.x * sn + t.y * cs};
    }
    for (int j = 0; j < nc; ++j)
      for (int i = 0; i < nc; ++i) v += block_d[i + nc * j] * (b[i].x * b[j].x + b[i].y * b[j].y);
  }
  if (r == 0) {
    diag[0] = v;
  } else {
    diag[2 * r - 1] = v;
    diag[2 * r] = v;
  }
}

// V_ion(G) = V_loc(|G|) sum_a e^{-i G.tau_a} per density-sphere G.
__global__ void k_dev_ionic(int64_t nd, int natoms, const double* __restrict__ dcart, const double* __restrict__ atoms,
                            const double* __restrict__ vloc, double* __restrict__ vre, double* __restrict__ vim) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= nd) return;
  double sr = 0.0, si = 0.0;
  for (int a = 0; a < natoms; ++a) {
    const double ang =
        dcart[3 * g] * atoms[3 * a] + dcart[3 * g + 1] * atoms[3 * a + 1] + dcart[3 * g + 2] * atoms[3 * a + 2];
    double sn, cs;
    sincos(ang, &sn, &cs);
    sr += cs;
    si -= sn;
  }
  vre[g] = vloc[g] * sr;
  vim[g] = vloc[g] * si;
}

template <typename T>
struct Buf {
  T* p = nullptr;
  explicit Buf(size_t count) { BD_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T))); }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  void up(const T* src, size_t count) { BD_HIP(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice)); }
  void down(T* dst, size_t count) const { BD_HIP(hipMemcpy(dst, p, count * sizeof(T), hipMemcpyDeviceToHost)); }
};

std::vector<double> flat3(const std::vector<std::array<double, 3>>& v) {
  std::vector<double> out(3 * v.size());
  for (size_t i = 0; i < v.size(); ++i)
    for (int k = 0; k < 3; ++k) out[3 * i + k] = v[i][k];
  return out;
}
}  // namespace

void device_projectors(System& s, const std::vector<std::complex<double>>& table) {
  const int nc = s.channel_count;
  const int na = int(s.atoms.size());
  if (nc > 16) fail(BD_ERR_INVALID_PARAMETER, "device set-up holds at most 16 channels per atom");
  const std::vector<double> cart = flat3(s.cart), atoms = flat3(s.atoms);
  Buf<int64_t> d_plus(s.nrep);
  Buf<double> d_cart(cart.size()), d_atoms(atoms.size()), d_g2(s.n), d_bd(s.block_d.size());
  Buf<c2<double>> d_table(table.size());
  d_plus.up(s.plus.data(), s.nrep);
  d_cart.up(cart.data(), cart.size());
  d_atoms.up(atoms.data(), atoms.size());
  d_g2.up(s.g2.data(), s.n);
  d_bd.up(s.block_d.data(), s.block_d.size());
  d_table.up(reinterpret_cast<const c2<double>*>(table.data()), table.size());
  Buf<double> d_orc(size_t(s.n) * s.nproj), d_diag(s.n);
  const unsigned blocks = unsigned((s.nrep + 255) / 256);
  hipLaunchKernelGGL(k_dev_projectors, dim3(blocks, na), dim3(256), 0, nullptr, s.n, s.nrep, nc, na, d_plus.p,
                     d_cart.p, d_atoms.p, d_table.p, d_orc.p);
  hipLaunchKernelGGL(k_dev_diagonal, dim3(blocks), dim3(256), 0, nullptr, s.n, s.nrep, nc, na, d_plus.p, d_cart.p,
                     d_g2.p, d_atoms.p, d_table.p, d_bd.p, d_diag.p);
  BD_HIP(hipGetLastError());
  s.projectors_orc.resize(size_t(s.n) * s.nproj);
  s.diagonal_static.resize(s.n);
  d_orc.down(s.projectors_orc.data(), s.projectors_orc.size());
  d_diag.down(s.diagonal_static.data(), s.n);
}

void device_ionic(System& s, const std::vector<double>& vloc) {
  const int64_t nd = int64_t(vloc.size());
  const std::vector<double> dcart = flat3(s.dens_cart), atoms = flat3(s.atoms);
  Buf<double> d_dcart(dcart.size()), d_atoms(atoms.size()), d_v(nd), d_re(nd), d_im(nd);
  d_dcart.up(dcart.data(), dcart.size());
  d_atoms.up(atoms.data(), atoms.size());
  d_v.up(vloc.data(), nd);
  hipLaunchKernelGGL(k_dev_ionic, dim3(unsigned((nd + 255) / 256)), dim3(256), 0, nullptr, nd, int(s.atoms.size()),
                     d_dcart.p, d_atoms.p, d_v.p, d_re.p, d_im.p);
  BD_HIP(hipGetLastError());
  s.vion_re.resize(nd);
  s.vion_im.resize(nd);
  d_re.down(s.vion_re.data(), nd);
  d_im.down(s.vion_im.data(), nd);
}

}  // namespace bd
