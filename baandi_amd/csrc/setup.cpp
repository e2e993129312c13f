// Host set-up of one Gamma cell: basis, grid, ORC half sphere, synthetic
// pseudopotential, radial transforms, projector table, ionic potential and the
// Davidson diagonal. Run once per cell (crates/dft-engine/src/refstart.rs:536-650).
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "system.hpp"

namespace bd {

namespace {
constexpr double kTau = 6.283185307179586476925286766559;
constexpr double kPi = 3.141592653589793238462643383279;
constexpr double kCutoffTol = 1e-12;  // crates/dft-basis/src/lib.rs:12

// 3x3 inverse by cofactors (the order of Cell::from_matrix's inverse,
// crates/dft-geometry/src/lib.rs:58-84); exact on 0/1 matrices.
void inverse3(const double m[3][3], double out[3][3], double* det_out) {
  double m11 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  double m12 = m[1][0] * m[2][2] - m[2][0] * m[1][2];
  double m13 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
  double det = m[0][0] * m11 - m[0][1] * m12 + m[0][2] * m13;
  out[0][0] = m11 / det;
  out[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) / det;
  out[0][2] = (m[0][1] * m[1][2] - m[1][1] * m[0][2]) / det;
  out[1][0] = -m12 / det;
  out[1][1] = (m[0][0] * m[2][2] - m[2][0] * m[0][2]) / det;
  out[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) / det;
  out[2][0] = m13 / det;
  out[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) / det;
  out[2][2] = (m[0][0] * m[1][1] - m[1][0] * m[0][1]) / det;
  *det_out = det;
}

bool smooth235(int n) {
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

// Composite Simpson in the mesh index, 3/8 tail for even counts, Kahan sum
// over panels (crates/dft-upf/src/radial.rs:56-81). v already includes rab.
double simpson(const double* v, size_t n) {
  if (n < 2) return 0.0;
  if (n == 2) return (v[0] + v[1]) * 0.5;
  size_t last = (n % 2 == 0) ? n - 4 : n - 1;
  double sum = 0.0, corr = 0.0;
  auto add = [&](double x) {
    double y = x - corr;
    double t = sum + y;
    corr = (t - sum) - y;
    sum = t;
  };
  for (size_t i = 0; i < last; i += 2) add((v[i] + 4.0 * v[i + 1] + v[i + 2]) / 3.0);
  if (n % 2 == 0)
    add(3.0 * (v[last] + 3.0 * v[last + 1] + 3.0 * v[last + 2] + v[last + 3]) / 8.0);
  return sum;
}

// Spherical Bessel j_l, l <= 3 (radial.rs:90-117).
double sph_bessel(int l, double x) {
  if (x < 4.0) {
    double term = 1.0;
    for (int k = 1; k <= l; ++k) term *= x / double(2 * k + 1);
    double s = term;
    for (int k = 1; k < 80; ++k) {
      term *= -x * x / double(2 * k * (2 * l + 2 * k + 1));
      s += term;
      if (std::fabs(term) <= 2.220446049250313e-16 * std::fabs(s)) break;
    }
    return s;
  }
  double sn = std::sin(x), cs = std::cos(x);
  double prev = sn / x;
  if (l == 0) return prev;
  double cur = (prev - cs) / x;
  for (int k = 1; k < l; ++k) {
    double next = double(2 * k + 1) / x * cur - prev;
    prev = cur;
    cur = next;
  }
  return cur;
}

// Real normalised harmonics with Condon-Shortley and the extra (-1)^m
// (radial.rs:121-164). Direction must be nonzero.
double real_ylm(int l, int m, const double d[3]) {
  double norm = std::hypot(std::hypot(d[0], d[1]), d[2]);
  double z = std::clamp(d[2] / norm, -1.0, 1.0);
  double phi = std::atan2(d[1], d[0]);
  int a = std::abs(m);
  double p = 1.0;
  for (int k = 1; k <= a; ++k) p *= -double(2 * k - 1) * std::sqrt((1.0 - z) * (1.0 + z));
  if (l > a) {
    double prev = p;
    p *= double(2 * a + 1) * z;
    for (int k = a + 2; k <= l; ++k) {
      double next = (double(2 * k - 1) * z * p - double(k + a - 1) * prev) / double(k - a);
      prev = p;
      p = next;
    }
  }
  double ratio = 1.0;
  for (int k = l - a + 1; k <= l + a; ++k) ratio /= double(k);
  double nrm = std::sqrt(double(2 * l + 1) / (4.0 * kPi) * ratio);
  double ang = 1.0;
  if (m != 0) {
    double sign = (a % 2 == 0) ? 1.0 : -1.0;
    ang = std::sqrt(2.0) * sign * (m < 0 ? std::sin(a * phi) : std::cos(a * phi));
  }
  return nrm * p * ang;
}

struct Sphere {
  std::vector<std::array<int32_t, 3>> miller;
  std::vector<std::array<double, 3>> cart;
  std::vector<double> g2;
};

// PlaneWaveBasis::new (crates/dft-basis/src/lib.rs:54-73, 234-269) at Gamma.
Sphere enumerate_sphere(const System& s, double ecut) {
  double radius = std::sqrt(2.0 * ecut * (1.0 + kCutoffTol));
  int bounds[3];
  for (int i = 0; i < 3; ++i) {
    double len = std::hypot(std::hypot(s.lat[0][i], s.lat[1][i]), s.lat[2][i]);
    bounds[i] = int(std::ceil(len * radius / kTau) + 1.0) + 1;
  }
  struct Entry {
    std::array<int32_t, 3> m;
    std::array<double, 3> c;
    double g2;
  };
  std::vector<Entry> v;
  const double limit = ecut * (1.0 + kCutoffTol);
  for (int m0 = -bounds[0]; m0 <= bounds[0]; ++m0)
    for (int m1 = -bounds[1]; m1 <= bounds[1]; ++m1)
      for (int m2 = -bounds[2]; m2 <= bounds[2]; ++m2) {
        std::array<double, 3> c;
        for (int i = 0; i < 3; ++i)
          c[i] = ((s.B[i][0] * double(m0)) + s.B[i][1] * double(m1)) + s.B[i][2] * double(m2);
        double g2 = (c[0] * c[0] + c[1] * c[1]) + c[2] * c[2];
        if (0.5 * g2 <= limit) v.push_back({{m0, m1, m2}, c, g2});
      }
  std::sort(v.begin(), v.end(), [](const Entry& a, const Entry& b) {
    if (a.g2 != b.g2) return a.g2 < b.g2;
    return a.m < b.m;
  });
  Sphere out;
  for (auto& e : v) {
    out.miller.push_back(e.m);
    out.cart.push_back(e.c);
    out.g2.push_back(e.g2);
  }
  return out;
}

int64_t flat_of(const std::array<int32_t, 3>& m, const int dims[3]) {
  int64_t idx[3];
  for (int i = 0; i < 3; ++i) {
    int n = dims[i];
    if (m[i] < -(n / 2) || m[i] > (n - 1) / 2)
      fail(BD_ERR_INVALID_PARAMETER, "Miller index outside the canonical FFT frequencies");
    idx[i] = m[i] >= 0 ? m[i] : n + m[i];
  }
  return (idx[0] * dims[1] + idx[1]) * dims[2] + idx[2];
}
}  // namespace

System::System(const bd_system_desc& d, bool on_device) : device(on_device) {
  if (!(d.ecut_ha > 0.0) || d.natoms <= 0 || d.mesh_points < 4 || d.z_valence <= 0.0 || d.rc <= 0.0)
    fail(BD_ERR_INVALID_PARAMETER, "system descriptor: cutoff, atoms, mesh, charge or rc");
  // Cell: user rows are lattice vectors -> columns (crates/dft-geometry/src/lib.rs:46-49).
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) lat[i][j] = d.lattice_rows[j][i];
  double scale = 0.0;
  for (auto& r : lat)
    for (double x : r) scale = std::max(scale, std::fabs(x));
  if (!(scale > 0.0)) fail(BD_ERR_INVALID_PARAMETER, "singular lattice");
  double scaled[3][3], sinv[3][3], det;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) scaled[i][j] = lat[i][j] / scale;
  inverse3(scaled, sinv, &det);
  if (!(std::fabs(det) > 1e-12)) fail(BD_ERR_INVALID_PARAMETER, "near-singular lattice");
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) inv[i][j] = sinv[i][j] / scale;
  volume = ((std::fabs(det) * scale) * scale) * scale;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B[i][j] = kTau * inv[j][i];
  for (int a = 0; a < d.natoms; ++a) {
    std::array<double, 3> c{};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) c[i] += lat[i][j] * d.frac_positions[3 * a + j];
    atoms.push_back(c);
  }
  electrons = d.z_valence * d.natoms;
  ecut = d.ecut_ha;

  build_pseudo(d);
  // Basis and grid.
  Sphere basis = enumerate_sphere(*this, d.ecut_ha);
  Sphere dens = enumerate_sphere(*this, 4.0 * d.ecut_ha);
  if (basis.miller.empty() || basis.miller[0] != std::array<int32_t, 3>{0, 0, 0})
    fail(BD_ERR_INVALID_PARAMETER, "basis has no G=0");
  int req[3] = {1, 1, 1};
  for (auto& m : dens.miller)
    for (int i = 0; i < 3; ++i) req[i] = std::max(req[i], 2 * std::abs(m[i]) + 1);
  bool override_on = d.grid_override[0] || d.grid_override[1] || d.grid_override[2];
  for (int i = 0; i < 3; ++i) {
    if (override_on) {
      if (d.grid_override[i] < req[i]) fail(BD_ERR_INVALID_PARAMETER, "FFT override smaller than required");
      dims[i] = d.grid_override[i];
    } else {
      int n = req[i];
      while (!smooth235(n)) ++n;
      dims[i] = n;
    }
  }
  n = int64_t(basis.miller.size());
  miller = basis.miller;
  g2 = basis.g2;
  cart = basis.cart;
  dens_miller = dens.miller;
  dens_g2 = dens.g2;
  dens_cart = dens.cart;
  for (auto& m : dens_miller) dens_flat.push_back(flat_of(m, dims));

  build_half_sphere();
  build_projectors();
  build_ionic();
  build_ewald();
}

// Ewald energy of point ions Z in a neutralising background, erfc(eta r)
// splitting with eta = sqrt(pi)/Omega^(1/3) (crates/dft-physics/src/ewald.rs:4,380);
// the total does not depend on eta.
void System::build_ewald() {
  const double eta = std::sqrt(kPi) / std::cbrt(volume);
  const double tol = 1e-20;
  double rcut = 1.0, gcut = 1.0;
  while (std::erfc(eta * rcut) > tol) rcut *= 1.1;
  while (std::exp(-gcut * gcut / (4.0 * eta * eta)) > tol) gcut *= 1.1;
  int nr[3], ng[3];
  for (int i = 0; i < 3; ++i) {
    double blen = std::hypot(std::hypot(B[0][i], B[1][i]), B[2][i]);
    double alen = std::hypot(std::hypot(lat[0][i], lat[1][i]), lat[2][i]);
    nr[i] = int(std::ceil(rcut * blen / kTau)) + 1;
    ng[i] = int(std::ceil(gcut * alen / kTau)) + 1;
  }
  const size_t na = atoms.size();
  double real = 0.0;
  for (int a = -nr[0]; a <= nr[0]; ++a)
    for (int b = -nr[1]; b <= nr[1]; ++b)
      for (int c = -nr[2]; c <= nr[2]; ++c) {
        double L[3];
        for (int i = 0; i < 3; ++i) L[i] = lat[i][0] * a + lat[i][1] * b + lat[i][2] * c;
        for (size_t i = 0; i < na; ++i)
          for (size_t j = 0; j < na; ++j) {
            double d2 = 0.0;
            for (int k = 0; k < 3; ++k) {
              double dk = atoms[i][k] - atoms[j][k] + L[k];
              d2 += dk * dk;
            }
            if (d2 == 0.0) continue;
            double r = std::sqrt(d2);
            if (r < rcut) real += 0.5 * z * z * std::erfc(eta * r) / r;
          }
      }
  double recip = 0.0;
  for (int a = -ng[0]; a <= ng[0]; ++a)
    for (int b = -ng[1]; b <= ng[1]; ++b)
      for (int c = -ng[2]; c <= ng[2]; ++c) {
        if (a == 0 && b == 0 && c == 0) continue;
        double G[3];
        for (int i = 0; i < 3; ++i) G[i] = B[i][0] * a + B[i][1] * b + B[i][2] * c;
        double g2v = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
        if (g2v > gcut * gcut) continue;
        double sr = 0.0, si = 0.0;
        for (auto& t : atoms) {
          double ph = G[0] * t[0] + G[1] * t[1] + G[2] * t[2];
          sr += z * std::cos(ph);
          si += z * std::sin(ph);
        }
        recip += 2.0 * kPi / volume * (sr * sr + si * si) * std::exp(-g2v / (4.0 * eta * eta)) / g2v;
      }
  const double ztot = z * double(na);
  const double self_e = -eta / std::sqrt(kPi) * z * z * double(na);
  const double background = -kPi * ztot * ztot / (2.0 * volume * eta * eta);
  ewald = real + recip + self_e + background;
}

void System::build_pseudo(const bd_system_desc& d) {
  z = d.z_valence;
  rc = d.rc;
  const double end = 12.0, r0 = 0.02;
  int points = d.mesh_points;
  double dx = std::log(1.0 + end / r0) / double(points - 1);
  r.resize(points);
  rab.resize(points);
  local_ha.resize(points);
  for (int i = 0; i < points; ++i) {
    r[i] = r0 * std::expm1(double(i) * dx);
    rab[i] = dx * (r[i] + r0);
    double ry = r[i] == 0.0 ? -4.0 * z / (rc * std::sqrt(kPi)) : -2.0 * z * std::erf(r[i] / rc) / r[i];
    local_ha[i] = ry / 2.0;  // PP_LOCAL Ry -> Ha (crates/dft-upf/src/parser.rs:765)
  }
  for (int p = 0; p < d.nproj; ++p) {
    Proj pj;
    pj.l = d.proj_l[p];
    if (pj.l < 0 || pj.l > 3) fail(BD_ERR_INVALID_PARAMETER, "projector l");
    pj.rbeta.resize(points);
    for (int i = 0; i < points; ++i)
      pj.rbeta[i] = std::pow(r[i], pj.l + 1) * std::exp(-r[i] * r[i] / d.proj_width[p]);
    projs.push_back(std::move(pj));
  }
  dij_ha.assign(size_t(d.nproj) * d.nproj, 0.0);
  for (int i = 0; i < d.nproj * d.nproj; ++i) dij_ha[i] = d.dij_ry[i] / 2.0;  // parser.rs:735
  for (int i = 0; i < d.nproj; ++i)
    for (int j = 0; j < d.nproj; ++j)
      if (projs[i].l != projs[j].l && dij_ha[i * d.nproj + j] != 0.0)
        fail(BD_ERR_INVALID_PARAMETER, "D couples different angular momenta");
}

void System::build_half_sphere() {
  // Mirror via a dense lookup over the basis box.
  int lo[3], hi[3];
  for (int i = 0; i < 3; ++i) {
    lo[i] = hi[i] = 0;
    for (auto& m : miller) {
      lo[i] = std::min(lo[i], m[i]);
      hi[i] = std::max(hi[i], m[i]);
    }
  }
  int ext[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
  std::vector<int64_t> lookup(size_t(ext[0]) * ext[1] * ext[2], -1);
  auto key = [&](const std::array<int32_t, 3>& m) -> int64_t {
    for (int i = 0; i < 3; ++i)
      if (m[i] < lo[i] || m[i] > hi[i]) return -1;
    return (int64_t(m[0] - lo[0]) * ext[1] + (m[1] - lo[1])) * ext[2] + (m[2] - lo[2]);
  };
  for (int64_t i = 0; i < n; ++i) lookup[key(miller[i])] = i;
  plus.assign(1, 0);
  minus.assign(1, 0);
  for (int64_t i = 0; i < n; ++i) {
    std::array<int32_t, 3> neg{-miller[i][0], -miller[i][1], -miller[i][2]};
    int64_t k = key(neg);
    int64_t j = k < 0 ? -1 : lookup[k];
    if (j < 0) fail(BD_ERR_INVALID_PARAMETER, "basis is not inversion symmetric");
    if (miller[i] > std::array<int32_t, 3>{0, 0, 0}) {
      plus.push_back(i);
      minus.push_back(j);
    }
  }
  nrep = int64_t(plus.size());
  if (2 * nrep - 1 != n) fail(BD_ERR_INVALID_PARAMETER, "half sphere size");
  plus_flat.resize(nrep);
  minus_flat.resize(nrep);
  for (int64_t i = 0; i < nrep; ++i) {
    plus_flat[i] = flat_of(miller[plus[i]], dims);
    minus_flat[i] = flat_of(miller[minus[i]], dims);
  }
  kinetic_orc.resize(n);
  kinetic_orc[0] = 0.5 * g2[0];
  for (int64_t i = 1; i < nrep; ++i) kinetic_orc[2 * i - 1] = kinetic_orc[2 * i] = 0.5 * g2[plus[i]];
}

namespace {
using Cplx = std::complex<double>;

// Channels (projector, l, m), ordered by projector then m = -l..l.
std::vector<std::array<int, 3>> channel_list(const System& s) {
  std::vector<std::array<int, 3>> channels;
  for (int i = 0; i < int(s.projs.size()); ++i)
    for (int m = -s.projs[i].l; m <= s.projs[i].l; ++m) channels.push_back({i, s.projs[i].l, m});
  return channels;
}

// Complex projector columns at the momenta q (hamiltonian.rs:930-1000):
// (-i)^l Y_lm(q^) beta_i(|q|)/sqrt(Omega) e^{-i q.tau}. The atom-independent
// part, channel-major (nc x n): radial transforms cached per exact |q|^2 bit
// pattern.
std::vector<Cplx> projector_table(const System& s, const std::vector<std::array<double, 3>>& q,
                                  const std::vector<double>& q2) {
  const auto channels = channel_list(s);
  const int npj = int(s.projs.size());
  const int nc = int(channels.size());
  const int64_t n = int64_t(q2.size());
  std::unordered_map<uint64_t, int> slot;
  std::vector<double> radial;
  std::vector<int> gslot(n);
  std::vector<double> buf(s.r.size());
  for (int64_t g = 0; g < n; ++g) {
    uint64_t k;
    std::memcpy(&k, &q2[g], 8);
    auto it = slot.find(k);
    if (it == slot.end()) {
      int sl = int(slot.size());
      slot.emplace(k, sl);
      double qn = std::sqrt(q2[g]);
      for (int i = 0; i < npj; ++i) {
        const auto& p = s.projs[i];
        for (size_t t = 0; t < s.r.size(); ++t)
          buf[t] = s.r[t] * p.rbeta[t] * sph_bessel(p.l, qn * s.r[t]) * s.rab[t];
        radial.push_back(4.0 * kPi * simpson(buf.data(), s.r.size()));
      }
      gslot[g] = sl;
    } else {
      gslot[g] = it->second;
    }
  }
  std::vector<Cplx> table(size_t(nc) * n);
  const double sv = std::sqrt(s.volume);
  for (int64_t g = 0; g < n; ++g) {
    double dir[3] = {q[g][0], q[g][1], q[g][2]};
    if (q2[g] == 0.0) dir[0] = 0.0, dir[1] = 0.0, dir[2] = 1.0;
    for (int c = 0; c < nc; ++c) {
      int i = channels[c][0], l = channels[c][1], m = channels[c][2];
      Cplx il = l == 0 ? Cplx(1, 0) : l == 1 ? Cplx(0, -1) : l == 2 ? Cplx(-1, 0) : Cplx(0, 1);
      table[size_t(c) * n + g] = il * (real_ylm(l, m, dir) * radial[size_t(gslot[g]) * npj + i] / sv);
    }
  }
  return table;
}

// The table times each atom's structure-factor phase e^{-i q.tau}: the
// columns atom-major (hamiltonian.rs:930-1000).
std::vector<Cplx> projector_columns(const System& s, const std::vector<std::array<double, 3>>& q,
                                    const std::vector<double>& q2) {
  const int nc = int(channel_list(s).size());
  const int64_t n = int64_t(q2.size());
  const std::vector<Cplx> table = projector_table(s, q, q2);
  std::vector<Cplx> full(size_t(n) * nc * s.atoms.size());
  for (size_t a = 0; a < s.atoms.size(); ++a) {
    for (int64_t g = 0; g < n; ++g) {
      double ang = q[g][0] * s.atoms[a][0] + q[g][1] * s.atoms[a][1] + q[g][2] * s.atoms[a][2];
      Cplx ph = std::polar(1.0, -ang);
      for (int c = 0; c < nc; ++c) full[(a * nc + c) * n + g] = table[size_t(c) * n + g] * ph;
    }
  }
  return full;
}

// T + sum_ij D_ij Re(b_i b_j*) of row g (refstart.rs:605-650 at Gamma,
// distributed_complex.rs:29-54 at k).
double static_row(const System& s, const std::vector<Cplx>& full, int64_t n, int64_t g, double kinetic) {
  const int nc = s.channel_count;
  double v = kinetic;
  for (size_t a = 0; a < s.atoms.size(); ++a)
    for (int j = 0; j < nc; ++j)
      for (int i = 0; i < nc; ++i) {
        Cplx bi = full[(a * nc + i) * n + g], bj = full[(a * nc + j) * n + g];
        v += s.block_d[i + nc * j] * (bi * std::conj(bj)).real();
      }
  return v;
}
}  // namespace

void System::build_projectors() {
  const auto channels = channel_list(*this);
  const int nc = int(channels.size());
  const int npj = int(projs.size());
  nproj = int64_t(nc) * int64_t(atoms.size());
  // Block-diagonal D over equal (l, m) channels (hamiltonian.rs:988-995).
  std::vector<double> block(size_t(nc) * nc, 0.0);
  for (int a = 0; a < nc; ++a)
    for (int b = 0; b < nc; ++b)
      if (channels[a][1] == channels[b][1] && channels[a][2] == channels[b][2])
        block[a * nc + b] = dij_ha[channels[a][0] * npj + channels[b][0]];
  channel_count = nc;
  block_d = block;
  coupling.assign(size_t(nproj) * nproj, 0.0);
  for (size_t at = 0; at < atoms.size(); ++at) {
    blocks.push_back({int64_t(at * nc), nc});
    for (int a = 0; a < nc; ++a)
      for (int b = 0; b < nc; ++b)
        coupling[(at * nc + b) * nproj + at * nc + a] = block[a * nc + b];  // column-major
  }
  if (device) {
    // The atom loops (phases, ORC packing, the Davidson diagonal) on the GPU.
    device_projectors(*this, projector_table(*this, cart, g2));
    return;
  }
  std::vector<Cplx> full = projector_columns(*this, cart, g2);
  // Per-atom columns packed to ORC rows (distributed_real_gamma.rs:363-402).
  projectors_orc.assign(size_t(n) * nproj, 0.0);
  const double s2 = std::sqrt(2.0);
  for (int64_t col = 0; col < nproj; ++col) {
    const Cplx* f = &full[size_t(col) * n];
    double* o = &projectors_orc[size_t(col) * n];
    o[0] = f[0].real();
    for (int64_t i = 1; i < nrep; ++i) {
      o[2 * i - 1] = s2 * f[plus[i]].real();
      o[2 * i] = s2 * f[plus[i]].imag();
    }
  }
  // Davidson diagonal T + sum_ij D_ij Re(b_i b_j*) per representative
  // (refstart.rs:605-650), shared by its two ORC rows.
  diagonal_static.resize(n);
  diagonal_static[0] = static_row(*this, full, n, 0, 0.5 * g2[0]);
  for (int64_t i = 1; i < nrep; ++i)
    diagonal_static[2 * i - 1] = diagonal_static[2 * i] = static_row(*this, full, n, plus[i], 0.5 * g2[plus[i]]);
}

void System::build_ionic() {
  // V_loc(|G|) on the density sphere (radial.rs:264-287, hamiltonian.rs:256-283).
  size_t nl = 0;
  double limit = std::min(r.back(), 10.0);
  while (nl < r.size() && r[nl] <= limit) ++nl;
  std::vector<double> shortv(nl), buf(nl);
  for (size_t i = 0; i < nl; ++i) shortv[i] = r[i] * r[i] * local_ha[i] + z * r[i] * std::erf(r[i]);
  for (size_t i = 0; i < nl; ++i) buf[i] = shortv[i] * rab[i];
  const double alpha = 4.0 * kPi / volume * simpson(buf.data(), nl) + kPi * z / volume;
  std::unordered_map<uint64_t, double> cache;
  const size_t nd = dens_g2.size();
  vion_re.assign(nd, 0.0);
  vion_im.assign(nd, 0.0);
  std::vector<double> vg(nd);
  for (size_t g = 0; g < nd; ++g) {
    double v;
    if (dens_miller[g] == std::array<int32_t, 3>{0, 0, 0}) {
      v = alpha;
    } else {
      uint64_t k;
      std::memcpy(&k, &dens_g2[g], 8);
      auto it = cache.find(k);
      if (it != cache.end()) {
        v = it->second;
      } else {
        double q = std::sqrt(dens_g2[g]);
        for (size_t i = 0; i < nl; ++i) buf[i] = shortv[i] * sph_bessel(0, q * r[i]) * rab[i];
        v = 4.0 * kPi / volume * (simpson(buf.data(), nl) - z * std::exp(-q * q / 4.0) / (q * q));
        cache.emplace(k, v);
      }
    }
    vg[g] = v;
  }
  if (device) {
    device_ionic(*this, vg);  // structure factors on the GPU
    return;
  }
  for (size_t g = 0; g < nd; ++g) {
    const double v = vg[g];
    double sr = 0.0, si = 0.0;
    for (auto& t : atoms) {
      double ang = dens_cart[g][0] * t[0] + dens_cart[g][1] * t[1] + dens_cart[g][2] * t[2];
      sr += std::cos(ang);
      si -= std::sin(ang);
    }
    vion_re[g] = v * sr;
    vion_im[g] = v * si;
  }
}

KBasis::KBasis(const System& s, const double k_fractional[3]) {
  // k wrapped into [-1/2, 1/2) (crates/dft-geometry/src/kpoints.rs:21-32).
  for (int i = 0; i < 3; ++i) {
    double x = k_fractional[i];
    if (!std::isfinite(x)) fail(BD_ERR_INVALID_PARAMETER, "k point must be finite");
    k_frac[i] = (x >= -0.5 && x < 0.5) ? x : x - std::floor(x + 0.5);
  }
  for (int i = 0; i < 3; ++i)
    k_cart[i] = ((s.B[i][0] * k_frac[0]) + s.B[i][1] * k_frac[1]) + s.B[i][2] * k_frac[2];
  // PlaneWaveBasis::at_k (crates/dft-basis/src/lib.rs:234-269, inclusion
  // test :99-120): |k+G|^2/2 <= Ecut (1 + 1e-12), order by |G|^2 then Miller.
  const double ecut = s.ecut;
  double radius = std::sqrt(2.0 * ecut * (1.0 + kCutoffTol));
  int bounds[3];
  for (int i = 0; i < 3; ++i) {
    double len = std::hypot(std::hypot(s.lat[0][i], s.lat[1][i]), s.lat[2][i]);
    bounds[i] = int(std::ceil(len * radius / kTau) + 1.0) + 1;
  }
  struct Entry {
    std::array<int32_t, 3> m;
    std::array<double, 3> c;
    double g2;
  };
  std::vector<Entry> v;
  const double limit = ecut * (1.0 + kCutoffTol);
  for (int m0 = -bounds[0]; m0 <= bounds[0]; ++m0)
    for (int m1 = -bounds[1]; m1 <= bounds[1]; ++m1)
      for (int m2 = -bounds[2]; m2 <= bounds[2]; ++m2) {
        std::array<double, 3> c;
        for (int i = 0; i < 3; ++i)
          c[i] = ((s.B[i][0] * double(m0)) + s.B[i][1] * double(m1)) + s.B[i][2] * double(m2);
        double g2v = (c[0] * c[0] + c[1] * c[1]) + c[2] * c[2];
        double a = c[0] + k_cart[0], b = c[1] + k_cart[1], d = c[2] + k_cart[2];
        double q2v = (a * a + b * b) + d * d;
        if (0.5 * q2v <= limit) v.push_back({{m0, m1, m2}, c, g2v});
      }
  std::sort(v.begin(), v.end(), [](const Entry& a, const Entry& b) {
    if (a.g2 != b.g2) return a.g2 < b.g2;
    return a.m < b.m;
  });
  nk = int64_t(v.size());
  if (nk == 0) fail(BD_ERR_INVALID_PARAMETER, "empty basis at this k point");
  std::vector<std::array<double, 3>> q(nk);
  for (int64_t g = 0; g < nk; ++g) {
    miller.push_back(v[g].m);
    g2.push_back(v[g].g2);
    for (int i = 0; i < 3; ++i) q[g][i] = v[g].c[i] + k_cart[i];
    q2.push_back((q[g][0] * q[g][0] + q[g][1] * q[g][1]) + q[g][2] * q[g][2]);
    flat.push_back(flat_of(v[g].m, s.dims));
  }
  kinetic2.resize(2 * nk);
  for (int64_t g = 0; g < nk; ++g) kinetic2[2 * g] = kinetic2[2 * g + 1] = 0.5 * q2[g];
  const int64_t p = s.nproj;
  proj2.assign(size_t(2 * nk) * size_t(2 * p), 0.0);
  diagonal2.resize(2 * nk);
  if (p > 0) {
    std::vector<Cplx> full = projector_columns(s, q, q2);
    const int64_t rows = 2 * nk;
    for (int64_t c = 0; c < p; ++c) {
      const Cplx* f = &full[size_t(c) * nk];
      double* a = &proj2[size_t(c) * rows];
      double* b = &proj2[size_t(p + c) * rows];
      for (int64_t g = 0; g < nk; ++g) {
        a[2 * g] = f[g].real();
        a[2 * g + 1] = f[g].imag();
        b[2 * g] = -f[g].imag();
        b[2 * g + 1] = f[g].real();
      }
    }
    for (int64_t g = 0; g < nk; ++g) diagonal2[2 * g] = diagonal2[2 * g + 1] = static_row(s, full, nk, g, 0.5 * q2[g]);
  } else {
    for (int64_t g = 0; g < nk; ++g) diagonal2[2 * g] = diagonal2[2 * g + 1] = 0.5 * q2[g];
  }
}

}  // namespace bd

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct bd_system {
  bd::System s;
  bd_system(const bd_system_desc& d, bool device) : s(d, device) {}
};

namespace bd {
const System& system_of(const bd_system* s) { return s->s; }
}  // namespace bd

extern "C" bd_status bd_system_create(const bd_system_desc* desc, bd_system** out) {
  BD_GUARD({
    if (!desc || !out) bd::fail(BD_ERR_INVALID_PARAMETER, "null descriptor");
    *out = new bd_system(*desc, false);
  });
}

extern "C" bd_status bd_system_create_dev(const bd_system_desc* desc, bd_system** out) {
  BD_GUARD({
    if (!desc || !out) bd::fail(BD_ERR_INVALID_PARAMETER, "null descriptor");
    *out = new bd_system(*desc, true);
  });
}

extern "C" void bd_system_destroy(bd_system* sys) { delete sys; }

extern "C" bd_status bd_system_get_info(const bd_system* sys, bd_system_info* out) {
  BD_GUARD({
    if (!sys || !out) bd::fail(BD_ERR_INVALID_PARAMETER, "null system");
    const auto& s = sys->s;
    out->n = s.n;
    out->nrep = s.nrep;
    for (int i = 0; i < 3; ++i) out->dims[i] = s.dims[i];
    out->nproj = s.nproj;
    out->natoms = int64_t(s.atoms.size());
    out->ndensity = int64_t(s.dens_g2.size());
    out->volume = s.volume;
    out->electrons = s.electrons;
  });
}

extern "C" bd_status bd_system_get(const bd_system* sys, const char* name, void* out, int64_t count) {
  BD_GUARD({
    if (!sys || !name || !out) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    const auto& s = sys->s;
    std::string k(name);
    auto put = [&](const void* src, size_t elems, size_t bytes_each) {
      if (int64_t(elems) != count) bd::fail(BD_ERR_INVALID_PARAMETER, "count mismatch for " + k);
      std::memcpy(out, src, elems * bytes_each);
    };
    if (k == "miller") put(s.miller.data(), s.miller.size() * 3, 4);
    else if (k == "g2") put(s.g2.data(), s.g2.size(), 8);
    else if (k == "kinetic_orc") put(s.kinetic_orc.data(), s.kinetic_orc.size(), 8);
    else if (k == "diagonal_static") put(s.diagonal_static.data(), s.diagonal_static.size(), 8);
    else if (k == "projectors_orc") put(s.projectors_orc.data(), s.projectors_orc.size(), 8);
    else if (k == "coupling") put(s.coupling.data(), s.coupling.size(), 8);
    else if (k == "plus_index") put(s.plus_flat.data(), s.plus_flat.size(), 8);
    else if (k == "minus_index") put(s.minus_flat.data(), s.minus_flat.size(), 8);
    else if (k == "plus_basis") put(s.plus.data(), s.plus.size(), 8);
    else if (k == "minus_basis") put(s.minus.data(), s.minus.size(), 8);
    else if (k == "density_miller") put(s.dens_miller.data(), s.dens_miller.size() * 3, 4);
    else if (k == "density_g2") put(s.dens_g2.data(), s.dens_g2.size(), 8);
    else if (k == "density_index") put(s.dens_flat.data(), s.dens_flat.size(), 8);
    else if (k == "ionic_density_re") put(s.vion_re.data(), s.vion_re.size(), 8);
    else if (k == "ionic_density_im") put(s.vion_im.data(), s.vion_im.size(), 8);
    else if (k == "atoms_cart") put(s.atoms.data(), s.atoms.size() * 3, 8);
    else if (k == "ewald") put(&s.ewald, 1, 8);
    else bd::fail(BD_ERR_INVALID_PARAMETER, "unknown array " + k);
  });
}
