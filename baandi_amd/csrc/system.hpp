// Set-up products of one Gamma cell (host memory).
#pragma once
#include <array>
#include <cstdint>
#include <complex>
#include <vector>

#include "../../include/baandi_amd.h"

namespace bd {

struct System {
  // Cell: lattice columns [cartesian][vector], inverse, B = 2 pi A^{-T}.
  double lat[3][3];
  double inv[3][3];
  double B[3][3];
  double volume = 0.0;
  std::vector<std::array<double, 3>> atoms;  // Cartesian, Bohr
  double electrons = 0.0;
  double ecut = 0.0;  // Ha
  // Synthetic pseudopotential on the log mesh (Ha).
  double z = 0.0, rc = 1.0;
  std::vector<double> r, rab, local_ha;
  struct Proj {
    int l;
    std::vector<double> rbeta;
  };
  std::vector<Proj> projs;
  std::vector<double> dij_ha;
  // Basis (sorted by |G|^2 then Miller) and grid.
  int64_t n = 0, nrep = 0;
  int dims[3] = {0, 0, 0};
  std::vector<std::array<int32_t, 3>> miller;
  std::vector<std::array<double, 3>> cart;
  std::vector<double> g2;
  // Half sphere: representative -> basis index of +G / -G, and flat FFT index.
  std::vector<int64_t> plus, minus, plus_flat, minus_flat;
  std::vector<double> kinetic_orc;
  // Nonlocal: ORC projector rows (n x nproj col-major), block-diagonal D.
  int64_t nproj = 0;
  int channel_count = 0;
  std::vector<double> projectors_orc, coupling, block_d;
  struct Block {
    int64_t start;
    int count;
  };
  std::vector<Block> blocks;
  std::vector<double> diagonal_static;
  // Density sphere (4 Ecut) and V_ion(G) on it.
  std::vector<std::array<int32_t, 3>> dens_miller;
  std::vector<std::array<double, 3>> dens_cart;
  std::vector<double> dens_g2;
  std::vector<int64_t> dens_flat;
  std::vector<double> vion_re, vion_im;
  double ewald = 0.0;  // Ha

  // on_device: the atom loops of the set-up (projector phases and ORC rows,
  // the Davidson diagonal, the ionic structure factors) run on the GPU.
  bool device = false;
  System(const bd_system_desc& d, bool on_device);

 private:
  void build_pseudo(const bd_system_desc& d);
  void build_half_sphere();
  void build_projectors();
  void build_ionic();
  void build_ewald();
};

// Device halves of the set-up (setup_dev.hip): from the atom-independent
// projector table (nc x n complex, channel-major), the ORC projector rows and
// the Davidson diagonal; from V_loc(|G|) per density-sphere G, the ionic
// potential V_ion(G) = V_loc sum_a e^{-i G.tau_a}.
void device_projectors(System& s, const std::vector<std::complex<double>>& table);
void device_ionic(System& s, const std::vector<double>& vloc);

// The Bloch basis of one k point of a System (k != 0 in general): the same
// cell, grid, pseudopotential and D blocks; complex coefficients stored as
// interleaved (Re, Im) real rows, 2 nk of them.
struct KBasis {
  double k_frac[3];  // wrapped into [-1/2, 1/2)
  double k_cart[3];  // Bohr^-1
  int64_t nk = 0;
  std::vector<std::array<int32_t, 3>> miller;  // |G|^2 then Miller order
  std::vector<double> g2, q2;                  // |G|^2 and |k+G|^2
  std::vector<int64_t> flat;                   // FFT index of each G
  std::vector<double> kinetic2;                // |k+G|^2/2 on both rows of a coefficient
  // Projectors as a real 2nk x 2p block (col-major): column c holds
  // (Re P_c, Im P_c) interleaved, column p + c holds (-Im P_c, Re P_c), so
  // that P2^T x gives (Re, Im) of P^H x and P2 w applies P to a complex w.
  std::vector<double> proj2;
  std::vector<double> diagonal2;  // static Davidson diagonal on both rows
  KBasis(const System& s, const double k_fractional[3]);
};

}  // namespace bd
