// Gamma Hamiltonian operator: FFT plan, device tables, scratch, launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "fft_device.hpp"
#include "gemm.hpp"
#include "system.hpp"

namespace bd {

#define BD_HIP(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      ::bd::fail(BD_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

// Device buffer owned by the op (allocated at construction only).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    n = count;
    if (count) BD_HIP(hipMalloc(&p, count * sizeof(T)));
  }
  void upload(const T* src, size_t count, hipStream_t s) {
    BD_HIP(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Stick decomposition of the sphere (sticks along grid axis 0).
struct StickPlan {
  int n0, n1, n2;
  int T;        // sticks per tile
  int ntiles;
  int S;        // occupied sticks
  int Spad;     // ntiles * T
  int Z;        // occupied z planes
  int chunk;    // y rows per z-pass chunk in the plane kernel
  int zb = 0;   // split plane pass: z lines per y-transform batch (0: fused k_plane)
  RadixPlan rx, ry, rz;
  std::vector<int> tile_off;       // ntiles+1 offsets into entries
  std::vector<int4> entries;       // Gamma: (rep, plus slot*n0+x, minus slot*n0+x, 0)
                                   // k: (basis index, slot*n0+x, -1, 0)
  std::vector<int2> stick_yz;      // per padded stick: (i1, zc), (-1,-1) pad
  std::vector<int> z_of_zc;        // Z
  std::vector<int> fpos;           // k only: stick * n0 + x per basis coefficient
};

// One k point's columns in the fused k-point kernels (a device array entry).
struct KWork {
  const double* x;
  double* y;
  int64_t ldx, ldy;
  int m, first;           // columns; index of the first work item
  int nk, S, Z;           // coefficients, occupied sticks, occupied z planes
  const int* fpos;        // nk: stick * n0 + x, the coefficient's LDS slot
  const int2* yz;         // S: (y, zc) of each stick
  const int* z_of_zc;     // Z
  const double* kinetic;  // 2nk
  const void* potential;  // N values of F
  const double* proj2;    // 2nk x np2, column-major (ld 2nk)
  const double* blockd;   // nc x nc
  int np2, nc;
};

// Launch shape of the fused kernel: threads per workgroup, sticks per x-FFT
// batch, y rows per z-FFT batch and y-z planes per pass (the per-line
// arithmetic does not depend on them, so every valid shape gives the same bits).
struct FusedShape {
  int threads = 256;
  int xlines = 0;
  int zrows = 0;
  int xplanes = 1;  // y-z planes transformed together
};

struct DavidsonWork;
void destroy_davidson_work(DavidsonWork* w);
struct ComplexWork;
void destroy_complex_work(ComplexWork* w);

// The H·X operator of one Gamma cell (ORC rows), or of one k point of it
// (kb != nullptr: rows are interleaved (Re, Im) of complex coefficients, one
// complex FFT per column instead of one per column pair).
struct GammaOp {
  const System* sys;
  const KBasis* kb = nullptr;
  DavidsonWork* dav = nullptr;  // Davidson workspace cached per (n, bands, cap)
  ComplexWork* cdav = nullptr;  // complex Davidson workspace (k-point ops), likewise
  int64_t n, nrep, N;
  int dims[3];
  int capacity;
  int precision;
  // The FFT stages (sticks, planes, the potential) run in FP32 in every mode
  // but FP64; the GEMMs are BF16x3 only in BD_MIXED.
  bool fft32() const { return precision != BD_FP64; }
  int64_t nproj;
  hipStream_t stream = nullptr;
  StickPlan plan;
  uint64_t epoch = 0;
  bool has_potential = false;
  // device tables
  DevBuf<int> d_tile_off, d_zc;
  DevBuf<int4> d_entries;
  DevBuf<int2> d_stick_yz;
  DevBuf<c2<double>> d_roots;  // [n0 | n1 | n2]
  DevBuf<double> d_kinetic, d_potential, d_proj, d_blockd;
  DevBuf<float> d_potential32;  // mixed tier: V(r) in float
  // scratch
  DevBuf<c2<double>> d_sticks;  // pairs_cap * n0 * Spad
  DevBuf<double> d_q, d_w, d_work, d_x, d_y, d_weights, d_rho_part;
  Gemm gemm;  // FP64, or BF16x3 in the mixed tier
  // k point: the fused one-workgroup-per-column H·X (whole 3D FFT in LDS)
  // when its LDS fits; the three-stage stick/plane path otherwise.
  bool fused = false;
  FusedShape fshape;
  int plane_threads = 256;  // k_plane workgroup size (256 or 512; tuned with plan.chunk)
  // Split plane pass (plan.zb > 0): plane buffer [pair][x][Z][n1] for
  // plane_batch FFT slots at a time.
  DevBuf<c2<double>> d_planes;
  int plane_batch = 0;
  size_t elem_bytes() const;
  bool plane_shape_valid(int threads, int chunk) const;
  void enable_split(int zb);
  template <int MODE, typename F, int NT>
  void split_pass(int pairs, const F* pot, const double* w, double* rho_part);
  // Warm start (KeptHx): real-space pair fields of the last band density,
  // [pair][N] of the op's precision, and the columns they belong to.
  DevBuf<c2<double>> d_kept;
  bool kept_on = false, kept_valid = false;
  const double* kept_x = nullptr;
  int64_t kept_ldx = 0;
  int kept_m = 0;
  void enable_kept(bool on);
  void apply_kept(const double* x, int64_t ldx, int m, const double* dv, double* y, int64_t ldy);
  template <typename F>
  void kept_apply(const double* x, int64_t ldx, int m, const double* dv, double* y, int64_t ldy);
  DevBuf<int> d_fpos;
  KWork kwork(const double* x, int64_t ldx, int m, double* y, int64_t ldy, int first) const;
  size_t fused_lds(const FusedShape& f) const;
  bool fused_valid(const FusedShape& f) const;
  void fused_apply(const double* x, int64_t ldx, int m, double* y, int64_t ldy);
  int nc = 0;  // channels per atom block
  int nblocks = 0;
  // profiling
  bool profiling = false;
  hipEvent_t ev[6] = {};
  double stage_ms[5] = {};
  double stage_total_ms[5] = {};
  double applies = 0.0, pairs_total = 0.0, columns_total = 0.0;

  GammaOp(const System& s, const KBasis* k, int capacity, int precision);
  // FFT slots that m columns occupy: column pairs at Gamma, columns at k.
  int slots(int m) const { return kb ? m : (m + 1) / 2; }
  ~GammaOp();
  void set_potential_dev(const double* v, uint64_t epoch);
  void apply(const double* x, int64_t ldx, int m, double* y, int64_t ldy);
  void band_density(const double* x, int64_t ldx, int m, const double* w, double* rho, bool accumulate = false);
  void projections(const double* x, int64_t ldx, int m);  // d_q = P^T X
  void refresh_potential32();
  template <typename F>
  void local_apply(const double* x, int64_t ldx, int m, double* y, int64_t ldy);
  size_t plane_lds_bytes(int chunk) const;
  template <typename F>
  void density(const double* x, int64_t ldx, int m, const double* w, double* rho, bool accumulate = false);
};

// Launch the fused H·X for `total` work items over `count` k points (works on
// the device); all share the grid, radix plans, roots and precision of `op`.
// works_dev == nullptr: the single entry `one` is used (passed by value).
void fused_launch(const GammaOp& op, const FusedShape& f, size_t lds, const KWork* works_dev, const KWork& one,
                  int count, int total, hipStream_t s);
bool fft_fits_threads(int nlines, const RadixPlan& rp, int threads);

StickPlan build_stick_plan(const System& s);
StickPlan build_k_stick_plan(const System& s, const KBasis& kb);
RadixPlan radix_plan(int n, bool big);

}  // namespace bd

// The C ABI handle (include/baandi_amd.h) owns one operator, and at k its basis.
struct bd_gamma_op {
  std::unique_ptr<bd::KBasis> kbasis;
  bd::GammaOp op;
  bd_gamma_op(const bd::System& s, int cap, int prec) : op(s, nullptr, cap, prec) {}
  bd_gamma_op(const bd::System& s, std::unique_ptr<bd::KBasis> k, int cap, int prec)
      : kbasis(std::move(k)), op(s, kbasis.get(), cap, prec) {}
};
