/*
 * baandi_amd — MI355X (gfx950) Davidson H·ψ path of baandi, C ABI.
 *
 * Plain pointers and sizes only. "_dev" entry points take device pointers
 * (hipMalloc'd, on the op's device); the others take host pointers and copy.
 * Matrices are column-major with the given leading dimension (RealMatrix
 * convention of crates/dft-numerics/src/real.rs). Energies are Hartree.
 * An op is not thread-safe (the reference operator is !Sync,
 * crates/dft-physics/src/hamiltonian.rs:1277); calls on one op serialise on
 * its HIP stream. Results are deterministic per precision mode.
 */
#ifndef BAANDI_AMD_H
#define BAANDI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: one per SolverError variant (crates/dft-solvers/src/lib.rs:41-57)
 * plus NotConverged and a device error. bd_last_error() returns the message. */
typedef enum {
  BD_OK = 0,
  BD_ERR_LINALG = 1,            /* SolverError::Linalg */
  BD_ERR_INVALID_PARAMETER = 2, /* SolverError::InvalidParameter */
  BD_ERR_APPLICATION = 3,       /* SolverError::Application */
  BD_NOT_CONVERGED = 4,         /* SolverError::NotConverged (exhausted iterations) */
  BD_ERR_DEVICE = 5             /* HIP runtime error or no gfx950 device */
} bd_status;

/* Arithmetic of the hot path. FP64: every stage in double (the reference
 * path). MIXED: the FFT stages of H.X (sticks, planes, V(r)) in FP32 and the
 * subspace GEMMs as BF16x3 split products. FP32_HX: the FFT stages in FP32,
 * the GEMMs in FP64 (no BF16x3 Gram noise, about 3e-7 Ha of Ritz error from
 * H.X). Factorisation, eigensolve and convergence control are FP64 in every
 * mode. */
typedef enum { BD_FP64 = 0, BD_MIXED = 1, BD_FP32_HX = 2 } bd_precision;

/* Last error message of the calling thread ("" when none). */
const char* bd_last_error(void);

/* ------------------------------------------------------------------------
 * Set-up of one Gamma cell with one synthetic species
 * Replaces refstart.rs:536-650 (crates/dft-engine/src/refstart.rs): plane-wave
 * basis (crates/dft-basis/src/lib.rs:234-269), FFT grid (lib.rs:414-455),
 * ORC half sphere (crates/dft-basis/src/half.rs:41-153), projector table
 * (crates/dft-physics/src/hamiltonian.rs:885-1035) packed to ORC rows
 * (crates/dft-physics/src/distributed_real_gamma.rs:363-402), ionic local
 * potential on the density sphere (hamiltonian.rs:211-340) and the Davidson
 * diagonal (refstart.rs:605-650). The pseudopotential is the reference's
 * synthetic Gaussian family (crates/dft-engine/tests/common/mod.rs:146-209).
 * ---------------------------------------------------------------------- */
typedef struct {
  double lattice_rows[3][3]; /* each row one lattice vector, Bohr */
  int32_t natoms;
  const double* frac_positions; /* 3*natoms */
  double z_valence;
  double rc;             /* local part -2 Z erf(r/rc)/r Ry */
  int32_t mesh_points;   /* log mesh to 12 Bohr */
  int32_t nproj;         /* 0 disables the nonlocal part */
  const int32_t* proj_l; /* nproj */
  const double* proj_width; /* beta = r^(l+1) exp(-r^2/width) */
  const double* dij_ry;  /* nproj*nproj row-major, Ry */
  double ecut_ha;
  int32_t grid_override[3]; /* all zero: automatic 2,3,5-smooth grid */
} bd_system_desc;

typedef struct bd_system bd_system;

typedef struct {
  int64_t n;          /* ORC rows = plane waves */
  int64_t nrep;       /* half-sphere representatives incl. G=0: (n+1)/2 */
  int32_t dims[3];    /* FFT grid */
  int64_t nproj;      /* projector columns (atoms x channels) */
  int64_t natoms;
  int64_t ndensity;   /* density-sphere size (4 Ecut) */
  double volume;      /* Bohr^3 */
  double electrons;
} bd_system_info;

bd_status bd_system_create(const bd_system_desc* desc, bd_system** out);
/* The same set-up with its atom loops on the GPU (the projector columns'
 * structure-factor phases and ORC packing, the Davidson diagonal, the ionic
 * potential's structure factors; hamiltonian.rs:885-1035, 211-340): the host
 * keeps the radial transforms and Y_lm. Same arrays to rounding; for large
 * cells (the Si512 cell's 192k rows and 1.5M density G against 512 atoms). */
bd_status bd_system_create_dev(const bd_system_desc* desc, bd_system** out);
void bd_system_destroy(bd_system* sys);
bd_status bd_system_get_info(const bd_system* sys, bd_system_info* out);
/* Copy one named set-up array into `out` (count elements; error if count
 * differs). Names (type, length):
 *   "miller"           int32, 3n   basis Miller indices, basis order
 *   "g2"               double, n   |G|^2, basis order
 *   "kinetic_orc"      double, n   T_G per ORC row (distributed_real_gamma.rs:352-361)
 *   "diagonal_static"  double, n   T_G + sum D_ij Re(b_i b_j*) per ORC row
 *   "projectors_orc"   double, n*nproj  ORC rows, column-major
 *   "coupling"         double, nproj*nproj  block-diagonal D (Ha), column-major
 *   "plus_index", "minus_index"  int64, nrep  flat FFT index of +G / -G of each rep
 *   "plus_basis", "minus_basis"  int64, nrep  basis index of +G / -G of each rep
 *                      (HalfSphere plus()/minus(), crates/dft-basis/src/half.rs:41-79)
 *   "density_miller"   int32, 3*ndensity; "density_g2" double, ndensity
 *   "density_index"    int64, ndensity  flat FFT index
 *   "ionic_density_re", "ionic_density_im" double, ndensity  V_ion(G), Ha
 *   "atoms_cart"       double, 3*natoms
 *   "ewald"            double, 1   Ewald energy of the ions (crates/dft-physics/src/ewald.rs)
 */
bd_status bd_system_get(const bd_system* sys, const char* name, void* out, int64_t count);

/* ------------------------------------------------------------------------
 * Gamma Hamiltonian operator (H·X on ORC columns)
 * Replaces DistributedRealHamiltonian::{new_owned, set_owned_potential,
 * apply_block} (crates/dft-physics/src/distributed_real_gamma.rs:222-1264) and
 * its DistributedRealOperator impl (:1581-1618, trait at
 * crates/dft-solvers/src/distributed_davidson.rs:19-40). Y is overwritten
 * (previous contents unspecified), NaN propagates (not scanned).
 * ---------------------------------------------------------------------- */
typedef struct bd_gamma_op bd_gamma_op;

/* Build the operator of a set-up system: FFT plan, device tables, scratch
 * for `capacity` columns. No device allocation happens after this call. */
bd_status bd_gamma_op_create(const bd_system* sys, int32_t capacity, int32_t precision,
                             bd_gamma_op** out);
void bd_gamma_op_destroy(bd_gamma_op* op);
/* V(r) on the full row-major n0*n1*n2 grid, Ha (set_owned_potential,
 * distributed_real_gamma.rs:476-492); epoch must increase. */
bd_status bd_gamma_op_set_potential(bd_gamma_op* op, const double* v_grid, uint64_t epoch);
bd_status bd_gamma_op_set_potential_dev(bd_gamma_op* op, const double* v_grid_dev, uint64_t epoch);
/* Y = H X for m <= capacity columns (apply_block, :1062-1264). */
bd_status bd_gamma_op_apply(bd_gamma_op* op, const double* x, int64_t ldx, int32_t m,
                            double* y, int64_t ldy);
bd_status bd_gamma_op_apply_dev(bd_gamma_op* op, const double* x_dev, int64_t ldx, int32_t m,
                                double* y_dev, int64_t ldy);
/* rho(r) = sum_b w_b psi_b(r)^2 / volume of the first m columns on the full
 * grid (band_density_sum, distributed_real_gamma.rs:741-769). */
bd_status bd_gamma_op_band_density_dev(bd_gamma_op* op, const double* x_dev, int64_t ldx,
                                       int32_t m, const double* weights, double* rho_dev);
/* The k points of a map (the k map's density sum, crates/dft-engine/src/refscf/
 * k_map.rs, one call instead of one per k point): rho = sum over k of the
 * band density of k-point op k (m[k] complex columns at x_dev[k], host
 * weights[k]). The k points run side by side on their ops' streams, each
 * into its own grid, summed in k order at the end: the same bits as the
 * per-k calls added in k order. k-point ops of one grid. `stream` (a HIP
 * stream, or NULL): the work is ordered after what it holds. Synchronises. */
bd_status bd_k_band_density_dev(bd_gamma_op* const* ops, int32_t count, const double* const* x_dev,
                                const int64_t* ldx, const int32_t* m, const double* const* weights,
                                double* rho_dev, void* stream);
/* Set every op's potential from one grid (bd_gamma_op_set_potential_dev per
 * op), ordered after `stream` (or NULL), which then waits for every copy. */
bd_status bd_k_set_potential_dev(bd_gamma_op* const* ops, int32_t count, const double* v_grid_dev,
                                 uint64_t epoch, void* stream);
/* Warm start of an SCF map (KeptHx: crates/dft-engine/src/refscf/kept_hx.rs,
 * apply_local_change of distributed_real_gamma.rs:1370-1438). With
 * keep_fields on, every band density also keeps its columns' real-space pair
 * fields on the device (capacity/2 x n0 n1 n2 complex values; Gamma ops on
 * the one-workgroup plane pass). apply_kept then adds dV X to Y for those
 * same columns (x_dev, ldx, m as passed to the density; an error otherwise):
 * only the r -> G half of the local transform runs, with no kinetic or
 * nonlocal term, so H_new X = H_old X + dV X with dV = V_new - V_old (full
 * grid, Ha). */
bd_status bd_gamma_op_keep_fields(bd_gamma_op* op, int32_t on);
bd_status bd_gamma_op_apply_kept_dev(bd_gamma_op* op, const double* x_dev, int64_t ldx, int32_t m,
                                     const double* delta_v_dev, double* y_dev, int64_t ldy);
/* x_b^T P D P^T x_b per column (nonlocal_band_energies, :878-928). */
bd_status bd_gamma_op_nonlocal_energies_dev(bd_gamma_op* op, const double* x_dev, int64_t ldx,
                                            int32_t m, double* out_host);
/* HIP stream the op runs on (hipStream_t as void*). */
void* bd_gamma_op_stream(bd_gamma_op* op);
/* Kernel timing with HIP events on the op's stream (adds one event wait per
 * apply). set_profiling(on) also resets the totals. stage_times fills 8
 * values: cumulative ms of [stick_inverse, plane, stick_forward, projections,
 * back_product] since profiling was switched on, then applies, pairs, columns. */
bd_status bd_gamma_op_set_profiling(bd_gamma_op* op, int32_t on);
/* FFT plan shape: n0, n1, n2, occupied sticks S, padded sticks, occupied z
 * planes Z, sticks per tile, tiles. */
bd_status bd_gamma_op_plan_info(bd_gamma_op* op, int64_t* out8);
bd_status bd_gamma_op_stage_times(bd_gamma_op* op, double* out8);

/* ------------------------------------------------------------------------
 * k-point (complex) operator
 * Replaces DistributedComplexHamiltonian::{new_owned, set_owned_potential,
 * apply_block, project} and static_diagonal
 * (crates/dft-physics/src/distributed_complex.rs:29-54, 88-386) on the basis
 * PlaneWaveBasis::at_k (crates/dft-basis/src/lib.rs:234-269) of the system's
 * cell, grid and pseudopotential. The handle is a bd_gamma_op whose n = 2 nk
 * rows are the interleaved (Re, Im) parts of nk complex coefficients: a
 * column is a complex vector, and the real symmetric operator on those rows
 * is [[Re H, -Im H], [Im H, Re H]] in interleaved order. Every bd_gamma_op_*
 * entry point and bd_davidson_solve_dev accept it: apply runs one complex FFT
 * per column; band_density sums w_b |psi_b(r)|^2 / volume; the Davidson on
 * it, started from pairs (x, i x), is the complex Davidson of the reference
 * (ref_davidson_complex.rs:641-1113) on the realified subspace.
 * k_frac is wrapped into [-1/2, 1/2) (crates/dft-geometry/src/kpoints.rs:21-32).
 * ---------------------------------------------------------------------- */
bd_status bd_k_op_create(const bd_system* sys, const double* k_frac3, int32_t capacity, int32_t precision,
                         bd_gamma_op** out);
/* Basis size nk and the wrapped fractional / Cartesian (Bohr^-1) k. */
bd_status bd_k_op_info(const bd_gamma_op* op, int64_t* nk, double* k_frac3, double* k_cart3);
/* Copy one named array (count elements, error if count differs):
 *   "miller" int32 3nk; "g2", "q2" double nk (|G|^2, |k+G|^2);
 *   "fft_index" int64 nk; "kinetic", "diagonal_static" double 2nk (per row);
 *   "projectors" double 2nk x 2p column-major: column c = (Re P_c, Im P_c)
 *   interleaved, column p + c = (-Im P_c, Re P_c). */
bd_status bd_k_op_get(const bd_gamma_op* op, const char* name, void* out, int64_t count);

/* ------------------------------------------------------------------------
 * Batched k points (no reference counterpart: the reference applies one k
 * point at a time, crates/dft-engine/src/refscf/k_map.rs). A k-point op whose
 * FFT fits one workgroup's LDS applies H with one workgroup per column (the
 * whole 3D FFT of the column in LDS); its local part is bit-identical to the
 * three-stage stick/plane path. A batch runs every column of every k point
 * in one persistent launch, nonlocal term included.
 * ---------------------------------------------------------------------- */
typedef struct bd_k_batch bd_k_batch;
/* All ops: k-point ops of one system and precision on the fused path. */
bd_status bd_k_batch_create(bd_gamma_op* const* ops, int32_t count, bd_k_batch** out);
void bd_k_batch_destroy(bd_k_batch* batch);
void* bd_k_batch_stream(bd_k_batch* batch);
/* Y_i = H_i X_i for each op i (device blocks, 16-byte aligned, even ld >= 2 nk_i,
 * m_i <= capacity_i). Ordered after each op's queued work; returns when done. */
bd_status bd_k_batch_apply_dev(bd_k_batch* batch, const double* const* x_dev, const int64_t* ldx,
                               const int32_t* m, double* const* y_dev, const int64_t* ldy);
/* Launch-shape autotuner of the fused kernel: times every valid (threads,
 * sticks per x-FFT batch, rows per z-FFT batch, y-z planes per pass) on m
 * columns, reps launches each, keeps the fastest.
 * out5: threads, xlines, zrows, xplanes, microseconds per launch. */
bd_status bd_gamma_op_autotune(bd_gamma_op* op, int32_t m, int32_t reps, double* out5);
/* Read (set = 0) or set the fused shape (threads, xlines, zrows, xplanes);
 * threads = 0 selects the three-stage path. */
bd_status bd_gamma_op_fused_shape(bd_gamma_op* op, int32_t* inout4, int32_t set);
/* Read (set = 0) or set the plane kernel's shape on the stick/plane path
 * (threads 256 or 512, y rows per z-FFT chunk); set accepts the shapes the
 * autotuner may pick. All shapes give the same bits. */
bd_status bd_gamma_op_plane_shape(bd_gamma_op* op, int32_t* inout2, int32_t set);
/* Read (set = 0) or set the plane pass: 0 = one workgroup per y-z plane
 * (only where the Z x n1 plane fits the LDS), zb > 0 = the split pass, which
 * moves each plane through HBM and transforms zb z-lines (then a chunk of y
 * rows) at a time. The op picks the split pass at create when the plane does
 * not fit (the 144^3 grid of the Si512 cell in FP64); setting it on a smaller
 * cell allocates its plane buffer. Both give the same bits. Replaces the
 * plane stage of distributed_fft.rs:956-1080, which handles any plane size. */
bd_status bd_gamma_op_plane_split(bd_gamma_op* op, int32_t* inout, int32_t set);

/* ------------------------------------------------------------------------
 * Reference block Davidson on the op
 * Replaces ref_davidson::solve_with_initial_hx
 * (crates/dft-solvers/src/ref_davidson.rs:117-366): bordered inverse Cholesky
 * (lean_davidson.rs:167-331), Ritz-change convergence, contraction at
 * capacity, rank repair. Options mirror ref_davidson::Options (:38-78).
 * ---------------------------------------------------------------------- */
/* BD_PRECOND_MLP: the learned row factor K on the QE correction,
 * e = -K r / d_QE (no reference counterpart; see the bd_mlp_* entry points),
 * with the QE correction (K = 1) as the fallback. */
typedef enum { BD_PRECOND_QE = 0, BD_PRECOND_TPA = 1, BD_PRECOND_MLP = 2 } bd_preconditioner;

typedef struct {
  int32_t bands;
  double threshold;      /* occupied Ritz-change threshold, Ha (strict) */
  int32_t occupied;
  int32_t first_step;    /* all bands use `threshold` */
  int32_t capacity;      /* 0: min(2*bands, n) */
  int32_t max_iterations;/* default 20 */
  int32_t return_hx;
  int32_t preconditioner;/* bd_preconditioner */
  const float* mlp_blob; /* BD_PRECOND_MLP: device blob from bd_mlp_quantize_dev */
  double mlp_vbar;       /* BD_PRECOND_MLP: mean local potential (Ha), a feature */
  int32_t profile;       /* nonzero: fill the Ritz history below (Options::profile) */
  /* Host arrays of max_iterations x bands (row per iteration), or NULL:
   * |theta_new - theta_old| per band (Solution::ritz_changes_ha), and the
   * pairs (theta_old, theta_new) (Solution::ritz_values_ha, 2 per entry).
   * Rows past record.iterations are left untouched. */
  double* ritz_changes_ha;
  double* ritz_values_ha;
  /* Nonzero: the columns are complex vectors (interleaved (re, im) rows, a
   * k-point op's layout) and the complex Hermitian Davidson runs
   * (ref_davidson_complex::solve_with_initial_hx,
   * crates/dft-solvers/src/ref_davidson_complex.rs:641-1113): `bands` and
   * `capacity` count complex columns (capacity 0: min(2 bands, n / 2)), the
   * Grams are Hermitian, the reduced problem is the complex one (each value
   * once). k-point ops only; QE or TPA preconditioner. */
  int32_t complex_columns;
} bd_davidson_options;

typedef struct {
  int32_t iterations;
  int32_t h_calls;
  int32_t h_columns;
  int32_t repairs;
  int32_t restarts;
  int32_t all_converged;
  double seconds_total;
  double seconds_apply;
  double seconds_dense;  /* subspace step: factor + reduced eigensolve (GPU, synchronised) */
  int32_t mlp_fallback;  /* BD_PRECOND_MLP: iteration that fell back to QE, -1 if none */
  int32_t eigen_retries; /* reduced eigensolves that lost a vector of a tight cluster and were
                            redone with the one-step tridiagonalisation */
  /* Restarts by kind (ResetStats, ref_davidson/ideal_reset.rs): ideal resets
   * (S = I, H_r = diag(theta)), periodic measured refreshes (every 4th), and
   * guard trips (after which the solve refreshes by measurement). */
  int32_t ideal_resets;
  int32_t periodic_refreshes;
  int32_t guard_fallbacks;
} bd_davidson_record;

/* X0 (n x bands, device), optional HX0 (device or NULL), diagonal (n, device).
 * Outputs: eigenvalues (host, bands), X and HX (device, n x bands; HX may be
 * NULL unless return_hx), converged flags (host, bands). Returns
 * BD_NOT_CONVERGED when the iteration limit was hit (outputs still valid). */
bd_status bd_davidson_solve_dev(bd_gamma_op* op, const double* x0_dev, int64_t ldx0,
                                const double* hx0_dev, const double* diag_dev,
                                const bd_davidson_options* options, double* eigenvalues,
                                double* x_dev, double* hx_dev, int64_t ldx, int32_t* converged,
                                bd_davidson_record* record);
/* Independent solves on `count` distinct ops (the k points of one SCF map;
 * no reference counterpart, the reference solves them in sequence,
 * crates/dft-engine/src/refscf/k_map.rs), `threads` host threads at a time,
 * each op on its own stream. Per-solve arrays; hxout and conv may be NULL
 * (or hold NULL entries). statuses[i] per solve; returns the first status
 * other than OK / NOT_CONVERGED (its message as the last error), else
 * BD_NOT_CONVERGED if any solve hit its limit, else BD_OK. */
bd_status bd_davidson_solve_many_dev(int32_t count, bd_gamma_op* const* ops, const double* const* x0_dev,
                                     const int64_t* ldx0, const double* const* diag_dev,
                                     const bd_davidson_options* options, double* const* eigenvalues,
                                     double* const* x_dev, double* const* hx_dev, const int64_t* ldx,
                                     int32_t* const* converged, bd_davidson_record* records,
                                     bd_status* statuses, int32_t threads);

/* The complex Davidson of every k point of a map in lock step
 * (ref_davidson_complex::solve_with_initial_hx,
 * crates/dft-solvers/src/ref_davidson_complex.rs:641-1113, for the k points
 * that crates/dft-engine/src/refscf/k_map.rs:181-208 solves one after the
 * other). Each k point keeps its own basis, active bands, convergence,
 * contractions and ideal resets; each stage of an iteration is one launch
 * over all k points, H of every k point's new columns one launch of the fused
 * k-point kernel, and the subspace step of a k point (bordered complex
 * Cholesky, L^-1 H L^-H, a complex Jacobi eigensolve, C = L^-H Y) one
 * workgroup with its matrices in LDS. Two status read-backs per iteration
 * cover all k points. The rank repair of a k point whose border the factor
 * rejects runs inside the lock step (counted in record.repairs). The ops:
 * k-point ops on the fused path of one
 * system and precision, each at most once; bands <= 24, capacity (0: 2 bands)
 * <= 48 complex columns and <= n_k. */
typedef struct bd_k_davidson bd_k_davidson;
bd_status bd_k_davidson_create(bd_gamma_op* const* ops, int32_t count, int32_t bands, int32_t capacity,
                               bd_k_davidson** out);
void bd_k_davidson_destroy(bd_k_davidson* solver);
/* Per k point: x0 (n x bands complex columns, ld ldx0 even), optional hx0
 * (NULL, or an array of H·x0 blocks with the same ld), diag (n rows); one
 * options struct for all (bands as created, complex_columns ignored, QE or TPA,
 * no Ritz histories); eigenvalues and converged: count x bands (host); x and
 * hx (hx when return_hx): n x bands blocks, ld ldx; records and statuses: one
 * per k point. Returns the first failed k point's status (its message as the
 * last error), else BD_NOT_CONVERGED if any k point hit the limit, else OK. */
bd_status bd_k_davidson_solve_dev(bd_k_davidson* solver, const double* const* x0_dev, const int64_t* ldx0,
                                  const double* const* hx0_dev, const double* const* diag_dev,
                                  const bd_davidson_options* options, double* eigenvalues,
                                  double* const* x_dev, double* const* hx_dev, const int64_t* ldx,
                                  int32_t* converged, bd_davidson_record* records, bd_status* statuses);

/* ------------------------------------------------------------------------
 * Fixed-shape block iteration (no reference counterpart; the iteration of
 * ref_davidson.rs:117-366 with every band active and a restart to the Ritz
 * block after each step): V = [X, W] with W the preconditioned residual
 * directions, Rayleigh-Ritz on V (bordered inverse Cholesky, FP64 GPU
 * eigensolve), X <- V C. Each step is device work with fixed shapes and no
 * host round trip; with use_graph the step is captured once as a HIP graph
 * and replayed, bit-identical to the eager step.
 * ---------------------------------------------------------------------- */
typedef struct bd_block bd_block;
/* bands: 2 bands <= n, bands <= the op's capacity; QE or TPA preconditioner. */
bd_status bd_block_create(bd_gamma_op* op, int32_t bands, int32_t preconditioner, bd_block** out);
void bd_block_destroy(bd_block* block);
/* Davidson diagonal (n rows, device), copied in; read by every later step. */
bd_status bd_block_set_diagonal_dev(bd_block* block, const double* diag_dev);
/* Rayleigh-Ritz of a start block (n x bands, ldx0): X, HX, values (host). */
bd_status bd_block_start_dev(bd_block* block, const double* x0_dev, int64_t ldx0, double* values);
/* `steps` iterations; values (host, bands) after the last. Synchronises. */
bd_status bd_block_step(bd_block* block, int32_t steps, int32_t use_graph, double* values);
/* The same without waiting; bd_block_finish waits and reads the values. */
bd_status bd_block_step_async(bd_block* block, int32_t steps, int32_t use_graph);
bd_status bd_block_finish(bd_block* block, double* values);
/* X (and HX unless NULL), n x bands, leading dimension ld. */
bd_status bd_block_vectors_dev(bd_block* block, double* x_dev, double* hx_dev, int64_t ld);

/* The block iterations of all k points of a map stepped together (no
 * reference counterpart: the batched form of the k map's per-k solves,
 * crates/dft-engine/src/refscf/k_map.rs:181-208): per-k work on the k
 * points' own streams, one fused launch of H·W over every column of every k
 * point, the whole iteration captured once as a HIP graph. The ops must be
 * k-point ops on the fused path of one system and precision (as for
 * bd_k_batch_create), each at most once; bands as for bd_block_create. */
typedef struct bd_kblocks bd_kblocks;
bd_status bd_kblocks_create(bd_gamma_op* const* ops, int32_t count, int32_t bands, int32_t preconditioner,
                            bd_kblocks** out);
void bd_kblocks_destroy(bd_kblocks* blocks);
/* The Davidson diagonal of every k point (device, n_k rows each), copied in. */
bd_status bd_kblocks_set_diagonals_dev(bd_kblocks* blocks, const double* const* diag_dev);
/* Rayleigh-Ritz of every start block (n_k x bands, ldx0[k]), ordered after
 * the work queued on the ops' streams; values: count x bands (host). */
bd_status bd_kblocks_start_dev(bd_kblocks* blocks, const double* const* x0_dev, const int64_t* ldx0,
                               double* values);
/* `steps` iterations of every k point; values (count x bands, host) after
 * the last. Synchronises once. */
bd_status bd_kblocks_step(bd_kblocks* blocks, int32_t steps, int32_t use_graph, double* values);
/* X (and HX where hx_dev and hx_dev[k] are not NULL) of every k point,
 * n_k x bands with leading dimension ld[k]. */
bd_status bd_kblocks_vectors_dev(bd_kblocks* blocks, double* const* x_dev, double* const* hx_dev,
                                 const int64_t* ld);

/* ------------------------------------------------------------------------
 * Dense products of the path as stand-alone entry points
 * C = alpha op(A) B + beta C, column-major, device pointers; op(A) = A^T when
 * trans_a (then a split-K reduction needs `work`, bd_gemm_workspace doubles;
 * the size returned for M, N also covers every smaller M, N at the same K).
 * mode 0: FP64; mode 1: BF16x3 split products on MFMA, FP32 accumulation.
 * These are the products behind PᵀX / P·W (distributed_real_gamma.rs:957-1251)
 * and the Davidson border Grams / rotations (ref_davidson/workspace.rs:300-381,
 * 982-1012); no single reference function corresponds to the BF16x3 mode.
 * ---------------------------------------------------------------------- */
int64_t bd_gemm_workspace(int32_t mode, int32_t trans_a, int64_t M, int64_t N, int64_t K);
bd_status bd_gemm_dev(int32_t mode, int32_t trans_a, int64_t M, int64_t N, int64_t K, double alpha,
                      const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                      int64_t ldc, double* work);
/* The same product reps times back to back on one stream between two HIP
 * events; *ms = the average per product (kernel time without the call's
 * host round trip). No reference counterpart (measurement). */
bd_status bd_gemm_time_dev(int32_t mode, int32_t trans_a, int64_t m, int64_t n, int64_t k, const double* a,
                           int64_t lda, const double* b, int64_t ldb, double* c, int64_t ldc, double* work,
                           int32_t reps, double* ms);
/* One v_mfma_f32_16x16x16_bf16 of row-major 16x16 bf16 bit patterns (a self
 * check of the MFMA operand layout the kernels assume). */
/* Lane-layout probe of v_mfma_f64_16x16x4f64 (the FP64 GEMMs' product):
 * D (16x16) = A (16x4) B (4x16), all row-major device doubles. */
bd_status bd_mfma_f64_probe_dev(const double* a_dev, const double* b_dev, double* d_dev);
bd_status bd_mfma_probe_dev(const uint16_t* a_bits, const uint16_t* b_bits, float* d);

/* ------------------------------------------------------------------------
 * The FP64 subspace step of bd_davidson_solve_dev as stand-alone entry points
 * (device pointers; synchronous on the null stream).
 * bd_sym_eigen_dev: lowest `want` eigenpairs of the symmetric m x m matrix A
 *   (column-major, lda; overwritten): values ascending, Y (m x want, ldy)
 *   orthonormal. The role of RealEighWorkspace::solve_into
 *   (crates/dft-numerics/src/real.rs:2286-2330) as called from
 *   ref_davidson/workspace.rs:435-499, which reads only the lowest `bands`
 *   pairs. `work` holds bd_sym_eigen_workspace(m, want) doubles.
 * bd_inverse_cholesky_dev: extends T = L^{-1} of the overlap S (both cap x cap,
 *   ld cap; T lower triangular with a zero upper triangle) from `from` to `to`
 *   columns (crates/dft-solvers/src/lean_davidson.rs:167-331). *accepted is 0
 *   when the pivot-ratio (eps/1e-12) or row-growth test rejects the border;
 *   *min_ratio is the smallest squared pivot over its overlap diagonal.
 * ---------------------------------------------------------------------- */
int64_t bd_sym_eigen_workspace(int32_t m, int32_t want);
bd_status bd_sym_eigen_dev(int32_t m, int32_t want, double* A, int64_t lda, double* values, double* Y,
                           int64_t ldy, double* work);
/* bd_herm_eigen_dev: lowest `want` eigenpairs of a complex Hermitian m x m
 * matrix (the k-point Davidson's reduced problem; replaces the complex
 * RealEighWorkspace counterpart, crates/dft-numerics/src/linalg.rs): A holds
 * both triangles as interleaved (re, im) doubles with leading dimension lda
 * complex entries and is overwritten; Y receives m x want complex vectors
 * (ld ldy complex entries); `work` holds bd_sym_eigen_workspace(m, want)
 * doubles. Complex Householder reduction to a real tridiagonal, the real
 * bisection / inverse iteration, complex back-transform. */
bd_status bd_herm_eigen_dev(int32_t m, int32_t want, double* A, int64_t lda, double* values, double* Y,
                            int64_t ldy, double* work);
/* Read (set = 0) or set the subspace size from which the tridiagonalisation
 * runs across the chip (one launch per reflector for its column, one for the
 * product with the trailing block on every CU, a tiled panel update) instead
 * of in one workgroup; default 1026. The solver takes m <= 2730. */
bd_status bd_sym_eigen_multi_from(int32_t* inout, int32_t set);
/* Read (set = 0) or set {workers, from, to, spins} of the cooperative
 * tridiagonalisation: for from <= m <= to (and below the multi_from size) the
 * matrix is split by columns over the LDS of `workers` workgroups (more
 * where the columns do not fit, at most 32; -1: m/10 clamped to 8..32),
 * which pass each reflector and each product through tagged words in device
 * memory; workers = 0 keeps the one-workgroup panel kernel. A worker that
 * polls a word more than `spins` times (at once when spins < 0) abandons the
 * solve, which then fails like any eigensolve (the Davidson redoes it with
 * the one-step kernel). Default {-1, 128, 1025, 2097152}. */
bd_status bd_sym_eigen_coop(int32_t* inout, int32_t set);
bd_status bd_inverse_cholesky_dev(const double* S, double* T, int32_t cap, int32_t from, int32_t to,
                                  int32_t* accepted, double* min_ratio);
/* bd_inverse_cholesky_static_dev: the same extension with the same results,
 * allocating nothing. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Its scratch is one file-scope
 * device array per GPU (16.4 MB, held until the process ends), so cap must
 * be <= 640 (BD_ERR_INVALID_PARAMETER otherwise). Calls that overlap on one
 * GPU share that array and race; serialise them. */
bd_status bd_inverse_cholesky_static_dev(const double* S, double* T, int32_t cap, int32_t from, int32_t to,
                                         int32_t* accepted, double* min_ratio);

/* ------------------------------------------------------------------------
 * SCF field step (the caller side of the path: the potential of an input
 * density, the residual and its metric, the mixing update and the energy
 * terms; crates/dft-engine/src/refscf/map.rs:41-198, metric.rs,
 * broyden.rs:106-113, 317-445; xc.rs:48-73 LDA-PZ). Coefficients live on the
 * density sphere (4 Ecut, bd_system_info.ndensity complex values, interleaved
 * Re/Im); fields on the full n0 n1 n2 grid (real). Device pointers; each call
 * returns when its work is done.
 * ---------------------------------------------------------------------- */
typedef struct bd_field bd_field;
bd_status bd_field_create(const bd_system* sys, bd_field** out);
void bd_field_destroy(bd_field* field);
void* bd_field_stream(bd_field* field);
/* grid (complex, N) = unnormalised inverse transform of sphere coefficients. */
bd_status bd_field_to_grid_dev(bd_field* field, const double* coeff, double* grid);
/* coeff = forward transform / N of a real grid field, on the sphere. */
bd_status bd_field_from_grid_dev(bd_field* field, const double* real_grid, double* coeff);
/* v = V_ion + V_H[x] + v_xc(|rho_in|) of input coefficients x; *vmean = mean(v). */
bd_status bd_field_potential_dev(bd_field* field, const double* x, const double* vion, double* v, double* vmean);
/* f = coefficients of rho_out (G = 0 set to electrons/volume) - x;
 * *dr2 = half-sphere Hartree-metric norm of f, Ry (metric.rs). */
bd_status bd_field_residual_dev(bd_field* field, const double* rho_out, const double* x, double* f, double* dr2);
/* a (p x p, row-major) = metric products of p <= 9 rows (p x nd complex),
 * c (p) = their products with f. */
bd_status bd_field_gram_dev(bd_field* field, int32_t p, const double* rows, const double* f, double* a, double* c);
/* out = (x - sum gamma_i dX_i) + beta (f - sum gamma_i dF_i) [* G^2/(G^2 + a_TF)
 * when screening]; gamma is a host array of p values. */
bd_status bd_field_mix_dev(bd_field* field, int32_t p, const double* dX, const double* dF, const double* gamma,
                           const double* x, const double* f, double beta, int32_t screening, double* out);
/* out3 = (Hartree, XC, local) energies (Ha) of a real grid density. */
bd_status bd_field_energies_dev(bd_field* field, const double* rho, const double* vion, double* out3);

/* ------------------------------------------------------------------------
 * Learned Davidson preconditioner (no reference counterpart: it takes the
 * place of the Teter-Payne-Allan weight of crates/dft-solvers/src/lib.rs:
 * 176-280 in e = -K r). K = exp(MLP(u)), a 4 -> 64 -> 64 -> 1 tanh network of
 * u = [log(1 + g/T), log T, (theta - vbar)/T, vbar] (g = |G|^2/2 of the row,
 * T the column's kinetic scale, theta its Ritz value, vbar the mean local
 * potential). Parameters are bd_mlp_param_count() FP32 values (W1 64x4, b1,
 * W2 64x64, b2, w3, b3; row-major); the device blob (bd_mlp_blob_floats()
 * floats) holds bf16 weights for the MFMA kernels.
 * Training data: `bands` residuals R and ideal corrections Tg (float, bands x
 * n, row-major per band), per-band (T, theta, vbar) in scal (bands x 3), the
 * row energies g (n). Loss per band: 1 - <K r, t>^2 / (|K r|^2 |t|^2). The
 * Davidson applies K to the QE-scaled residual r / d_QE, so training pairs
 * for it carry that residual in R.
 * ---------------------------------------------------------------------- */
int32_t bd_mlp_param_count(void);
int32_t bd_mlp_blob_floats(void);
bd_status bd_mlp_quantize_dev(const float* theta, float* blob);
bd_status bd_mlp_forward_dev(int64_t n, const float* g, float T, float ritz, float vbar, const float* blob,
                             float* K);
int64_t bd_mlp_train_workspace(int64_t n, int32_t batch);
bd_status bd_mlp_loss_grad_dev(int64_t n, int32_t bands, const float* g, const float* R, const float* Tg,
                               const float* scal, const int32_t* batch_dev, int32_t batch, const float* blob,
                               float* grad, double* loss, double* work);
bd_status bd_mlp_adam_dev(float* theta, const float* grad, float* m, float* v, int32_t step, float lr);
/* `steps` Adam steps on batches of `batch` bands drawn by a fixed LCG from
 * `seed`; losses (host, steps) receives each step's batch loss. batch_dev
 * holds `batch` ints; theta, m, v are the FP32 masters and moments. */
bd_status bd_mlp_train_dev(int64_t n, int32_t bands, const float* g, const float* R, const float* Tg,
                           const float* scal, float* theta, float* m, float* v, int32_t steps, int32_t batch,
                           float lr, uint64_t seed, double* losses, int32_t* batch_dev, float* blob, float* grad,
                           double* work);

#ifdef __cplusplus
}
#endif
#endif /* BAANDI_AMD_H */
