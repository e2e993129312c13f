// FP64 symmetric eigensolver and bordered inverse Cholesky of the Davidson
// subspace step, on the GPU (stream ordered; no host round trips inside).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace bd {

// Lets launches of kernel `fn` ask for up to the whole 160 KB of LDS (set once
// per kernel; `name` labels the error).
void allow_full_lds(const void* fn, const char* name);

// Doubles of workspace needed by sym_eigen_lowest for an m x m matrix and
// `want` eigenpairs.
int64_t sym_eigen_work(int m, int want);

// Lowest `want` eigenpairs of the symmetric m x m matrix A (column-major,
// leading dimension lda; overwritten by the Householder vectors). values
// (device, `want`) ascending; Y (m x want, ldy) orthonormal eigenvectors.
// *info (device) is set to 0, or to 1 + j when vector j could not be made
// orthogonal to its cluster. `blocked` selects the panel tridiagonalisation
// where it fits (else the one-step kernel).
void sym_eigen_lowest(hipStream_t s, int m, int want, double* A, int64_t lda, double* values, double* Y,
                      int64_t ldy, double* work, int* info, bool blocked = true);
// The same for a complex Hermitian A (both triangles, interleaved (re, im),
// ld lda complex entries): Y complex m x want (ld ldy complex entries); the
// same workspace size.
void herm_eigen_lowest(hipStream_t s, int m, int want, double* A, int64_t lda, double* values, double* Y,
                       int64_t ldy, double* work, int* info);

// Device state of the bordered inverse Cholesky of the subspace overlap.
struct FactorWork {
  double* W = nullptr;  // cap x cap
  double* G = nullptr;  // cap x cap
  double* U = nullptr;  // cap x cap
  double* gwork = nullptr;
  int* flag = nullptr;          // 0 ok, else rejected
  double* min_ratio = nullptr;  // smallest pivot ratio of the last call
  // Row growth guard of the real factor (lean_davidson.rs); off for the
  // complex factor, whose extension checks only the pivots and finiteness
  // (crates/dft-numerics/src/linalg.rs:277-340).
  int growth_guard = 1;
};

// Extend T = L^{-1} of S (both cap x cap, ld cap, lower triangular T with a
// zero upper triangle) from `from` to `to` columns. Writes the verdict into
// fw.flag / fw.min_ratio; the caller reads them back. Semantics of
// crates/dft-solvers/src/lean_davidson.rs:167-331 (pivot ratio eps/1e-12,
// row growth guard).
void extend_inverse_cholesky_dev(hipStream_t s, const double* S, double* T, int cap, int from, int to,
                                 const FactorWork& fw);
int64_t factor_gwork(int cap);

// M[j, i] = M[i, j] for i < j, j in [from, to) (ld cap): mirror a bordered
// column block into its rows, taking the upper triangle of the new block.
void mirror_border(hipStream_t s, double* M, int cap, int from, int to);
// B (n x m, ldb) = A^T for A (m x n, lda).
void transpose(hipStream_t s, int m, int n, const double* A, int64_t lda, double* B, int64_t ldb);
// A = (A + A^T) / 2 in place (m x m, lda).
void symmetrize(hipStream_t s, int m, double* A, int64_t lda);
// A (m x n, lda) = identity block (1 on the diagonal, 0 elsewhere).
void set_identity(hipStream_t s, int m, int n, double* A, int64_t lda);
// Zero `count` ints / doubles (a fill kernel, so that the call may sit in a captured graph).
void zero_ints(hipStream_t s, int* p, int count);
void zero_doubles(hipStream_t s, double* p, int64_t count);

}  // namespace bd
