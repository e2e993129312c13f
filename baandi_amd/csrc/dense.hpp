// Small dense FP64 linear algebra of the Davidson subspace step (host).
#pragma once
#include <cstdint>
#include <vector>

namespace bd {

// Bordered inverse Cholesky (crates/dft-solvers/src/lean_davidson.rs:167-331):
// S is symmetric, column-major with leading dimension ld; T holds L^{-1} for
// the leading `from` columns on entry and for `to` columns on success.
// Rejects when a new squared pivot over its overlap diagonal falls below
// eps/1e-12 or the row growth test fails (T is then valid for `from` only).
bool extend_inverse_cholesky(const double* S, int64_t ld, int from, int to, double* T, double* min_ratio);

// All eigenpairs of the symmetric m x m matrix A (column-major, lda):
// Householder tridiagonalisation and implicit-shift QL (Wilkinson & Reinsch,
// Handbook for Automatic Computation II, tred2/tql2). Values ascending;
// vectors column-major m x m (ldv = m). Returns false if QL fails to converge.
bool symmetric_eigen(const double* A, int64_t lda, int m, double* values, double* vectors);

}  // namespace bd
