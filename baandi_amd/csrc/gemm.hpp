// FP64 tiled GEMMs used by the projector and subspace products.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace bd {

// C(MxN, ldc) = alpha * A(MxK, lda) * B(KxN, ldb) + beta * C; all column-major.
void gemm_nn(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
             int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc);

// C(MxN, ldc) = alpha * A^T B + beta * C with A (KxM, lda), B (KxN, ldb): a
// tall-skinny reduction over K, split into fixed chunks whose partial tiles
// are summed in ascending chunk order (deterministic). `work` must hold
// gemm_tn_work(M, N, K) doubles.
void gemm_tn(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
             int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
             double* work);
int64_t gemm_tn_work(int64_t M, int64_t N, int64_t K);

}  // namespace bd
