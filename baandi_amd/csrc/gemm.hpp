// FP64 tiled GEMMs used by the projector and subspace products.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace bd {

// C(MxN, ldc) = alpha * A(MxK, lda) * B(KxN, ldb) + beta * C; all column-major.
void gemm_nn(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
             int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc);

// C(MxN, ldc) = alpha * A^T B + beta * C with A (KxM, lda), B (KxN, ldb): a
// tall-skinny reduction over K, split into fixed chunks whose partial tiles
// are summed in ascending chunk order (deterministic). `work` must hold
// gemm_tn_work(M, N, K) doubles, a size that also covers every smaller M, N
// at the same K.
void gemm_tn(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
             int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
             double* work);
int64_t gemm_tn_work(int64_t M, int64_t N, int64_t K);

// BF16x3 split-precision variants on MFMA (FP32 accumulation, FP64 in/out).
void gemm_nn_bf16x3(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                    const double* B, int64_t ldb, double beta, double* C, int64_t ldc);
void gemm_tn_bf16x3(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                    const double* B, int64_t ldb, double beta, double* C, int64_t ldc, double* work);
int64_t gemm_tn_bf16x3_work(int64_t M, int64_t N, int64_t K);
void mfma_probe(hipStream_t s, const unsigned short* A, const unsigned short* B, float* D);

// Precision-dispatching pair used by the operator and the Davidson solver.
struct Gemm {
  bool bf16x3 = false;
  void nn(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
          const double* B, int64_t ldb, double beta, double* C, int64_t ldc) const {
    if (bf16x3) gemm_nn_bf16x3(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    else gemm_nn(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  }
  void tn(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
          const double* B, int64_t ldb, double beta, double* C, int64_t ldc, double* work) const {
    if (bf16x3) gemm_tn_bf16x3(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, work);
    else gemm_tn(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, work);
  }
  static int64_t tn_work(int64_t M, int64_t N, int64_t K) {
    return std::max(gemm_tn_work(M, N, K), gemm_tn_bf16x3_work(M, N, K));
  }
};

}  // namespace bd
