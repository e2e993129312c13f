// Learned Davidson preconditioner: a 4 -> 64 -> 64 -> 1 tanh network giving
// a row factor K that scales the QE correction, e = -K r / d_QE(h, theta)
// (no reference counterpart; the QE denominator is ref_davidson/stream.rs:16-22
// and the preconditioner it joins is crates/dft-solvers/src/lib.rs:176-280).
// K = 1 is the QE correction itself, so an untrained output layer starts at QE.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace bd {
namespace mlp {

constexpr int kHidden = 64;
constexpr int kFeatures = 4;
// Packed FP32 parameters: W1 (64x4, row-major), b1, W2 (64x64, row-major), b2, w3, b3.
constexpr int oW1 = 0;
constexpr int ob1 = oW1 + kHidden * kFeatures;
constexpr int oW2 = ob1 + kHidden;
constexpr int ob2 = oW2 + kHidden * kHidden;
constexpr int ow3 = ob2 + kHidden;
constexpr int ob3 = ow3 + kHidden;
constexpr int kCount = ob3 + 1;  // 4545
// Inference blob (floats): bf16-rounded W1, b1, b2, bf16-rounded w3, b3
// (first 512 floats), then W2 and W2^T as bf16 bit patterns (2048 floats each).
constexpr int kBlobW2 = 512;
constexpr int kBlobW2T = kBlobW2 + kHidden * kHidden / 2;
constexpr int kBlobFloats = kBlobW2T + kHidden * kHidden / 2;
constexpr float kKineticFloor = 0.1f;  // kinetic_scale floor, lib.rs:178-193

// QE smooth denominator in Ha for the Davidson diagonal h and Ritz value e
// (ref_davidson/stream.rs:16-22).
__device__ __forceinline__ double qe_denominator(double h, double e) {
  const double x = 2.0 * (h - e);
  const double z = fabs(x - 1.0);
  return 0.25 * (2.0 * fmax(x, 1.0) + 1.0 / (sqrt(1.0 + z * z) + z));
}

// Blob from FP32 master parameters.
void quantize(hipStream_t s, const float* theta, float* blob);

// Davidson corrections with the network: for each active column j,
// r = HX_j - theta_j X_j, e = -K(g, T_j, theta_j, vbar) r / d_QE(diag, theta_j),
// where T_j is the column's kinetic scale and diag the Davidson diagonal.
// Writes E (n x k) and |e|^2, |r|^2 per column.
// `work` holds mlp_correct_work(n, k) doubles.
int64_t correct_work(int64_t n, int k);
void correct(hipStream_t s, int64_t n, int k, const double* X, const double* HX, const double* theta_ritz,
             const double* kinetic, const double* diag, double vbar, const float* blob, double* E,
             double* norm2_e, double* norm2_r, double* work);

// Training on a set of bands, each a residual r and target t (float, n each)
// with scalars (T, theta, vbar), sharing the row energies g (float, n).
struct TrainSet {
  int64_t n = 0;
  int bands = 0;
  const float* g = nullptr;      // n
  const float* R = nullptr;      // bands x n
  const float* Tg = nullptr;     // bands x n
  const float* scal = nullptr;   // bands x 3: T, theta, vbar
};
int64_t train_work(int64_t n, int batch);
// Mean band loss and its gradient (FP32, kCount) over the bands listed in
// `batch_dev` (device ints), for the network in `blob`.
void loss_grad(hipStream_t s, const TrainSet& ts, const int* batch_dev, int batch, const float* blob, float* grad,
               double* loss, double* work);
// One Adam step on the FP32 masters (Kingma & Ba 2015).
void adam(hipStream_t s, float* theta, const float* grad, float* m, float* v, int step, float lr);
// Forward only: K for one band's rows (tests).
void forward_k(hipStream_t s, int64_t n, const float* g, float T, float theta_ritz, float vbar, const float* blob,
               float* K);

}  // namespace mlp
}  // namespace bd
