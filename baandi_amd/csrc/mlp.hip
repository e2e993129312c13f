// Learned Davidson preconditioner on gfx950: inference fused into the
// correction step, and the training kernels (forward, per-band loss
// coefficients, backward, ordered gradient reduction, Adam).
//
// One wave evaluates the network for 16 rows at a time. Lane l works on row
// e = l % 16 and on the hidden units 16*b + 4*(l/16) + i (b, i < 4): that is
// the B-operand layout of v_mfma_f32_16x16x16_bf16 for the 64 x 64 layer, and
// its D layout for the outputs, so h1 feeds the MFMA straight from registers
// and the layer-2 outputs land where the lane needs them. W2 is read as bf16
// from LDS; sums are FP32 (layers 1 and 3 on the VALU).
//
// Backward: delta2 (D layout) is the B operand of W2^T delta2; the weight
// gradient dW2 = sum_rows delta2 h1^T is an MFMA over the 16 rows of a group
// after a 16 x 16 transpose of delta2 and h1 through LDS. Gradients are
// accumulated per wave over a fixed set of groups and reduced in a fixed
// order, so a training step is bitwise deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"
#include "mlp.hpp"

namespace bd {
namespace mlp {

namespace {
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Blob header offsets (floats).
constexpr int hW1 = 0, hb1 = 256, hb2 = 320, hw3 = 384, hb3 = 448;

__device__ __forceinline__ unsigned short bf16_bits(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(u >> 16);
}
__device__ __forceinline__ float bf16_value(float f) { return __uint_as_float(unsigned(bf16_bits(f)) << 16); }

// tanh(x) = 1 - 2 / (1 + e^{2x}) with the hardware reciprocal (1 ulp, no
// division sequence); saturates to +-1 for large |x| (rcp(inf) = 0).
__device__ __forceinline__ float tanh_f(float x) {
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + __expf(2.0f * x));
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

struct Group {
  float h1[16];  // unit 16*kb + 4*q + i at [kb*4 + i]
  float h2[16];  // unit 16*rb + 4*q + i at [rb*4 + i]
  bf16x4 hb[4];  // bf16(h1), the MFMA B operand
  float o;       // network output of the lane's row (same in all four quads)
};

// Network output for the lane's row from features u (sb: the blob in LDS).
__device__ __forceinline__ void forward_group(const float* sb, float u0, float u1, float u2, float u3, Group& G) {
  const int lane = threadIdx.x & 63, q = lane >> 4, e = lane & 15;
  const unsigned short* W2 = reinterpret_cast<const unsigned short*>(sb + kBlobW2);
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int n = 16 * kb + 4 * q + i;
      const float z = fmaf(sb[hW1 + 4 * n], u0,
                           fmaf(sb[hW1 + 4 * n + 1], u1, fmaf(sb[hW1 + 4 * n + 2], u2,
                                                              fmaf(sb[hW1 + 4 * n + 3], u3, sb[hb1 + n]))));
      const float h = tanh_f(z);
      G.h1[kb * 4 + i] = h;
      G.hb[kb][i] = static_cast<short>(bf16_bits(h));
    }
  float part = 0.0f;
#pragma unroll
  for (int rb = 0; rb < 4; ++rb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(&W2[(16 * rb + e) * 64 + 16 * kb + 4 * q]);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, G.hb[kb], acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int n = 16 * rb + 4 * q + i;
      const float h2 = tanh_f(acc[i] + sb[hb2 + n]);
      G.h2[rb * 4 + i] = h2;
      part = fmaf(sb[hw3 + n], h2, part);
    }
  }
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  G.o = part + sb[hb3];
}

// The first `count` floats of the blob into LDS: inference reads W1, the
// biases, w3 and W2 (kBlobW2T floats); training also reads W2^T.
template <int count = kBlobFloats>
__device__ __forceinline__ void load_blob(float* sb, const float* blob) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) sb[i] = blob[i];
  __syncthreads();
}

__device__ __forceinline__ float band_u1(float T) { return logf(T); }

// ---------------------------------------------------------------------------
// Parameters -> inference blob.
// ---------------------------------------------------------------------------
__global__ void k_quantize(const float* __restrict__ th, float* __restrict__ blob) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned short* W2 = reinterpret_cast<unsigned short*>(blob + kBlobW2);
  unsigned short* W2T = reinterpret_cast<unsigned short*>(blob + kBlobW2T);
  if (t < kHidden * kHidden) {
    const int out = t / kHidden, in = t % kHidden;
    const unsigned short b = bf16_bits(th[oW2 + t]);
    W2[out * kHidden + in] = b;
    W2T[in * kHidden + out] = b;
  }
  if (t < 512) {
    float v = 0.0f;
    if (t < 256) v = bf16_value(th[oW1 + t]);
    else if (t < 320) v = th[ob1 + t - 256];
    else if (t < 384) v = th[ob2 + t - 320];
    else if (t < 448) v = bf16_value(th[ow3 + t - 384]);
    else if (t == 448) v = th[ob3];
    blob[t] = v;
  }
}

// ---------------------------------------------------------------------------
// Inference inside the Davidson correction step.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_kinetic_scale(const double* __restrict__ X, int64_t n,
                                                       const double* __restrict__ kinetic, double* __restrict__ T) {
  __shared__ double red[2][4];
  const int j = blockIdx.x;
  const double* x = X + int64_t(j) * n;
  double a = 0.0, b = 0.0;
  const int64_t S = blockDim.x;
  int64_t i = threadIdx.x;
  // Four rows a pass with their loads issued first; the sums run in row
  // order, as in the one-row loop below.
  for (; i + 3 * S < n; i += 4 * S) {
    double xv[4], kv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xv[u] = x[i + u * S];
      kv[u] = kinetic[i + u * S];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double w = xv[u] * xv[u];
      a += w * kv[u];
      b += w;
    }
  }
  for (; i < n; i += S) {
    const double w = x[i] * x[i];
    a += w * kinetic[i];
    b += w;
  }
  a = wave_sum_d(a);
  b = wave_sum_d(b);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = a;
    red[1][wave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double sa = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    const double sb = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    T[j] = fmax(sa / sb, double(kKineticFloor));
  }
}

__global__ __launch_bounds__(256) void k_correct(int64_t n, const double* __restrict__ X, const double* __restrict__ HX,
                                                 const double* __restrict__ theta_ritz, const double* __restrict__ Tb,
                                                 const double* __restrict__ kinetic,
                                                 const double* __restrict__ diag, double vbar,
                                                 const float* __restrict__ blob, double* __restrict__ E,
                                                 double* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) float sb[kBlobW2T];
  __shared__ double red[2][4];
  load_blob<kBlobW2T>(sb, blob);
  const int j = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, e = lane & 15;
  const double th = theta_ritz[j];
  const float T = float(Tb[j]);
  const float u1 = band_u1(T), u2 = float((th - vbar) / Tb[j]), u3 = float(vbar);
  const double* x = X + int64_t(j) * n;
  const double* hx = HX + int64_t(j) * n;
  double* ej = E + int64_t(j) * n;
  double se = 0.0, sr = 0.0;
  const int64_t stride = int64_t(gridDim.x) * 64;
  int64_t row = int64_t(blockIdx.x) * 64 + 16 * wave + e;
  float g_next = row < n ? float(kinetic[row]) : 0.0f;
  for (int64_t base = int64_t(blockIdx.x) * 64; base < n; base += stride, row += stride) {
    const bool ok = row < n;
    const float g = g_next;
    // This pass's operands and the next pass's row energy are loaded before
    // the network runs, so their latency hides behind it.
    double xr = 0.0, hr = 0.0, dr = 0.0;
    if (ok && q == 0) {
      xr = x[row];
      hr = hx[row];
      dr = diag[row];
    }
    const int64_t next = row + stride;
    g_next = next < n ? float(kinetic[next]) : 0.0f;
    Group G;
    forward_group(sb, log1pf(g / T), u1, u2, u3, G);
    if (ok && q == 0) {
      const double r = hr - th * xr;
      const double v = -double(expf(G.o)) * r / qe_denominator(dr, th);
      ej[row] = v;
      se += v * v;
      sr += r * r;
    }
  }
  se = wave_sum_d(se);
  sr = wave_sum_d(sr);
  if (lane == 0) {
    red[0][wave] = se;
    red[1][wave] = sr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double* p = part + (int64_t(j) * gridDim.x + blockIdx.x) * 2;
    p[0] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    p[1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

__global__ void k_reduce_norms(int k, int nblk, const double* __restrict__ part, double* __restrict__ n2e,
                               double* __restrict__ n2r) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  double a = 0.0, b = 0.0;
  for (int p = 0; p < nblk; ++p) {
    a += part[(int64_t(j) * nblk + p) * 2];
    b += part[(int64_t(j) * nblk + p) * 2 + 1];
  }
  n2e[j] = a;
  n2r[j] = b;
}

__global__ __launch_bounds__(256) void k_forward_k(int64_t n, const float* __restrict__ g, float T, float ritz,
                                                   float vbar, const float* __restrict__ blob, float* __restrict__ K) {
  __shared__ __attribute__((aligned(16))) float sb[kBlobW2T];
  load_blob<kBlobW2T>(sb, blob);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, e = lane & 15;
  const float u1 = band_u1(T), u2 = (ritz - vbar) / T, u3 = vbar;
  for (int64_t base = int64_t(blockIdx.x) * 64; base < n; base += int64_t(gridDim.x) * 64) {
    const int64_t row = base + 16 * wave + e;
    const bool ok = row < n;
    Group G;
    forward_group(sb, log1pf((ok ? g[row] : 0.0f) / T), u1, u2, u3, G);
    if (ok && q == 0) K[row] = expf(G.o);
  }
}

// ---------------------------------------------------------------------------
// Training.
// ---------------------------------------------------------------------------
struct BandIn {
  float u1, u2, u3, T;
};
__device__ __forceinline__ BandIn band_in(const float* scal, int b) {
  BandIn B;
  B.T = scal[3 * b];
  B.u1 = band_u1(B.T);
  B.u2 = (scal[3 * b + 1] - scal[3 * b + 2]) / B.T;
  B.u3 = scal[3 * b + 2];
  return B;
}

// Per band of the batch and block: sum a t, sum a a, sum t t (a = K r).
__global__ __launch_bounds__(256) void k_train_fwd(TrainSet ts, const int* __restrict__ batch,
                                                   const float* __restrict__ blob, double* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) float sb[kBlobFloats];
  __shared__ double red[3][4];
  load_blob(sb, blob);
  const int bi = blockIdx.y, b = batch[bi];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, e = lane & 15;
  const BandIn B = band_in(ts.scal, b);
  const float* R = ts.R + int64_t(b) * ts.n;
  const float* Tg = ts.Tg + int64_t(b) * ts.n;
  double c = 0.0, aa = 0.0, tt = 0.0;
  for (int64_t base = int64_t(blockIdx.x) * 64; base < ts.n; base += int64_t(gridDim.x) * 64) {
    const int64_t row = base + 16 * wave + e;
    const bool ok = row < ts.n;
    Group G;
    forward_group(sb, log1pf((ok ? ts.g[row] : 0.0f) / B.T), B.u1, B.u2, B.u3, G);
    if (ok && q == 0) {
      const double a = double(expf(G.o)) * double(R[row]);
      const double t = Tg[row];
      c += a * t;
      aa += a * a;
      tt += t * t;
    }
  }
  c = wave_sum_d(c);
  aa = wave_sum_d(aa);
  tt = wave_sum_d(tt);
  if (lane == 0) {
    red[0][wave] = c;
    red[1][wave] = aa;
    red[2][wave] = tt;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int s = threadIdx.x;
    part[(int64_t(bi) * gridDim.x + blockIdx.x) * 3 + s] = red[s][0] + red[s][1] + red[s][2] + red[s][3];
  }
}

// Loss and dL/da = alpha t + beta a per batch band (mean over the batch).
__global__ void k_band_coef(int batch, int nblk, const double* __restrict__ part, double* __restrict__ coef,
                            double* __restrict__ loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double total = 0.0;
  for (int bi = 0; bi < batch; ++bi) {
    double c = 0.0, A = 0.0, B = 0.0;
    for (int p = 0; p < nblk; ++p) {
      c += part[(int64_t(bi) * nblk + p) * 3];
      A += part[(int64_t(bi) * nblk + p) * 3 + 1];
      B += part[(int64_t(bi) * nblk + p) * 3 + 2];
    }
    double L = 1.0, alpha = 0.0, beta = 0.0;
    if (A > 0.0 && B > 0.0) {
      L = 1.0 - c * c / (A * B);
      alpha = -2.0 * c / (A * B) / batch;
      beta = 2.0 * c * c / (A * A * B) / batch;
    }
    coef[2 * bi] = alpha;
    coef[2 * bi + 1] = beta;
    total += L;
  }
  loss[0] = total / batch;
}

// Backward over all rows of the batch; per-wave gradient partials (kCount floats each).
__global__ __launch_bounds__(256) void k_train_bwd(TrainSet ts, const int* __restrict__ batch, int nbatch,
                                                   const float* __restrict__ blob, const double* __restrict__ coef,
                                                   float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) float sb[kBlobFloats];
  __shared__ __attribute__((aligned(16))) unsigned short d2T[4][64 * 16];  // per wave: delta2[unit][row]
  __shared__ __attribute__((aligned(16))) unsigned short h1T[4][64 * 16];  // per wave: bf16(h1)[unit][row]
  load_blob(sb, blob);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, e = lane & 15;
  const unsigned short* W2T = reinterpret_cast<const unsigned short*>(sb + kBlobW2T);
  const int64_t gpb = (ts.n + 15) / 16;  // row groups per band
  const int64_t total = gpb * nbatch;
  const int64_t per_iter = int64_t(gridDim.x) * 4;
  const int64_t iters = (total + per_iter - 1) / per_iter;
  f32x4 dW2[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) dW2[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  float dW1[16][4], db1[16], db2[16], dw3[16], db3 = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    dW1[k][0] = dW1[k][1] = dW1[k][2] = dW1[k][3] = 0.0f;
    db1[k] = db2[k] = dw3[k] = 0.0f;
  }
  for (int64_t it = 0; it < iters; ++it) {
    const int64_t gidx = (it * gridDim.x + blockIdx.x) * 4 + wave;
    const bool valid = gidx < total;
    const int bi = valid ? int(gidx / gpb) : 0;
    const int64_t row = (valid ? gidx % gpb : 0) * 16 + e;
    const bool ok = valid && row < ts.n;
    const int b = batch[bi];
    const BandIn B = band_in(ts.scal, b);
    const float u0 = log1pf((ok ? ts.g[row] : 0.0f) / B.T);
    Group G;
    forward_group(sb, u0, B.u1, B.u2, B.u3, G);
    float dLdo = 0.0f;
    if (ok) {
      const float K = expf(G.o);
      const float r = ts.R[int64_t(b) * ts.n + row], t = ts.Tg[int64_t(b) * ts.n + row];
      const double a = double(K) * double(r);
      dLdo = float((coef[2 * bi] * t + coef[2 * bi + 1] * a) * double(r) * double(K));
    }
    if (q == 0) db3 += dLdo;
    bf16x4 d2b[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int n = 16 * rb + 4 * q + i;
        const float h2 = G.h2[rb * 4 + i];
        dw3[rb * 4 + i] = fmaf(dLdo, h2, dw3[rb * 4 + i]);
        const float d2 = dLdo * sb[hw3 + n] * (1.0f - h2 * h2);
        db2[rb * 4 + i] += d2;
        const unsigned short bits = bf16_bits(d2);
        d2b[rb][i] = static_cast<short>(bits);
        d2T[wave][n * 16 + e] = bits;
      }
    // delta1 = (W2^T delta2) * (1 - h1^2).
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(&W2T[(16 * rb + e) * 64 + 16 * kb + 4 * q]);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, d2b[kb], acc, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = rb * 4 + i;
        const float h1 = G.h1[k];
        const float d1 = acc[i] * (1.0f - h1 * h1);
        db1[k] += d1;
        dW1[k][0] = fmaf(d1, u0, dW1[k][0]);
        dW1[k][1] = fmaf(d1, B.u1, dW1[k][1]);
        dW1[k][2] = fmaf(d1, B.u2, dW1[k][2]);
        dW1[k][3] = fmaf(d1, B.u3, dW1[k][3]);
        h1T[wave][(16 * rb + 4 * q + i) * 16 + e] = static_cast<unsigned short>(G.hb[rb][i]);
      }
    }
    __syncthreads();
    // dW2[out][in] += sum_rows delta2[out][row] h1[in][row].
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(&d2T[wave][(16 * rb + e) * 16 + 4 * q]);
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        const bf16x4 bb = *reinterpret_cast<const bf16x4*>(&h1T[wave][(16 * cb + e) * 16 + 4 * q]);
        dW2[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, dW2[rb][cb], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float* P = part + (int64_t(blockIdx.x) * 4 + wave) * kCount;
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int i = 0; i < 4; ++i) P[oW2 + (16 * rb + 4 * q + i) * kHidden + 16 * cb + e] = dW2[rb][cb][i];
  // The other gradients: sum over the 16 rows held by lanes of the same quad.
#pragma unroll
  for (int k = 0; k < 16; ++k) {
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      dW1[k][0] += __shfl_xor(dW1[k][0], off, 64);
      dW1[k][1] += __shfl_xor(dW1[k][1], off, 64);
      dW1[k][2] += __shfl_xor(dW1[k][2], off, 64);
      dW1[k][3] += __shfl_xor(dW1[k][3], off, 64);
      db1[k] += __shfl_xor(db1[k], off, 64);
      db2[k] += __shfl_xor(db2[k], off, 64);
      dw3[k] += __shfl_xor(dw3[k], off, 64);
    }
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) db3 += __shfl_xor(db3, off, 64);
  if (e == 0) {
#pragma unroll
    for (int b4 = 0; b4 < 4; ++b4)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = b4 * 4 + i, n = 16 * b4 + 4 * q + i;
        P[oW1 + 4 * n] = dW1[k][0];
        P[oW1 + 4 * n + 1] = dW1[k][1];
        P[oW1 + 4 * n + 2] = dW1[k][2];
        P[oW1 + 4 * n + 3] = dW1[k][3];
        P[ob1 + n] = db1[k];
        P[ob2 + n] = db2[k];
        P[ow3 + n] = dw3[k];
      }
  }
  if (lane == 0) P[ob3] = db3;
}

__global__ void k_grad_reduce(int waves, const float* __restrict__ part, float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kCount) return;
  double s = 0.0;
  for (int w = 0; w < waves; ++w) s += part[int64_t(w) * kCount + p];
  grad[p] = float(s);
}

__global__ void k_adam(float* __restrict__ th, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float c1, float c2, float lr) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kCount) return;
  const float b1 = 0.9f, b2 = 0.999f;
  const float mm = b1 * m[p] + (1.0f - b1) * g[p];
  const float vv = b2 * v[p] + (1.0f - b2) * g[p] * g[p];
  m[p] = mm;
  v[p] = vv;
  th[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
}

constexpr int kCorrectBlocks = 128;
constexpr int kTrainFwdBlocks = 64;
constexpr int kTrainBwdBlocks = 128;
}  // namespace

void quantize(hipStream_t s, const float* theta, float* blob) {
  hipLaunchKernelGGL(k_quantize, dim3((kHidden * kHidden + 255) / 256), dim3(256), 0, s, theta, blob);
}

int64_t correct_work(int64_t n, int k) {
  const int64_t nblk = std::min<int64_t>((n + 63) / 64, kCorrectBlocks);
  return k + 2 * int64_t(k) * nblk;
}

void correct(hipStream_t s, int64_t n, int k, const double* X, const double* HX, const double* theta_ritz,
             const double* kinetic, const double* diag, double vbar, const float* blob, double* E,
             double* norm2_e, double* norm2_r, double* work) {
  if (k <= 0) return;
  // About 4096 workgroups in all (each loads the network once), at most
  // kCorrectBlocks per column and at least 16.
  const int per_column = std::min(kCorrectBlocks, std::max(16, 4096 / k));
  const int nblk = int(std::min<int64_t>((n + 63) / 64, per_column));
  double* Tb = work;
  double* part = work + k;
  hipLaunchKernelGGL(k_kinetic_scale, dim3(k), dim3(256), 0, s, X, n, kinetic, Tb);
  hipLaunchKernelGGL(k_correct, dim3(nblk, k), dim3(256), 0, s, n, X, HX, theta_ritz, Tb, kinetic, diag, vbar, blob,
                     E, part);
  hipLaunchKernelGGL(k_reduce_norms, dim3((k + 63) / 64), dim3(64), 0, s, k, nblk, part, norm2_e, norm2_r);
}

int64_t train_work(int64_t n, int batch) {
  (void)n;
  // fwd partials, coefficients, loss, then float gradient partials (as doubles).
  const int64_t fwd = 3 * int64_t(batch) * kTrainFwdBlocks;
  const int64_t bwd = (int64_t(kTrainBwdBlocks) * 4 * kCount + 1) / 2;
  return fwd + 2 * int64_t(batch) + 2 + bwd + 1;  // + the loss word
}

void loss_grad(hipStream_t s, const TrainSet& ts, const int* batch_dev, int batch, const float* blob, float* grad,
               double* loss, double* work) {
  if (batch <= 0 || ts.n <= 0) fail(BD_ERR_INVALID_PARAMETER, "empty training batch");
  double* fwd = work;
  double* coef = fwd + 3 * int64_t(batch) * kTrainFwdBlocks;
  float* gpart = reinterpret_cast<float*>(coef + 2 * int64_t(batch) + 2);
  hipLaunchKernelGGL(k_train_fwd, dim3(kTrainFwdBlocks, batch), dim3(256), 0, s, ts, batch_dev, blob, fwd);
  hipLaunchKernelGGL(k_band_coef, dim3(1), dim3(64), 0, s, batch, kTrainFwdBlocks, fwd, coef, loss);
  hipLaunchKernelGGL(k_train_bwd, dim3(kTrainBwdBlocks), dim3(256), 0, s, ts, batch_dev, batch, blob, coef, gpart);
  hipLaunchKernelGGL(k_grad_reduce, dim3((kCount + 255) / 256), dim3(256), 0, s, kTrainBwdBlocks * 4, gpart, grad);
}

void adam(hipStream_t s, float* theta, const float* grad, float* m, float* v, int step, float lr) {
  const float c1 = 1.0f - powf(0.9f, float(step));
  const float c2 = 1.0f - powf(0.999f, float(step));
  hipLaunchKernelGGL(k_adam, dim3((kCount + 255) / 256), dim3(256), 0, s, theta, grad, m, v, c1, c2, lr);
}

void forward_k(hipStream_t s, int64_t n, const float* g, float T, float theta_ritz, float vbar, const float* blob,
               float* K) {
  const int nblk = int(std::min<int64_t>((n + 63) / 64, 1024));
  hipLaunchKernelGGL(k_forward_k, dim3(nblk), dim3(256), 0, s, n, g, T, theta_ritz, vbar, blob, K);
}

}  // namespace mlp
}  // namespace bd

// ---------------------------------------------------------------------------
// C ABI.
// ---------------------------------------------------------------------------
namespace {
bd::mlp::TrainSet train_set(int64_t n, int32_t bands, const float* g, const float* R, const float* Tg,
                            const float* scal) {
  if (n <= 0 || bands <= 0 || !g || !R || !Tg || !scal) bd::fail(BD_ERR_INVALID_PARAMETER, "training set");
  bd::mlp::TrainSet ts;
  ts.n = n;
  ts.bands = bands;
  ts.g = g;
  ts.R = R;
  ts.Tg = Tg;
  ts.scal = scal;
  return ts;
}
void device_check(const char* what) {
  if (hipDeviceSynchronize() != hipSuccess || hipGetLastError() != hipSuccess)
    bd::fail(BD_ERR_DEVICE, std::string(what) + " failed on the device");
}
}  // namespace

extern "C" int32_t bd_mlp_param_count(void) { return bd::mlp::kCount; }
extern "C" int32_t bd_mlp_blob_floats(void) { return bd::mlp::kBlobFloats; }

extern "C" bd_status bd_mlp_quantize_dev(const float* theta, float* blob) {
  BD_GUARD({
    if (!theta || !blob) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    bd::mlp::quantize(nullptr, theta, blob);
    device_check("bd_mlp_quantize_dev");
  });
}

extern "C" bd_status bd_mlp_forward_dev(int64_t n, const float* g, float T, float ritz, float vbar, const float* blob,
                                        float* K) {
  BD_GUARD({
    if (n <= 0 || !g || !blob || !K || !(T > 0.0f)) bd::fail(BD_ERR_INVALID_PARAMETER, "mlp forward arguments");
    bd::mlp::forward_k(nullptr, n, g, T, ritz, vbar, blob, K);
    device_check("bd_mlp_forward_dev");
  });
}

extern "C" int64_t bd_mlp_train_workspace(int64_t n, int32_t batch) {
  return batch > 0 ? bd::mlp::train_work(n, batch) : 0;
}

extern "C" bd_status bd_mlp_loss_grad_dev(int64_t n, int32_t bands, const float* g, const float* R, const float* Tg,
                                          const float* scal, const int32_t* batch_dev, int32_t batch,
                                          const float* blob, float* grad, double* loss, double* work) {
  BD_GUARD({
    const bd::mlp::TrainSet ts = train_set(n, bands, g, R, Tg, scal);
    if (!batch_dev || batch <= 0 || !blob || !grad || !loss || !work)
      bd::fail(BD_ERR_INVALID_PARAMETER, "mlp loss arguments");
    double* dloss = work + bd::mlp::train_work(n, batch) - 1;
    bd::mlp::loss_grad(nullptr, ts, batch_dev, batch, blob, grad, dloss, work);
    if (hipMemcpy(loss, dloss, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      bd::fail(BD_ERR_DEVICE, "mlp loss readback");
    device_check("bd_mlp_loss_grad_dev");
  });
}

extern "C" bd_status bd_mlp_adam_dev(float* theta, const float* grad, float* m, float* v, int32_t step, float lr) {
  BD_GUARD({
    if (!theta || !grad || !m || !v || step < 1) bd::fail(BD_ERR_INVALID_PARAMETER, "adam arguments");
    bd::mlp::adam(nullptr, theta, grad, m, v, step, lr);
    device_check("bd_mlp_adam_dev");
  });
}

extern "C" bd_status bd_mlp_train_dev(int64_t n, int32_t bands, const float* g, const float* R, const float* Tg,
                                      const float* scal, float* theta, float* m, float* v, int32_t steps,
                                      int32_t batch, float lr, uint64_t seed, double* losses, int32_t* batch_dev,
                                      float* blob, float* grad, double* work) {
  BD_GUARD({
    const bd::mlp::TrainSet ts = train_set(n, bands, g, R, Tg, scal);
    if (!theta || !m || !v || steps < 0 || batch <= 0 || batch > bands || !losses || !batch_dev || !blob || !grad ||
        !work)
      bd::fail(BD_ERR_INVALID_PARAMETER, "mlp training arguments");
    double* dloss = work + bd::mlp::train_work(n, batch) - 1;
    std::vector<int32_t> idx(batch);
    uint64_t state = seed * 6364136223846793005ull + 1442695040888963407ull;
    for (int step = 1; step <= steps; ++step) {
      // Partial Fisher-Yates draw of `batch` distinct bands (64-bit LCG).
      std::vector<int32_t> pool(bands);
      for (int i = 0; i < bands; ++i) pool[i] = i;
      for (int i = 0; i < batch; ++i) {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        const int j = i + int((state >> 33) % uint64_t(bands - i));
        std::swap(pool[i], pool[j]);
        idx[i] = pool[i];
      }
      if (hipMemcpy(batch_dev, idx.data(), sizeof(int32_t) * batch, hipMemcpyHostToDevice) != hipSuccess)
        bd::fail(BD_ERR_DEVICE, "batch upload");
      bd::mlp::quantize(nullptr, theta, blob);
      bd::mlp::loss_grad(nullptr, ts, batch_dev, batch, blob, grad, dloss, work);
      bd::mlp::adam(nullptr, theta, grad, m, v, step, lr);
      if (hipMemcpy(&losses[step - 1], dloss, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        bd::fail(BD_ERR_DEVICE, "loss readback");
    }
    bd::mlp::quantize(nullptr, theta, blob);
    device_check("bd_mlp_train_dev");
  });
}
