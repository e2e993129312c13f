// FP64 symmetric eigensolver for the Rayleigh-Ritz step (the role of
// RealEighWorkspace::solve_into, crates/dft-numerics/src/real.rs:2286-2330,
// called from ref_davidson/workspace.rs:435-499). Only the lowest `want`
// pairs are formed, which is all the Davidson step reads.
//
//  1. Householder tridiagonalisation (Golub & Van Loan, Matrix Computations,
//     4th ed., Alg. 8.3.1) in one 1024-thread workgroup: the matrix stays in
//     the L2 of one XCD, and each step updates the trailing block and
//     multiplies it by the next reflector in a single pass.
//  2. Sturm-sequence bisection (Barth, Martin & Wilkinson 1967), one wave per
//     eigenvalue with 64-point multisection per step.
//  3. Inverse iteration with a partially pivoted tridiagonal LU (the scheme
//     of LAPACK dlagtf/dlagts/dstein), one lane per eigenvector.
//  4. Two passes of classical Gram-Schmidt inside clusters of close
//     eigenvalues (gap <= 1e-3 ||T||, the dstein criterion), in ascending
//     order, with a fresh inverse-iteration solve when a vector loses most of
//     its norm.
//  5. Back-transformation by the stored reflectors, one wave per vector.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <set>
#include <string>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <string>

#include "common.hpp"
#include "eigen.hpp"
#include "gemm.hpp"
#include "wave.hpp"

namespace bd {

namespace {

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

constexpr int kTT = 1024;      // tridiagonalisation workgroup
constexpr int kTW = kTT / 64;  // its waves
constexpr int kLC = 2;         // columns per wave pass of the lower-triangle update

// Sum over a kTT-thread block, returned to every thread (fixed order).
// `red` holds kTW doubles of LDS; ends with a barrier so it can be reused.
__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kTW; ++i) t += red[i];
  __syncthreads();
  return t;
}

// block_sum that also reads *peek into *peeked between its two barriers (after
// every thread's writes before the call, before any write after it).
__device__ double block_sum_peek(double v, double* red, const double* peek, double* peeked) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kTW; ++i) t += red[i];
  *peeked = *peek;
  __syncthreads();
  return t;
}

// Deterministic start vector entries in (-1, 1).
__device__ __forceinline__ double hash_unit(uint32_t i, uint32_t j, uint32_t salt) {
  uint32_t h = i * 0x9E3779B1u ^ (j * 0x85EBCA77u + salt * 0xC2B2AE3Du + 0x27D4EB2Fu);
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return (double(h) + 0.5) * (2.0 / 4294967296.0) - 1.0;
}

// ---------------------------------------------------------------------------
// 1. Tridiagonalisation. Reflector k is H_k = I - tau_k x x^T with x stored in
// A[k+1:, k] (x[0] = a - alpha, unnormalised); d, e the tridiagonal
// (e[k] couples d[k] and d[k+1]).
// ---------------------------------------------------------------------------
// LOWER: after the first matvec only the lower triangle of the trailing
// block is read and written (half the L2 traffic); the matvec's row part
// B[j, j'] x_j' (j' < j) is accumulated per wave in LDS and summed after the
// pass. Needs kTW more vectors of m in LDS.
template <bool LOWER>
__global__ __launch_bounds__(kTT) void k_tridiag(int m, double* __restrict__ A, int64_t lda, double* __restrict__ d,
                                                 double* __restrict__ e, double* __restrict__ tau) {
  // LDS: two reflector vectors, w, q = B x, the next step's first column,
  // two reduction slots per wave and (LOWER) a row accumulator per wave.
  extern __shared__ double sh[];
  double* xa = sh;
  double* xb = sh + m;
  double* w = sh + 2 * m;
  double* q = sh + 3 * m;
  double* nc = sh + 4 * m;
  double* red = sh + 5 * m;
  double* red2 = red + kTW;
  double* pw = red2 + kTW;  // LOWER: kTW x m
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  auto a = [&](int i, int j) -> double& { return A[i + int64_t(j) * lda]; };
  if (m < 3) {
    if (tid == 0) {
      d[0] = a(0, 0);
      tau[0] = 0.0;
      if (m == 2) {
        e[0] = a(1, 0);
        d[1] = a(1, 1);
        tau[1] = 0.0;
      }
    }
    return;
  }
  // Reflector of column k from x (its rows k+1.., in LDS) and sum(x^2);
  // returns tau_k (0 when the column is already zero).
  auto reflect = [&](double ss, double* xv, int k) -> double {
    const double sigma = sqrt(ss);
    const double x0 = xv[0];
    double alpha = 0.0, tk = 0.0;
    if (sigma > 0.0) {
      alpha = x0 >= 0.0 ? -sigma : sigma;
      tk = 1.0 / (sigma * (sigma + fabs(x0)));  // 2 / |x - alpha e1|^2
    }
    __syncthreads();
    if (tid == 0) {
      if (sigma > 0.0) {
        xv[0] = x0 - alpha;
        a(k + 1, k) = x0 - alpha;
      }
      e[k] = alpha;
      tau[k] = tk;
    }
    __syncthreads();
    return tk;
  };

  // The same with x0 = xv[0] already read by every thread (block_sum_peek),
  // so only the write needs a barrier.
  auto reflect_read = [&](double ss, double x0, double* xv, int k) -> double {
    const double sigma = sqrt(ss);
    double alpha = 0.0, tk = 0.0;
    if (sigma > 0.0) {
      alpha = x0 >= 0.0 ? -sigma : sigma;
      tk = 1.0 / (sigma * (sigma + fabs(x0)));
    }
    if (tid == 0) {
      if (sigma > 0.0) {
        xv[0] = x0 - alpha;
        a(k + 1, k) = x0 - alpha;
      }
      e[k] = alpha;
      tau[k] = tk;
    }
    __syncthreads();
    return tk;
  };

  double* x = xa;
  double* xn = xb;
  int L = m - 1;
  double qx_carry = 0.0;  // LOWER: q'.x' of the step before, in every thread
  double ss = 0.0;
  for (int r = tid; r < L; r += kTT) {
    const double v = a(1 + r, 0);
    x[r] = v;
    ss += v * v;
  }
  if (tid == 0) d[0] = a(0, 0);
  ss = block_sum(ss, red);
  double t = reflect(ss, x, 0);
  // q = B_0 x (B_0 = A[1:, 1:], four columns per wave, lanes along rows);
  // the first column of B_0 is kept in nc; q.x is summed on the way.
  {
    double qx = 0.0;
    for (int j0 = 4 * wave; j0 < L; j0 += 4 * kTW) {
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      const double* col[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) col[c] = &a(1, 1 + (j0 + c < L ? j0 + c : j0));
      for (int r = lane; r < L; r += 64) {
        const double xr = x[r];
        double b[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) b[c] = col[c][r];
        if (j0 == 0) nc[r] = b[0];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] += b[c] * xr;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double v = wave_sum(acc[c]);
        if (j0 + c < L) {
          if (lane == 0) q[j0 + c] = v;
          qx += v * x[j0 + c];
        }
      }
    }
    if (lane == 0) red2[wave] = qx;
  }
  if (LOWER)
    for (int i = tid; i < kTW * m; i += kTT) pw[i] = 0.0;
  __syncthreads();

  for (int k = 0; k + 2 < m; ++k) {
    L = m - k - 1;
    const int c0 = k + 1;  // global index of the trailing block's first row/column
    double qx = 0.0;
    if (LOWER && k > 0) {
      qx = qx_carry;
    } else {
#pragma unroll
      for (int i = 0; i < kTW; ++i) qx += red2[i];
    }
    // p = t q, K = t/2 p.x = t^2/2 q.x, w = p - K x.
    const double K = 0.5 * t * t * qx;
    for (int r = tid; r < L; r += kTT) w[r] = t * q[r] - K * x[r];
    __syncthreads();
    if (L == 2) {
      if (tid == 0) {
        d[m - 2] = nc[0] - 2.0 * x[0] * w[0];
        e[m - 2] = nc[1] - (x[1] * w[0] + w[1] * x[0]);
        d[m - 1] = a(c0 + 1, c0 + 1) - 2.0 * x[1] * w[1];
        tau[m - 2] = 0.0;
        tau[m - 1] = 0.0;
      }
      break;
    }
    // First column of B_k (held in nc): rows 1.. become the next reflector.
    const double x0 = x[0], w0 = w[0];
    ss = 0.0;
    for (int r = tid; r < L; r += kTT) {
      const double b = nc[r] - (x[r] * w0 + w[r] * x0);
      if (r == 0) {
        d[c0] = b;
      } else {
        xn[r - 1] = b;
        a(c0 + r, c0) = b;
        ss += b * b;
      }
    }
    double xn0 = 0.0;
    ss = block_sum_peek(ss, red, xn, &xn0);
    const double tn = reflect_read(ss, xn0, xn, c0);
    if constexpr (LOWER) {
      // Rest of B_k, lower triangle (rows r >= j of column j): B -= x w^T +
      // w x^T and q' = B_{k+1} x', the column part sum_{r>=j} b xn_{r-1} per
      // column, the row part b xn_{j-1} into the wave's row accumulator.
      double* pwv = pw + int64_t(wave) * m;
      for (int j0 = 1 + kLC * wave; j0 < L; j0 += kLC * kTW) {
        double acc[kLC], xj[kLC], wj[kLC], xnj[kLC];
        double* col[kLC];
#pragma unroll
        for (int c = 0; c < kLC; ++c) {
          acc[c] = 0.0;
          const bool ok = j0 + c < L;
          xj[c] = ok ? x[j0 + c] : 0.0;
          wj[c] = ok ? w[j0 + c] : 0.0;
          xnj[c] = ok ? xn[j0 + c - 1] : 0.0;
          col[c] = &a(c0, c0 + (ok ? j0 + c : j0));
        }
        for (int r = j0 + lane; r < L; r += 64) {
          const double xr = x[r], wr = w[r], xnr = xn[r - 1];
          double pr = 0.0;
#pragma unroll
          for (int c = 0; c < kLC; ++c) {
            if (j0 + c < L && r >= j0 + c) {
              const double b = col[c][r] - (xr * wj[c] + wr * xj[c]);
              col[c][r] = b;
              acc[c] += b * xnr;
              if (r > j0 + c) pr += b * xnj[c];
              if (j0 + c == 1) nc[r - 1] = b;
            }
          }
          pwv[r - 1] += pr;
        }
#pragma unroll
        for (int c = 0; c < kLC; ++c) {
          const double v = wave_sum(acc[c]);
          if (j0 + c < L && lane == 0) q[j0 + c - 1] = v;
        }
      }
      __syncthreads();
      // q'_i += the row parts of every wave (waves in fixed order); q'.x'.
      double qxs = 0.0;
      for (int i = tid; i < L - 1; i += kTT) {
        double sacc = q[i];
        for (int ww = 0; ww < kTW; ++ww) {
          sacc += pw[int64_t(ww) * m + i];
          pw[int64_t(ww) * m + i] = 0.0;
        }
        q[i] = sacc;
        qxs += sacc * xn[i];
      }
      qx_carry = block_sum(qxs, red);
    } else {
      // Rest of B_k: B -= x w^T + w x^T and q' = B_{k+1} x' in one pass, four
      // columns per wave (four loads in flight per lane). Local column 1 is the
      // next step's first column: it goes to nc as well.
      qx = 0.0;
      for (int j0 = 1 + 4 * wave; j0 < L; j0 += 4 * kTW) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0}, xj[4], wj[4];
        double* col[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const bool ok = j0 + c < L;
          xj[c] = ok ? x[j0 + c] : 0.0;
          wj[c] = ok ? w[j0 + c] : 0.0;
          col[c] = &a(c0, c0 + (ok ? j0 + c : j0));
        }
        for (int r = 1 + lane; r < L; r += 64) {
          const double xr = x[r], wr = w[r], xnr = xn[r - 1];
          double b[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) b[c] = col[c][r];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            if (j0 + c < L) {
              b[c] -= xr * wj[c] + wr * xj[c];
              col[c][r] = b[c];
              acc[c] += b[c] * xnr;
            }
          }
          if (j0 == 1) nc[r - 1] = b[0];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const double v = wave_sum(acc[c]);
          if (j0 + c < L) {
            if (lane == 0) q[j0 + c - 1] = v;
            qx += v * xn[j0 + c - 1];
          }
        }
      }
      if (lane == 0) red2[wave] = qx;
      __syncthreads();
    }
    double* tmp = x;
    x = xn;
    xn = tmp;
    t = tn;
  }
}

// Blocked form (the panel scheme of LAPACK's dsytrd/dlatrd): reflectors are
// generated kPNB at a time against the matrix as it stood at the panel's
// start, A_cur = A - sum_{i<s} (v_i w_i^T + w_i v_i^T), with the panel's
// v_i, w_i held in LDS row by row (VW[row][i] = v_i, VW[row][kPNB + i] = w_i).
// Both triangles of the trailing block are kept, so the per-step product
// q = A_cur x is a plain column sweep with lanes along rows (no cross-lane
// sums); the global matrix is only read during a panel. The rank-2 kPNB
// update of the trailing block and the reflectors' write-back run once per
// panel; the block starts symmetric (lower triangle mirrored) and the
// update keeps it so bit for bit. Same outputs as k_tridiag: reflector k in A[k+1:, k]
// (x[0] = a - alpha), d, e, tau, with tau = 0 for the last two indices.
constexpr int kPNB = 8;             // reflectors per panel
constexpr int kPSW = 2 * kPNB + 2;  // VW row stride (doubles; 16-byte rows)
constexpr int kPSB = 16;            // columns per task of the product
constexpr int kPUC = 16;            // columns per tile of the panel update

// Upper triangle <- lower triangle (the panel kernel reads full columns).
__global__ void k_lower_to_upper(int m, double* __restrict__ A, int64_t lda) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < j && j < m) A[i + int64_t(j) * lda] = A[j + int64_t(i) * lda];
}

size_t tridiag_panel_lds(int m) { return size_t((kPSW + 5 + kTW) * m + 2 * kTW + 2 * kPNB) * 8; }

// Sum over the workgroup with one barrier: `slot` (kTW doubles) must not be
// written again before every thread has passed a later barrier (callers
// alternate two slots).
__device__ double block_sum1(double v, double* slot) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) slot[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kTW; ++i) t += slot[i];
  return t;
}

__global__ __launch_bounds__(kTT) void k_tridiag_panel(int m, double* __restrict__ A, int64_t lda,
                                                       double* __restrict__ d, double* __restrict__ e,
                                                       double* __restrict__ tau) {
  extern __shared__ __align__(16) double sh[];
  double* VW = sh;               // m rows of kPSW
  double* q = VW + kPSW * m;     // A_cur x of the current step
  double* dl = q + m;            // d, e, tau, written out at the end
  double* el = dl + m;
  double* tl = el + m;
  double* cb = tl + m;           // the current column below the diagonal
  double* pw = cb + m;           // kTW x m: per-wave row sums of the product
  double* red = pw + kTW * m;    // two reduction slots of kTW
  double* coef = red + 2 * kTW;  // 2 kPNB: w_i.x at i, v_i.x at kPNB + i
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  auto a = [&](int i, int j) -> double& { return A[i + int64_t(j) * lda]; };
  if (m < 3) {
    if (tid == 0) {
      d[0] = a(0, 0);
      tau[0] = 0.0;
      if (m == 2) {
        e[0] = a(1, 0);
        d[1] = a(1, 1);
        tau[1] = 0.0;
      }
    }
    return;
  }
  for (int i = tid; i < kTW * m; i += kTT) pw[i] = 0.0;
  __syncthreads();
  const int last = m - 2;  // reflector steps k = 0 .. m-3
  double* pwv = pw + int64_t(wv) * m;
  int slot = 0;
  for (int p0 = 0;;) {
    const int se = min(kPNB, last - p0);
    for (int s = 0; s < se; ++s) {
      const int k = p0 + s, c0 = k + 1, L = m - k - 1;
      // Column k of A_cur, rows k..m-1: d[k] and the raw reflector in cb.
      double ss = 0.0;
      const double* vk = VW + int64_t(k) * kPSW;
      for (int r = tid; r <= L; r += kTT) {
        double v = a(k + r, k);
        const double* vr = VW + int64_t(k + r) * kPSW;
        // Unrolled and branch-free (columns i >= s selected out), so the LDS
        // reads are issued together.
#pragma unroll
        for (int i = 0; i < kPNB; ++i) {
          const double t = vr[i] * vk[kPNB + i] + vr[kPNB + i] * vk[i];
          v -= i < s ? t : 0.0;
        }
        if (r == 0) {
          dl[k] = v;
        } else {
          cb[r - 1] = v;
          ss += v * v;
        }
      }
      ss = block_sum1(ss, red + slot * kTW);
      slot ^= 1;
      const double x0 = cb[0];
      const double sigma = sqrt(ss);
      double alpha = 0.0, tk = 0.0;
      if (sigma > 0.0) {
        alpha = x0 >= 0.0 ? -sigma : sigma;
        tk = 1.0 / (sigma * (sigma + fabs(x0)));  // 2 / |x - alpha e1|^2
      }
      const double x0m = sigma > 0.0 ? x0 - alpha : x0;  // x[0]; x[r] = cb[r] for r > 0
      if (tid == 0) {
        el[k] = alpha;
        tl[k] = tk;
      }
      // Row sums of A[c0:, c0:] x: tasks of 64 rows x kPSB columns dealt to
      // the waves in turn; a task's loads are issued together.
      const int nrc = (L + 63) / 64, ntask = nrc * ((L + kPSB - 1) / kPSB);
      for (int t = wv; t < ntask; t += kTW) {
        const int rc = t % nrc, j0 = (t / nrc) * kPSB;
        const int r = 64 * rc + lane, rr = min(r, L - 1);
        const double* base = A + c0 + rr + int64_t(c0) * lda;
        double bl[kPSB], xs[kPSB];
#pragma unroll
        for (int jj = 0; jj < kPSB; ++jj) bl[jj] = base[int64_t(min(j0 + jj, L - 1)) * lda];
#pragma unroll
        for (int jj = 0; jj < kPSB; ++jj) xs[jj] = cb[min(j0 + jj, L - 1)];
        double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
        for (int jj = 0; jj < kPSB; jj += 2) {
          const int j = j0 + jj;
          acc0 += bl[jj] * (j == 0 ? x0m : (j < L ? xs[jj] : 0.0));
          acc1 += bl[jj + 1] * (j + 1 < L ? xs[jj + 1] : 0.0);
        }
        if (r < L) pwv[r] += acc0 + acc1;
      }
      // w_i.x (t < s) and v_i.x (t >= s) of the panel's earlier steps.
      for (int t = wv; t < 2 * s; t += kTW) {
        const int cidx = t < s ? kPNB + t : t - s;  // w_t, else v_{t-s}
        double acc = 0.0;
        for (int r = lane; r < L; r += 64) acc += VW[int64_t(c0 + r) * kPSW + cidx] * (r == 0 ? x0m : cb[r]);
        acc = wave_sum(acc);
        if (lane == 0) coef[t < s ? t : kPNB + t - s] = acc;
      }
      __syncthreads();
      // q = y - sum_i (v_i (w_i.x) + w_i (v_i.x)); q.x.
      double qxs = 0.0;
      for (int i = tid; i < L; i += kTT) {
        double v = 0.0;
        for (int ww = 0; ww < kTW; ++ww) {
          v += pw[int64_t(ww) * m + i];
          pw[int64_t(ww) * m + i] = 0.0;
        }
        const double* vr = VW + int64_t(c0 + i) * kPSW;
#pragma unroll
        for (int i2 = 0; i2 < kPNB; ++i2) {
          const double t = vr[i2] * coef[i2] + vr[kPNB + i2] * coef[kPNB + i2];
          v -= i2 < s ? t : 0.0;
        }
        q[i] = v;
        qxs += v * (i == 0 ? x0m : cb[i]);
      }
      const double qx = block_sum1(qxs, red + slot * kTW);
      slot ^= 1;
      // v_s = x and w_s = tau q - (tau^2/2 q.x) x into the panel rows.
      const double K = 0.5 * tk * tk * qx;
      for (int i = tid; i < L; i += kTT) {
        const double xi = i == 0 ? x0m : cb[i];
        double* row = VW + int64_t(c0 + i) * kPSW;
        row[s] = xi;
        row[kPNB + s] = tk * q[i] - K * xi;
      }
      __syncthreads();
    }
    const int p1 = p0 + se;
    // The panel's reflectors to A[k+1:, k].
    for (int s = 0; s < se; ++s) {
      const int k = p0 + s;
      for (int r = k + 1 + tid; r < m; r += kTT) a(r, k) = VW[int64_t(r) * kPSW + s];
    }
    if (p1 >= last) {
      // Trailing 2 x 2 of A_cur.
      if (tid == 0) {
        double v00 = a(m - 2, m - 2), v10 = a(m - 1, m - 2), v11 = a(m - 1, m - 1);
        const double* r2 = VW + int64_t(m - 2) * kPSW;
        const double* r1 = VW + int64_t(m - 1) * kPSW;
        for (int i = 0; i < se; ++i) {
          v00 -= 2.0 * r2[i] * r2[kPNB + i];
          v10 -= r1[i] * r2[kPNB + i] + r1[kPNB + i] * r2[i];
          v11 -= 2.0 * r1[i] * r1[kPNB + i];
        }
        dl[m - 2] = v00;
        el[m - 2] = v10;
        dl[m - 1] = v11;
        tl[m - 2] = 0.0;
        tl[m - 1] = 0.0;
      }
      break;
    }
    // Trailing block A[p1:, p1:] -= sum_i (v_i w_i^T + w_i v_i^T), both
    // triangles: tiles of 64 rows (a lane per row) x kPUC columns dealt to
    // the waves in turn; a tile's loads are issued before its stores.
    {
      const int Lp = m - p1;
      const int nrc = (Lp + 63) / 64, ntile = nrc * ((Lp + kPUC - 1) / kPUC);
      for (int t = wv; t < ntile; t += kTW) {
        const int rc = t % nrc, J0 = p1 + (t / nrc) * kPUC;
        const int R = p1 + 64 * rc + lane, Rc = min(R, m - 1);
        double rv[2 * kPNB], val[kPUC];
        const double* vr = VW + int64_t(Rc) * kPSW;
#pragma unroll
        for (int i = 0; i < 2 * kPNB; ++i) rv[i] = vr[i];
#pragma unroll
        for (int jj = 0; jj < kPUC; ++jj) val[jj] = A[Rc + int64_t(min(J0 + jj, m - 1)) * lda];
#pragma unroll
        for (int jj = 0; jj < kPUC; ++jj) {
          const int J = J0 + jj;
          const double* vc = VW + int64_t(min(J, m - 1)) * kPSW;
          // Each product rounded on its own, so that (R, J) and (J, R) get
          // the same value: the trailing block stays exactly symmetric.
          double dlt = 0.0;
#pragma unroll
          for (int i = 0; i < kPNB; ++i)
            dlt = __dadd_rn(dlt, __dadd_rn(__dmul_rn(rv[i], vc[kPNB + i]), __dmul_rn(rv[kPNB + i], vc[i])));
          if (R < m && J < m) A[R + int64_t(J) * lda] = val[jj] - dlt;
        }
      }
    }
    __syncthreads();
    p0 = p1;
  }
  __syncthreads();
  for (int i = tid; i < m; i += kTT) {
    d[i] = dl[i];
    tau[i] = tl[i];
    if (i < m - 1) e[i] = el[i];
  }
}

// ---------------------------------------------------------------------------
// 1c. Tridiagonalisation across the chip, for subspaces whose trailing block
// one CU cannot stream fast enough (the Si512 subspace of 2252). The panel
// scheme of k_tridiag_panel with kMB reflectors per panel, one launch per
// reflector for its column (one workgroup: the corrections of the earlier
// panel reflectors, the reflector and, for the step before, w) and one for
// the product of the panel-start trailing block with it (a wave per column,
// every CU), then a tiled rank-2kMB update of the trailing block that writes
// both triangles from one value (the block stays exactly symmetric). V and W
// (m x kMB, ld m) hold the panel's v_i, w_i, zero above their first row; y
// the last product; sc[0] the last tau. Same outputs as k_tridiag.
// ---------------------------------------------------------------------------
constexpr int kMB = 16;  // reflectors per panel

// Finish w of step s - 1 (when s > 0): q = y - sum_{t < s-1} (v_t (w_t.v) +
// w_t (v_t.v)), w = tau q - (tau^2/2 q.v) v. Then by `mode`: 0, column k =
// p0 + s of A_cur = A - sum_{t < s} (v_t w_t^T + w_t v_t^T): d[k], and the
// reflector of its rows k+1.. into V[:, s], A[k+1:, k], e[k], tau[k]; 1,
// nothing more (the panel's last w); 2 (k = m - 2), the trailing 2 x 2 of
// A_cur: d[m-2], e[m-2], d[m-1].
__global__ __launch_bounds__(kTT) void k_td_column(int m, double* __restrict__ A, int64_t lda, int p0, int s,
                                                   int mode, double* __restrict__ V, double* __restrict__ W,
                                                   const double* __restrict__ y, double* __restrict__ sc,
                                                   double* __restrict__ d, double* __restrict__ e,
                                                   double* __restrict__ tau) {
  extern __shared__ double sh[];
  double* wl = sh;                // w of step s-1 (rows >= k)
  double* xl = sh + m;            // the new reflector
  double* coef = sh + 2 * m;      // 2 kMB: w_t.v at t, v_t.v at kMB + t
  double* rowk = coef + 2 * kMB;  // 2 kMB: V[k, t] at t, W[k, t] at kMB + t
  double* red = rowk + 2 * kMB;   // kTW
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = p0 + s;
  if (s > 0) {
    const int c0 = k;  // first trailing row of the previous step
    const double* v = V + int64_t(s - 1) * m;
    const double t1 = sc[0];
    for (int t = wave; t < 2 * (s - 1); t += kTW) {
      const double* a = t < s - 1 ? W + int64_t(t) * m : V + int64_t(t - (s - 1)) * m;
      double acc = 0.0;
      for (int r = c0 + lane; r < m; r += 64) acc += a[r] * v[r];
      acc = wave_sum(acc);
      if (lane == 0) coef[t < s - 1 ? t : kMB + t - (s - 1)] = acc;
    }
    __syncthreads();
    double qx = 0.0;
    for (int r = c0 + tid; r < m; r += kTT) {
      double q = y[r];
      for (int t = 0; t < s - 1; ++t) q -= V[r + int64_t(t) * m] * coef[t] + W[r + int64_t(t) * m] * coef[kMB + t];
      wl[r] = q;
      qx += q * v[r];
    }
    qx = block_sum(qx, red);
    const double K = 0.5 * t1 * t1 * qx;
    double* wg = W + int64_t(s - 1) * m;
    for (int r = tid; r < m; r += kTT) {
      const double w = r >= c0 ? t1 * wl[r] - K * v[r] : 0.0;
      wl[r] = w;
      wg[r] = w;
    }
    __syncthreads();
  }
  if (mode == 1) return;  // panel end: only w of the last step
  // Row k of the panel's V and W (t < s; t = s-1 from LDS).
  for (int t = tid; t < s; t += kTT) {
    rowk[t] = V[k + int64_t(t) * m];
    rowk[kMB + t] = t == s - 1 ? wl[k] : W[k + int64_t(t) * m];
  }
  __syncthreads();
  auto acur = [&](int r, int j) {  // A_cur[r, j] for r, j >= k (j == k or the final 2 x 2)
    double a = A[r + int64_t(j) * lda];
    for (int t = 0; t < s; ++t) {
      const double vr = V[r + int64_t(t) * m], vj = V[j + int64_t(t) * m];
      const double wr = t == s - 1 ? wl[r] : W[r + int64_t(t) * m];
      const double wj = t == s - 1 ? wl[j] : W[j + int64_t(t) * m];
      a -= vr * wj + wr * vj;
    }
    return a;
  };
  if (mode == 2) {
    if (tid == 0) {
      d[m - 2] = acur(m - 2, m - 2);
      e[m - 2] = acur(m - 1, m - 2);
      d[m - 1] = acur(m - 1, m - 1);
      tau[m - 2] = 0.0;
      tau[m - 1] = 0.0;
    }
    return;
  }
  double ss = 0.0;
  for (int r = k + tid; r < m; r += kTT) {
    double a = A[r + int64_t(k) * lda];
    for (int t = 0; t < s; ++t) {
      const double wr = t == s - 1 ? wl[r] : W[r + int64_t(t) * m];
      a -= V[r + int64_t(t) * m] * rowk[kMB + t] + wr * rowk[t];
    }
    if (r == k) {
      d[k] = a;
    } else {
      xl[r] = a;
      ss += a * a;
    }
  }
  ss = block_sum(ss, red);
  const double x0 = xl[k + 1];
  const double sigma = sqrt(ss);
  double alpha = 0.0, tk = 0.0;
  if (sigma > 0.0) {
    alpha = x0 >= 0.0 ? -sigma : sigma;
    tk = 1.0 / (sigma * (sigma + fabs(x0)));  // 2 / |x - alpha e1|^2
  }
  double* vs = V + int64_t(s) * m;
  for (int r = tid; r < m; r += kTT) {
    double x = 0.0;
    if (r > k) x = (r == k + 1 && sigma > 0.0) ? x0 - alpha : xl[r];
    vs[r] = x;
    if (r > k) A[r + int64_t(k) * lda] = x;
  }
  if (tid == 0) {
    e[k] = alpha;
    tau[k] = tk;
    sc[0] = tk;
  }
}

// y[r] = sum_{j > k} A[j, r] v[j] for rows r > k (column r of the symmetric
// panel-start block against v, a wave per column).
__global__ __launch_bounds__(256) void k_td_symv(int m, const double* __restrict__ A, int64_t lda, int k,
                                                 const double* __restrict__ v, double* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int r = k + 1 + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= m) return;
  const double* col = A + int64_t(r) * lda;
  double acc = 0.0;
  for (int j = k + 1 + lane; j < m; j += 64) acc += col[j] * v[j];
  acc = wave_sum(acc);
  if (lane == 0) y[r] = acc;
}

// A[p1:, p1:] -= sum_{t < se} (v_t w_t^T + w_t v_t^T): 64 x 64 tiles of the
// lower triangle, each written to (i, j) and, transposed, to (j, i).
__global__ __launch_bounds__(256) void k_td_update(int m, double* __restrict__ A, int64_t lda, int p1, int se,
                                                   const double* __restrict__ V, const double* __restrict__ W) {
  __shared__ double rv[64][kMB], rw[64][kMB], cv[64][kMB], cw[64][kMB];
  __shared__ double tile[64][65];
  // Lower-triangle tile index b -> (bi, bj), bi >= bj.
  int bi = int((sqrt(8.0 * double(blockIdx.x) + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= int(blockIdx.x)) ++bi;
  while (bi * (bi + 1) / 2 > int(blockIdx.x)) --bi;
  const int bj = int(blockIdx.x) - bi * (bi + 1) / 2;
  const int i0 = p1 + 64 * bi, j0 = p1 + 64 * bj;
  for (int e = threadIdx.x; e < 64 * kMB; e += 256) {
    const int r = e / kMB, t = e - r * kMB;
    const int ri = min(i0 + r, m - 1), rj = min(j0 + r, m - 1);
    const bool ok = t < se;
    rv[r][t] = ok ? V[ri + int64_t(t) * m] : 0.0;
    rw[r][t] = ok ? W[ri + int64_t(t) * m] : 0.0;
    cv[r][t] = ok ? V[rj + int64_t(t) * m] : 0.0;
    cw[r][t] = ok ? W[rj + int64_t(t) * m] : 0.0;
  }
  __syncthreads();
  const int li = threadIdx.x & 63;  // row within the tile (coalesced)
  for (int lj = threadIdx.x >> 6; lj < 64; lj += 4) {
    const int i = i0 + li, j = j0 + lj;
    if (i < m && j < m && i >= j) {
      double dlt = 0.0;
#pragma unroll
      for (int t = 0; t < kMB; ++t) dlt += rv[li][t] * cw[lj][t] + rw[li][t] * cv[lj][t];
      const double nv = A[i + int64_t(j) * lda] - dlt;
      A[i + int64_t(j) * lda] = nv;
      tile[li][lj] = nv;
    }
  }
  __syncthreads();
  // Mirror: (j, i) = (i, j), rows j of the tile's columns written coalesced.
  const int lj = threadIdx.x & 63;
  for (int li2 = threadIdx.x >> 6; li2 < 64; li2 += 4) {
    const int i = i0 + li2, j = j0 + lj;
    if (i < m && j < m && i > j) A[j + int64_t(i) * lda] = tile[li2][lj];
  }
}

// ---------------------------------------------------------------------------
// 1d. Tridiagonalisation by cooperating workgroups ("workers"), for subspaces
// whose matrix fits the workers' LDS together (m up to ~720). Column c lives
// in the LDS of worker c mod nw, so the products and updates of every step
// read LDS instead of streaming the trailing block through one CU. Every
// worker also keeps copies of columns k + 1 and k + 2 and forms each
// reflector itself, so a step needs one exchange between workers, not two
// (the exchange, a trip through memory, is what bounds a step). Step k, with
// v_k, tau_k and the copy of column k + 1 (as of step k - 1) in every worker:
//  A. each worker multiplies its columns c > k by v_k, its part of
//     y = A22 v_k, and publishes it; the owner of column k + 2 also
//     publishes that column (as of step k - 1);
//  B. every worker reads all of y and column k + 2, forms
//     w = tau y - (tau^2/2 y.v) v and applies -(v w^T + w v^T) to its
//     columns and to its copies of columns k + 1 and k + 2, each product
//     rounded on its own: the copies get the owners' bits, and the matrix
//     stays exactly symmetric;
//  C. every worker forms reflector k + 1 from its copy of column k + 1 (the
//     same code on the same data: the same bits everywhere); worker 0
//     records d, e and tau, the owner of column k + 1 keeps the reflector.
// Values travel as two 8-byte words, each holding 32 bits of the double and
// the 32-bit tag k + 1 of their step (written and read as single atomic
// words at agent scope): a reader polls the words until both carry the tag,
// with no separate flag and no cache flush. Parts and columns alternate
// between two slots by the parity of k (a worker reaches step k + 2 only
// after every worker has read step k), and the slots are cleared before the
// launch, so a tag is never stale. The workers are blocks 0 .. nw-1, spread
// over the XCDs by the round-robin dispatch. Every wait is bounded: a worker
// that polls a word more than `spins` times (at once when spins < 0), or
// sees another worker's abort, sets *info = -1 and every worker leaves (the
// caller redoes the solve with the one-workgroup kernels). A is only read
// until the loop ends; the reflectors are then written to A[k+1:, k] as
// k_tridiag leaves them.
// ---------------------------------------------------------------------------
constexpr int kCT = 256;
constexpr int kCoopRd = 5;  // words of each slot a thread reads per step (m <= 5 kCT)

__device__ __forceinline__ void ll_put(uint64_t* p, double v, uint32_t tag) {
  const uint64_t b = uint64_t(__double_as_longlong(v));
  const uint64_t t = uint64_t(tag) << 32;
  __hip_atomic_store(p, t | (b & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p + 1, t | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The double at p once both words carry `tag`; false on a timeout or abort.
__device__ __forceinline__ bool ll_get(const uint64_t* p, uint32_t tag, double* out, long long limit,
                                       const int* info) {
  if (limit < 0) return false;
  long long spins = 0;
  for (;;) {
    const uint64_t lo = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t hi = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (uint32_t(lo >> 32) == tag && uint32_t(hi >> 32) == tag) {
      *out = __longlong_as_double((long long)((hi << 32) | (lo & 0xffffffffull)));
      return true;
    }
    if (++spins > limit) return false;
    if ((spins & 63) == 0 && __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    __builtin_amdgcn_s_sleep(1);
  }
}

// Sum over the kCT-thread block in a fixed order (every worker gets the same
// bits for the same data). `red` holds kCT/64 doubles; ends with a barrier.
__device__ double coop_sum(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kCT / 64; ++i) t += red[i];
  __syncthreads();
  return t;
}

size_t coop_lds(int m, int nw) { return size_t((int64_t((m + nw - 1) / nw) * m + 4 * m + kCT / 64 + 2) * 8); }
// (the last double of the reduction slots holds the workgroup's go flag)

// Words of the exchange slots: two slots of m parts of y and two of m column
// values, two words each (8 m of the 8 m + 4 reserved).
int64_t coop_words(int m) { return 4 * int64_t(m + 1) + 4 * int64_t(m); }

// Reflector k from column `col` (rows k .. m-1 current; every worker): v
// into vv[0 .. m-k-2] (vv[i] for row k + 1 + i), tau into *tkp, and
// x0 - alpha (or x0) into *v1 for the owner's copy. Worker `record` writes
// d, e, tau. Ends with a barrier.
__device__ void coop_reflector(int m, int k, const double* col, double* vv, double* red, double* tkp, double* v1,
                               bool record, double* d, double* e, double* tau) {
  const int tid = threadIdx.x;
  double ss = 0.0;
  for (int r = k + 1 + tid; r < m; r += kCT) ss += col[r] * col[r];
  ss = coop_sum(ss, red);
  const double x0 = col[k + 1];
  const double sigma = sqrt(ss);
  double alpha = 0.0, tk = 0.0;
  if (sigma > 0.0) {
    alpha = x0 >= 0.0 ? -sigma : sigma;
    tk = 1.0 / (sigma * (sigma + fabs(x0)));
  }
  const double first = sigma > 0.0 ? x0 - alpha : x0;
  for (int r = k + 1 + tid; r < m; r += kCT) vv[r - (k + 1)] = r == k + 1 ? first : col[r];
  if (tid == 0) {
    *tkp = tk;
    *v1 = first;
    if (record) {
      d[k] = col[k];
      e[k] = alpha;
      tau[k] = tk;
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(kCT) void k_td_coop(int m, int nw, double* __restrict__ A, int64_t lda,
                                                 double* __restrict__ d, double* __restrict__ e,
                                                 double* __restrict__ tau, uint64_t* __restrict__ words,
                                                 int* __restrict__ info, long long spins) {
  const int me = blockIdx.x;
  extern __shared__ __align__(16) double sh[];
  const int ncl = (m - me + nw - 1) / nw;  // columns me, me + nw, ...
  double* Al = sh;                         // ncl columns of m rows
  double* vv = Al + int64_t((m + nw - 1) / nw) * m;
  double* yw = vv + m;
  double* c1 = yw + m;  // copy of column k + 1
  double* c2 = c1 + m;  // copy of column k + 2
  double* red = c2 + m;
  constexpr int NWV = kCT / 64;
  int& go = *reinterpret_cast<int*>(red + NWV + 1);
  double& v1 = yw[m - 1];  // scratch for the reflector's first entry (yw is idle then)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint64_t* yslot[2] = {words, words + 2 * int64_t(m)};
  uint64_t* cslot[2] = {words + 4 * int64_t(m), words + 6 * int64_t(m)};
  if (tid == 0) go = 1;
  for (int lc = 0; lc < ncl; ++lc) {
    const int c = me + lc * nw;
    for (int r = tid; r < m; r += kCT)
      Al[int64_t(lc) * m + r] = r >= c ? A[r + int64_t(c) * lda] : A[c + int64_t(r) * lda];
  }
  for (int r = tid; r < m; r += kCT) {
    c2[r] = A[r];  // column 0
    c1[r] = r >= 1 ? A[r + lda] : A[1];
  }
  __syncthreads();
  if (m > 2) {
    coop_reflector(m, 0, c2, vv, red, &red[NWV], &v1, me == 0, d, e, tau);
    if (me == 0 && tid == 0) Al[1] = v1;  // column 0's copy keeps its reflector
    __syncthreads();
  }
  for (int k = 0; k + 2 < m; ++k) {
    const int L = m - k - 1;
    const uint32_t tag = uint32_t(k + 1);
    uint64_t* yb = yslot[k & 1];
    uint64_t* cb = cslot[k & 1];
    const double tk = red[NWV];
    const int lc0 = k + 1 <= me ? 0 : (k + 1 - me + nw - 1) / nw;  // first own column > k
    // A. This worker's parts of y (a wave's columns three at a time, so
    // their wave sums overlap); column k + 2 from its owner.
    for (int lc = lc0 + wave; lc < ncl; lc += 3 * NWV) {
      const double* col = Al + int64_t(lc) * m + k + 1;
      const bool has1 = lc + NWV < ncl, has2 = lc + 2 * NWV < ncl;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0;
      for (int i = lane; i < L; i += 64) {
        const double v = vv[i];
        a0 += col[i] * v;
        if (has1) a1 += col[i + NWV * int64_t(m)] * v;
        if (has2) a2 += col[i + 2 * NWV * int64_t(m)] * v;
      }
      a0 = wave_sum(a0);
      a1 = wave_sum(a1);
      a2 = wave_sum(a2);
      if (lane == 0) {
        ll_put(yb + 2 * (me + lc * nw), a0, tag);
        if (has1) ll_put(yb + 2 * (me + (lc + NWV) * nw), a1, tag);
        if (has2) ll_put(yb + 2 * (me + (lc + 2 * NWV) * nw), a2, tag);
      }
    }
    if ((k + 2) % nw == me) {
      const double* col = Al + int64_t((k + 2 - me) / nw) * m;
      for (int r = k + 1 + tid; r < m; r += kCT) ll_put(cb + 2 * r, col[r], tag);
    }
    // B. All of y and column k + 2, then w and the updates. Every load of
    // this thread's words is issued before any is waited on (one trip
    // through memory when the values are there); late words are polled one
    // at a time (polling them all again each time loads the memory system
    // enough to slow the writers: measured slower).
    bool ok = true;
    double s = 0.0;
    {
      uint64_t w4[kCoopRd][4];
#pragma unroll
      for (int j = 0; j < kCoopRd; ++j) {
        const int i = tid + j * kCT;
        if (i < L) {
          const uint64_t* py = yb + 2 * (k + 1 + i);
          const uint64_t* pc = cb + 2 * (k + 1 + i);
          w4[j][0] = __hip_atomic_load(py, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          w4[j][1] = __hip_atomic_load(py + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          w4[j][2] = __hip_atomic_load(pc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          w4[j][3] = __hip_atomic_load(pc + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
#pragma unroll
      for (int j = 0; j < kCoopRd; ++j) {
        const int i = tid + j * kCT;
        if (i < L && ok) {
          if (spins >= 0 && uint32_t(w4[j][0] >> 32) == tag && uint32_t(w4[j][1] >> 32) == tag)
            yw[i] = __longlong_as_double((long long)((w4[j][1] << 32) | (w4[j][0] & 0xffffffffull)));
          else
            ok = ll_get(yb + 2 * (k + 1 + i), tag, &yw[i], spins, info);
          if (ok && uint32_t(w4[j][2] >> 32) == tag && uint32_t(w4[j][3] >> 32) == tag)
            c2[k + 1 + i] = __longlong_as_double((long long)((w4[j][3] << 32) | (w4[j][2] & 0xffffffffull)));
          else if (ok)
            ok = ll_get(cb + 2 * (k + 1 + i), tag, &c2[k + 1 + i], spins, info);
          s += yw[i] * vv[i];
        }
      }
    }
    if (!ok) go = 0;
    s = coop_sum(s, red);
    if (!go) break;
    const double K = 0.5 * tk * tk * s;
    for (int i = tid; i < L; i += kCT) yw[i] = tk * yw[i] - K * vv[i];
    __syncthreads();
    const int ncols = ncl - lc0;
    {
      // (column, row) of flat index t = tid, tid + kCT, ... advanced without
      // a division per element
      const int dq = kCT / L, dr = kCT - dq * L;
      int lc = lc0 + tid / L, i = tid - (tid / L) * L;
      for (int t = tid; t < ncols * L; t += kCT) {
        const int j = me + lc * nw - (k + 1);
        double* a = Al + int64_t(lc) * m + k + 1 + i;
        *a = *a - __dadd_rn(__dmul_rn(vv[i], yw[j]), __dmul_rn(yw[i], vv[j]));
        lc += dq;
        i += dr;
        if (i >= L) {
          i -= L;
          ++lc;
        }
      }
    }
    for (int i = tid; i < L; i += kCT) {
      c1[k + 1 + i] = c1[k + 1 + i] - __dadd_rn(__dmul_rn(vv[i], yw[0]), __dmul_rn(yw[i], vv[0]));
      c2[k + 1 + i] = c2[k + 1 + i] - __dadd_rn(__dmul_rn(vv[i], yw[1]), __dmul_rn(yw[i], vv[1]));
    }
    __syncthreads();
    // C. Reflector k + 1 (or, after the last step, the trailing 2 x 2).
    if (k + 3 < m) {
      coop_reflector(m, k + 1, c1, vv, red, &red[NWV], &v1, me == 0, d, e, tau);
      if ((k + 1) % nw == me && tid == 0) Al[int64_t((k + 1 - me) / nw) * m + k + 2] = v1;
      __syncthreads();
    } else if (me == 0 && tid == 0) {
      d[m - 2] = c1[m - 2];
      e[m - 2] = c1[m - 1];
      tau[m - 2] = 0.0;
      d[m - 1] = c2[m - 1];
      tau[m - 1] = 0.0;
    }
    double* t = c1;
    c1 = c2;
    c2 = t;
  }
  if (!go && m > 2) {
    // Abandoned: the failure is in *info; a zero tridiagonal keeps the
    // later stages' inputs finite.
    if (tid == 0) __hip_atomic_store(info, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int i = tid; i < m; i += kCT) {
      d[i] = 0.0;
      e[i] = 0.0;
      tau[i] = 0.0;
    }
    return;
  }
  // The reflectors into A.
  for (int lc = 0; lc < ncl; ++lc) {
    const int c = me + lc * nw;
    if (c >= m - 2) continue;
    const double* col = Al + int64_t(lc) * m;
    for (int r = c + 1 + tid; r < m; r += kCT) A[r + int64_t(c) * lda] = col[r];
  }
}

// ---------------------------------------------------------------------------
// 1e. Complex Hermitian tridiagonalisation (the k-point Davidson's reduced
// problem): Householder reflectors H_k = I - tau_k v_k v_k^H with v_k[0] = 1
// chosen so that H_k^H x = beta e1 with beta real (the zlarfg convention of
// LAPACK's zhetd2), so the tridiagonal T = Q^H A Q is real symmetric and the
// real bisection / inverse iteration serve it; eigenvectors of A are Q z.
// One workgroup; A (complex, both triangles, interleaved re/im, ld lda
// complex entries) stays in global memory; each step updates both triangles
// from the same rounded products, so A stays exactly Hermitian. v_k ends in
// A[k+1:, k] with A[k+1, k] = 1.
// ---------------------------------------------------------------------------
constexpr int kZT = 256;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(__dadd_rn(__dmul_rn(a.x, b.x), -__dmul_rn(a.y, b.y)),
                      __dadd_rn(__dmul_rn(a.x, b.y), __dmul_rn(a.y, b.x)));
}
// a conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(__dadd_rn(__dmul_rn(a.x, b.x), __dmul_rn(a.y, b.y)),
                      __dadd_rn(__dmul_rn(a.y, b.x), -__dmul_rn(a.x, b.y)));
}

// Sum over the kZT-thread block of (re, im) pairs, to every thread.
__device__ double2 zblock_sum(double2 v, double* red) {
  v.x = wave_sum(v.x);
  v.y = wave_sum(v.y);
  if ((threadIdx.x & 63) == 0) {
    red[2 * (threadIdx.x >> 6)] = v.x;
    red[2 * (threadIdx.x >> 6) + 1] = v.y;
  }
  __syncthreads();
  double2 t = make_double2(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < kZT / 64; ++i) {
    t.x += red[2 * i];
    t.y += red[2 * i + 1];
  }
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(kZT) void k_zhe_tridiag(int m, double2* __restrict__ A, int64_t lda,
                                                     double* __restrict__ d, double* __restrict__ e,
                                                     double2* __restrict__ tau) {
  extern __shared__ double2 zsh[];
  double2* v = zsh;      // m: the reflector
  double2* w = v + m;    // m: p, then w
  double* red = reinterpret_cast<double*>(w + m);  // 2 * kZT / 64
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  auto a = [&](int i, int j) -> double2& { return A[i + int64_t(j) * lda]; };
  for (int k = 0; k + 1 < m; ++k) {
    const int L = m - k - 1;
    double xs = 0.0;
    for (int i = k + 2 + tid; i < m; i += kZT) {
      const double2 x = a(i, k);
      xs += x.x * x.x + x.y * x.y;
    }
    const double xnorm2 = zblock_sum(make_double2(xs, 0.0), red).x;
    const double2 alpha = a(k + 1, k);
    double beta = alpha.x;
    double2 tk = make_double2(0.0, 0.0);
    double2 scale = make_double2(0.0, 0.0);
    if (xnorm2 > 0.0 || alpha.y != 0.0) {
      const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
      beta = alpha.x >= 0.0 ? -nrm : nrm;
      tk = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      // scale = 1 / (alpha - beta)
      const double2 den = make_double2(alpha.x - beta, alpha.y);
      const double dd = den.x * den.x + den.y * den.y;
      scale = make_double2(den.x / dd, -den.y / dd);
    }
    // v = (1, x[1:] * scale) into LDS and A[k+1:, k].
    for (int i = tid; i < L; i += kZT) {
      double2 vi = make_double2(1.0, 0.0);
      if (i > 0) vi = (tk.x != 0.0 || tk.y != 0.0) ? cmul(a(k + 1 + i, k), scale) : a(k + 1 + i, k);
      v[i] = vi;
    }
    __syncthreads();
    for (int i = tid; i < L; i += kZT) a(k + 1 + i, k) = v[i];
    if (tid == 0) {
      d[k] = a(k, k).x;
      e[k] = beta;
      tau[k] = tk;
    }
    if (tk.x == 0.0 && tk.y == 0.0) {
      __syncthreads();
      continue;
    }
    // p = tau A22 v: a wave per row i, lanes down column i (A(i, j) =
    // conj(A(j, i)), contiguous reads).
    for (int i = wave; i < L; i += kZT / 64) {
      double2 acc = make_double2(0.0, 0.0);
      for (int j = lane; j < L; j += 64) {
        const double2 aji = a(k + 1 + j, k + 1 + i);
        const double2 vj = v[j];
        acc.x += aji.x * vj.x + aji.y * vj.y;
        acc.y += aji.x * vj.y - aji.y * vj.x;
      }
      acc.x = wave_sum(acc.x);
      acc.y = wave_sum(acc.y);
      if (lane == 0) w[i] = cmul(tk, acc);
    }
    __syncthreads();
    // alpha2 = -tau/2 (p^H v); w = p + alpha2 v.
    double2 pv = make_double2(0.0, 0.0);
    for (int i = tid; i < L; i += kZT) {
      const double2 c = cmulc(v[i], w[i]);  // conj(p_i) v_i
      pv.x += c.x;
      pv.y += c.y;
    }
    pv = zblock_sum(pv, red);
    const double2 al2 = cmul(make_double2(-0.5 * tk.x, -0.5 * tk.y), pv);
    for (int i = tid; i < L; i += kZT) {
      const double2 p = w[i], t = cmul(al2, v[i]);
      w[i] = make_double2(p.x + t.x, p.y + t.y);
    }
    __syncthreads();
    // A22 -= v w^H + w v^H, both triangles (entry (j, i) is the conjugate of
    // entry (i, j) to the bit).
    for (int t = tid; t < L * L; t += kZT) {
      const int i = t % L, j = t / L;
      const double2 p1 = cmulc(v[i], w[j]), p2 = cmulc(w[i], v[j]);
      double2& aij = a(k + 1 + i, k + 1 + j);
      aij = make_double2(aij.x - __dadd_rn(p1.x, p2.x), aij.y - __dadd_rn(p1.y, p2.y));
    }
    __syncthreads();
  }
  if (tid == 0) {
    d[m - 1] = a(m - 1, m - 1).x;
    tau[m - 1] = make_double2(0.0, 0.0);
  }
}

// Y (complex, m x want) = Q Z with Z real (m x want, ld ldz): a wave per
// vector, reflectors from A[k+1:, k] (v[0] = 1) applied last to first.
__global__ __launch_bounds__(256) void k_zback_transform(int m, int want, const double2* __restrict__ A, int64_t lda,
                                                         const double2* __restrict__ tau,
                                                         const double* __restrict__ Z, int64_t ldz,
                                                         double2* __restrict__ Y, int64_t ldy) {
  extern __shared__ double2 ysh[];
  const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
  const int col = blockIdx.x * (blockDim.x >> 6) + wl;
  if (col >= want) return;
  double2* y = ysh + int64_t(wl) * m;
  for (int i = lane; i < m; i += 64) y[i] = make_double2(Z[i + int64_t(col) * ldz], 0.0);
  for (int k = m - 2; k >= 0; --k) {
    const double2 t = tau[k];
    if (t.x == 0.0 && t.y == 0.0) continue;
    const double2* v = A + (k + 1) + int64_t(k) * lda;
    double2 acc = make_double2(0.0, 0.0);
    for (int i = lane; i < m - k - 1; i += 64) {
      const double2 c = cmulc(y[k + 1 + i], v[i]);  // conj(v_i) y_i
      acc.x += c.x;
      acc.y += c.y;
    }
    acc.x = wave_sum(acc.x);
    acc.y = wave_sum(acc.y);
    const double2 sc = cmul(t, acc);
    for (int i = lane; i < m - k - 1; i += 64) {
      const double2 u = cmul(sc, v[i]);
      y[k + 1 + i] = make_double2(y[k + 1 + i].x - u.x, y[k + 1 + i].y - u.y);
    }
  }
  for (int i = lane; i < m; i += 64) Y[i + int64_t(col) * ldy] = y[i];
}

// ---------------------------------------------------------------------------
// 2. Bisection: one wave per wanted eigenvalue index j (ascending).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int sturm_count(int m, const double* ds, const double* e2, double x, double pivmin) {
  double q = ds[0] - x;
  if (fabs(q) <= pivmin) q = -pivmin;
  int c = q < 0.0 ? 1 : 0;
  for (int i = 1; i < m; ++i) {
    q = ds[i] - x - e2[i - 1] / q;
    if (fabs(q) <= pivmin) q = -pivmin;
    c += q < 0.0 ? 1 : 0;
  }
  return c;
}

__global__ __launch_bounds__(256) void k_bisect(int m, int want, const double* __restrict__ d,
                                                const double* __restrict__ e, double* __restrict__ values,
                                                double* __restrict__ tnorm_out) {
  extern __shared__ double sh[];
  double* ds = sh;
  double* e2 = sh + m;
  double* ea = sh + 2 * m;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    ds[i] = d[i];
    const double v = i + 1 < m ? e[i] : 0.0;
    e2[i] = v * v;
    ea[i] = fabs(v);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (j >= want) return;  // whole waves leave together; no barrier follows
  double gl = INFINITY, gu = -INFINITY, emax2 = 0.0;
  for (int i = lane; i < m; i += 64) {
    const double r = (i > 0 ? ea[i - 1] : 0.0) + ea[i];
    gl = fmin(gl, ds[i] - r);
    gu = fmax(gu, ds[i] + r);
    emax2 = fmax(emax2, e2[i]);
  }
  gl = wave_min(gl);
  gu = wave_max(gu);
  emax2 = wave_max(emax2);
  const double tnorm = fmax(fabs(gl), fabs(gu));
  const double pivmin = DBL_MIN * fmax(1.0, emax2);
  const double pad = 2.1 * tnorm * DBL_EPSILON * m + 4.2 * pivmin;
  double lo = gl - pad, hi = gu + pad;
  // Invariant: count(lo) <= j < count(hi). 64 points per step shrink the
  // bracket 65x; the cap bounds the loop.
  for (int it = 0; it < 48; ++it) {
    const double width = hi - lo;
    if (width <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    const double xl = lo + width * (double(lane + 1) / 65.0);
    const int c = sturm_count(m, ds, e2, xl, pivmin);
    const unsigned long long above = __ballot(c >= j + 1);
    if (above == 0ull) {
      lo = __shfl(xl, 63, 64);
    } else {
      const int f = __ffsll(static_cast<long long>(above)) - 1;
      const double nhi = __shfl(xl, f, 64);
      const double nlo = __shfl(xl, f > 0 ? f - 1 : 0, 64);
      hi = nhi;
      if (f > 0) lo = nlo;
    }
  }
  if (lane == 0) {
    values[j] = 0.5 * (lo + hi);
    if (j == 0) tnorm_out[0] = tnorm;
  }
}

// ---------------------------------------------------------------------------
// 3. Inverse iteration. Per eigenvector j the LU of (T - lambda_j I) with
// partial pivoting: U has the reciprocal diagonal U0i and two superdiagonals
// U1, U2; Lm the multipliers, piv the row swaps. Arrays are laid out
// [i * want + j] so lanes (consecutive j) access consecutive addresses.
// ---------------------------------------------------------------------------
struct LU {
  double* U0i;
  double* U1;
  double* U2;
  double* Lm;
  unsigned char* piv;
};

// Solve (T - lambda_j I) y = z in place (z[i * zs]); returns max |y|.
__device__ double lu_solve(int m, int want, int j, const LU& f, double* z, int64_t zs) {
  auto at = [&](int i) { return int64_t(i) * want + j; };
  double cur = z[0];
  for (int i = 0; i + 1 < m; ++i) {
    double nxt = z[int64_t(i + 1) * zs];
    if (f.piv[at(i)]) {
      const double tmp = cur;
      cur = nxt;
      nxt = tmp;
    }
    nxt -= f.Lm[at(i)] * cur;
    z[int64_t(i) * zs] = cur;
    cur = nxt;
  }
  z[int64_t(m - 1) * zs] = cur;
  double y1 = 0.0, y2 = 0.0, big = 0.0;
  for (int i = m - 1; i >= 0; --i) {
    const double y = (z[int64_t(i) * zs] - f.U1[at(i)] * y1 - f.U2[at(i)] * y2) * f.U0i[at(i)];
    z[int64_t(i) * zs] = y;
    big = fmax(big, fabs(y));
    y2 = y1;
    y1 = y;
  }
  return big;
}

// (T - lambda I)^{-1} z in place for the inverse-iteration kernel: the LU
// factors and z are fetched kInvChunk rows ahead into registers, so the
// sequential sweep waits on memory once per chunk. Returns sum y^2.
constexpr int kInvChunk = 8;

__device__ double lu_solve_prefetch(int m, int want, int j, const LU& f, double* z) {
  auto at = [&](int i) { return int64_t(i) * want + j; };
  double cur = z[at(0)];
  for (int i0 = 0; i0 + 1 < m; i0 += kInvChunk) {
    const int cnt = min(kInvChunk, m - 1 - i0);
    double zn[kInvChunk], lm[kInvChunk];
    unsigned char pv[kInvChunk];
#pragma unroll
    for (int c = 0; c < kInvChunk; ++c)
      if (c < cnt) {
        zn[c] = z[at(i0 + c + 1)];
        lm[c] = f.Lm[at(i0 + c)];
        pv[c] = f.piv[at(i0 + c)];
      }
#pragma unroll
    for (int c = 0; c < kInvChunk; ++c)
      if (c < cnt) {
        double nxt = zn[c];
        if (pv[c]) {
          const double tmp = cur;
          cur = nxt;
          nxt = tmp;
        }
        nxt -= lm[c] * cur;
        z[at(i0 + c)] = cur;
        cur = nxt;
      }
  }
  z[at(m - 1)] = cur;
  double y1 = 0.0, y2 = 0.0, nn = 0.0;
  for (int i1 = m - 1; i1 >= 0; i1 -= kInvChunk) {
    const int cnt = min(kInvChunk, i1 + 1);
    double zz[kInvChunk], u0[kInvChunk], u1[kInvChunk], u2[kInvChunk];
#pragma unroll
    for (int c = 0; c < kInvChunk; ++c)
      if (c < cnt) {
        const int64_t ix = at(i1 - c);
        zz[c] = z[ix];
        u0[c] = f.U0i[ix];
        u1[c] = f.U1[ix];
        u2[c] = f.U2[ix];
      }
#pragma unroll
    for (int c = 0; c < kInvChunk; ++c)
      if (c < cnt) {
        const double y = (zz[c] - u1[c] * y1 - u2[c] * y2) * u0[c];
        z[at(i1 - c)] = y;
        nn += y * y;
        y2 = y1;
        y1 = y;
      }
  }
  return nn;
}

// The per-vector work of inverse iteration: LU of (T - lambda' I) into g
// (entries at i * stride + lane), a hashed start, two solves, unit norm.
// Column j of the caller's output numbering seeds the start vector.
__device__ void invit_vector(int m, int stride, int lane, int j, const double* ds, const double* es, double lam,
                             double tiny, const LU& g, double* z) {
  auto at = [&](int i) { return int64_t(i) * stride + lane; };
  double a = ds[0] - lam, b = es[0];
  for (int i = 0; i + 1 < m; ++i) {
    const double c = es[i];
    const double dn = ds[i + 1] - lam;
    const double en = es[i + 1];
    if (fabs(a) >= fabs(c)) {
      double p = a;
      if (fabs(p) < tiny) p = p >= 0.0 ? tiny : -tiny;
      const double r = 1.0 / p;
      g.U0i[at(i)] = r;
      g.U1[at(i)] = b;
      g.U2[at(i)] = 0.0;
      g.Lm[at(i)] = c * r;
      g.piv[at(i)] = 0;
      a = dn - c * r * b;
      b = en;
    } else {
      const double r = 1.0 / c;
      const double l = a * r;
      g.U0i[at(i)] = r;
      g.U1[at(i)] = dn;
      g.U2[at(i)] = en;
      g.Lm[at(i)] = l;
      g.piv[at(i)] = 1;
      a = b - l * dn;
      b = -l * en;
    }
  }
  if (fabs(a) < tiny) a = a >= 0.0 ? tiny : -tiny;
  g.U0i[at(m - 1)] = 1.0 / a;
  // The last row has no superdiagonals: the back substitution multiplies
  // these by zero, so they must not be left as whatever the memory held (an
  // Inf or NaN there would poison the vector).
  g.U1[at(m - 1)] = 0.0;
  g.U2[at(m - 1)] = 0.0;
  g.Lm[at(m - 1)] = 0.0;
  g.piv[at(m - 1)] = 0;
  for (int i = 0; i < m; ++i) z[at(i)] = hash_unit(i, j, 0);
  // Two solves (the second from the unscaled first: growth per solve is at
  // most ~1/(eps ||T||), far from overflow), then unit norm.
  lu_solve_prefetch(m, stride, lane, g, z);
  const double nn = lu_solve_prefetch(m, stride, lane, g, z);
  const double sc = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
  for (int i = 0; i < m; ++i) z[at(i)] *= sc;
}

// Shift of vector j: shifts of (numerically) equal eigenvalues are spread
// 10 eps ||T|| apart, as LAPACK's dstein does: an exactly repeated eigenvalue
// (the realified complex problem has only such pairs) would otherwise give
// every vector of the pair one singular factorisation and one direction.
__device__ double invit_shift(const double* values, int j, double tnorm) {
  const double pertol = 10.0 * DBL_EPSILON * tnorm;
  double lam = values[0];
  for (int i = 1; i <= j; ++i) lam = fmax(values[i], lam + pertol);
  return lam;
}

// Lane per vector, factors and vector in global memory (any m).
__global__ __launch_bounds__(64) void k_invit(int m, int want, const double* __restrict__ d,
                                              const double* __restrict__ e, const double* __restrict__ values,
                                              const double* __restrict__ tnorm_in, LU f, double* __restrict__ Z) {
  extern __shared__ double sh[];  // d, e of the tridiagonal (every lane reads the same entries)
  double* ds = sh;
  double* es = sh + m;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    ds[i] = d[i];
    es[i] = i + 1 < m ? e[i] : 0.0;
  }
  __syncthreads();
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= want) return;
  const double tiny = DBL_EPSILON * fmax(tnorm_in[0], DBL_MIN);
  invit_vector(m, want, j, j, ds, es, invit_shift(values, j, tnorm_in[0]), tiny, f, Z);
}

// The same with kInvLds vectors per workgroup whose factors and vector stay in
// LDS (a dependent chain of LDS accesses instead of L2 round trips), copied
// to the global layout at the end (the cluster re-iteration reads them there).
// 8 vectors per workgroup up to m = 476, 4 from there to m = 910.
template <int kInvLds>
__global__ __launch_bounds__(kInvLds) void k_invit_lds(int m, int want, const double* __restrict__ d,
                                                       const double* __restrict__ e,
                                                       const double* __restrict__ values,
                                                       const double* __restrict__ tnorm_in, LU f,
                                                       double* __restrict__ Z) {
  extern __shared__ double sh[];
  double* ds = sh;
  double* es = sh + m;
  const int64_t mv = int64_t(m) * kInvLds;
  LU g;
  g.U0i = sh + 2 * m;
  g.U1 = g.U0i + mv;
  g.U2 = g.U1 + mv;
  g.Lm = g.U2 + mv;
  double* z = g.Lm + mv;
  g.piv = reinterpret_cast<unsigned char*>(z + mv);
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    ds[i] = d[i];
    es[i] = i + 1 < m ? e[i] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x;
  const int j = blockIdx.x * kInvLds + lane;
  if (j >= want) return;
  const double tiny = DBL_EPSILON * fmax(tnorm_in[0], DBL_MIN);
  invit_vector(m, kInvLds, lane, j, ds, es, invit_shift(values, j, tnorm_in[0]), tiny, g, z);
  for (int i = 0; i < m; ++i) {
    const int64_t src = int64_t(i) * kInvLds + lane, dst = int64_t(i) * want + j;
    Z[dst] = z[src];
    f.U0i[dst] = g.U0i[src];
    f.U1[dst] = g.U1[src];
    f.U2[dst] = g.U2[src];
    f.Lm[dst] = g.Lm[src];
    f.piv[dst] = g.piv[src];
  }
}

// ---------------------------------------------------------------------------
// 4. Orthogonality. Inverse-iteration vectors of eigenvalues a gap g apart
// overlap by about eps ||T|| / g, so vectors whose eigenvalues lie within
// kClusterRel ||T|| of a neighbour form a cluster and are made orthogonal by
// two passes of Gram-Schmidt in ascending order, one wave per cluster (a
// vector that loses most of its norm is re-iterated from the remainder plus
// a fresh direction). The remaining overlaps, <= ~1e-9, are removed by one
// symmetric (Loewdin) step Y <- Y (3 I - Y^T Y) / 2 in the host driver.
// Reads the inverse-iteration vectors from Z ([i * want + j]); writes Y.
// ---------------------------------------------------------------------------
constexpr double kClusterRel = 1e-6;

__global__ __launch_bounds__(256) void k_cluster_gs(int m, int want, const double* __restrict__ values,
                                                    const double* __restrict__ tnorm_in, LU f,
                                                    const double* __restrict__ Z, double* __restrict__ Y,
                                                    int64_t ldy, int* __restrict__ info) {
  const int lane = threadIdx.x & 63;
  const int h = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (h >= want) return;  // whole waves; no barrier follows
  const double tol = kClusterRel * tnorm_in[0];
  if (h > 0 && values[h] - values[h - 1] <= tol) return;  // not the head of its cluster
  int end = h + 1;
  while (end < want && values[end] - values[end - 1] <= tol) ++end;
  for (int j = h; j < end; ++j) {
    double* y = Y + int64_t(j) * ldy;
    for (int i = lane; i < m; i += 64) y[i] = Z[int64_t(i) * want + j];
    double scale = 0.0;
    for (int attempt = 0; attempt < 4; ++attempt) {
      for (int pass = 0; pass < 2 && j > h; ++pass)
        for (int l = h; l < j; ++l) {
          const double* yl = Y + int64_t(l) * ldy;
          double acc = 0.0;
          for (int i = lane; i < m; i += 64) acc += yl[i] * y[i];
          const double c = wave_sum(acc);
          for (int i = lane; i < m; i += 64) y[i] -= c * yl[i];
        }
      double nn = 0.0;
      for (int i = lane; i < m; i += 64) nn += y[i] * y[i];
      nn = wave_sum(nn);
      if (nn > 1e-2 || attempt == 3) {
        if (!(nn > 1e-2) && lane == 0) info[0] = 1 + j;
        scale = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
        break;
      }
      // Most of the vector lay in the span of its cluster: restart from the
      // remainder plus a fresh direction and apply (T - lambda_j I)^{-1},
      // projecting the cluster's accepted vectors out after every solve. A
      // nearly singular factor sends every right-hand side towards the same
      // direction; without the projection in between, a second solve would
      // swamp the wanted direction in that one.
      const double rs = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
      for (int i = lane; i < m; i += 64) y[i] = y[i] * rs + 0.1 * hash_unit(i, j, 1 + attempt);
      for (int it = 0; it < 2; ++it) {
        __threadfence_block();
        if (lane == 0) {
          const double big = lu_solve(m, want, j, f, y, 1);
          const double sc = big > 0.0 ? 1.0 / big : 1.0;
          for (int i = 0; i < m; ++i) y[i] *= sc;
        }
        __threadfence_block();
        for (int l = h; l < j; ++l) {
          const double* yl = Y + int64_t(l) * ldy;
          double acc = 0.0;
          for (int i = lane; i < m; i += 64) acc += yl[i] * y[i];
          const double c = wave_sum(acc);
          for (int i = lane; i < m; i += 64) y[i] -= c * yl[i];
        }
        double n2 = 0.0;
        for (int i = lane; i < m; i += 64) n2 += y[i] * y[i];
        n2 = wave_sum(n2);
        const double sc = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
        for (int i = lane; i < m; i += 64) y[i] *= sc;
      }
      __threadfence_block();
    }
    for (int i = lane; i < m; i += 64) y[i] *= scale;
    __threadfence_block();
  }
}

// M = (3 I - G) / 2 in place (n x n, ld n).
__global__ void k_loewdin_factor(int n, double* __restrict__ G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < n) G[i + int64_t(j) * n] = (i == j ? 1.5 : 0.0) - 0.5 * G[i + int64_t(j) * n];
}

// ---------------------------------------------------------------------------
// 5. Y <- H_0 H_1 ... H_{m-3} Y, one wave per column (its column in LDS,
// lane-owned rows); the block stages kBtChunk reflectors at a time in LDS.
// ---------------------------------------------------------------------------
// CH reflectors staged at a time: 8, or 4 where (4 + 8) m doubles exceed the LDS.
template <int CH>
__global__ __launch_bounds__(256) void k_back_transform(int m, int want, const double* __restrict__ A, int64_t lda,
                                                        const double* __restrict__ tau, double* __restrict__ Y,
                                                        int64_t ldy) {
  constexpr int kBtChunk = CH;
  extern __shared__ double sh[];
  const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
  const int col = blockIdx.x * (blockDim.x >> 6) + wl;
  const bool active = col < want;
  double* y = sh + int64_t(wl) * m;
  double* vs = sh + 4 * int64_t(m);
  double* ts = vs + int64_t(kBtChunk) * m;
  double* yg = Y + int64_t(active ? col : 0) * ldy;
  if (active)
    for (int i = lane; i < m; i += 64) y[i] = yg[i];
  for (int k1 = m - 3; k1 >= 0; k1 -= kBtChunk) {
    const int cnt = k1 + 1 < kBtChunk ? k1 + 1 : kBtChunk;  // reflectors k1, k1-1, ..., k1-cnt+1
    __syncthreads();
    // Stage the chunk's reflectors: 16 independent loads per thread per batch.
    for (int e0 = 0; e0 < cnt * m; e0 += 16 * int(blockDim.x)) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int e = e0 + u * int(blockDim.x) + int(threadIdx.x);
        const int c = e / m, i = e - c * m, k = k1 - c;
        v[u] = (e < cnt * m && i > k) ? A[i + int64_t(k) * lda] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int e = e0 + u * int(blockDim.x) + int(threadIdx.x);
        if (e < cnt * m) vs[e] = v[u];
      }
    }
    if (int(threadIdx.x) < cnt) ts[threadIdx.x] = tau[k1 - int(threadIdx.x)];
    __syncthreads();
    if (!active) continue;
    for (int c = 0; c < cnt; ++c) {
      const double t = ts[c];
      if (t == 0.0) continue;
      const int k = k1 - c;
      const double* v = vs + int64_t(c) * m;
      double acc = 0.0;
      for (int i = k + 1 + ((lane - (k + 1)) & 63); i < m; i += 64) acc += v[i] * y[i];
      const double sc = t * wave_sum(acc);
      for (int i = k + 1 + ((lane - (k + 1)) & 63); i < m; i += 64) y[i] -= sc * v[i];
    }
  }
  if (active)
    for (int i = lane; i < m; i += 64) yg[i] = y[i];
}

// The same for m <= 64 YP with the vector in registers (YP rows per lane):
// per reflector the chain is the lane's dot, one wave sum and the update,
// with the chunk's reflector rows read from LDS ahead of it. The next chunk
// of 8 reflectors is read from memory into registers while the current one
// is applied (two LDS buffers).
constexpr int kBtRegCh = 8;
size_t bt_reg_lds(int m) { return size_t(2 * kBtRegCh * m + 2 * kBtRegCh) * 8; }

template <int YP>
__global__ __launch_bounds__(256) void k_back_transform_reg(int m, int want, const double* __restrict__ A,
                                                            int64_t lda, const double* __restrict__ tau,
                                                            double* __restrict__ Y, int64_t ldy) {
  constexpr int CH = kBtRegCh;
  constexpr int NPT = (CH * 64 * YP + 255) / 256;  // staged values per thread (m <= 64 YP)
  extern __shared__ double sh[];
  double* vb[2] = {sh, sh + CH * m};
  double* tb[2] = {sh + 2 * CH * m, sh + 2 * CH * m + CH};
  const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
  const int col = blockIdx.x * (blockDim.x >> 6) + wl;
  const bool active = col < want;
  double* yg = Y + int64_t(active ? col : 0) * ldy;
  double y[YP];
#pragma unroll
  for (int r = 0; r < YP; ++r) {
    const int i = lane + 64 * r;
    y[r] = (active && i < m) ? yg[i] : 0.0;
  }
  // Chunk starting at reflector k1 (k1, k1 - 1, ...): values into registers.
  double nv[NPT];
  double nt = 0.0;
  auto fetch = [&](int k1) {
    const int cnt = k1 + 1 < CH ? k1 + 1 : CH;
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
      const int e = u * 256 + int(threadIdx.x);
      const int c = e / m, i = e - c * m, k = k1 - c;
      nv[u] = (c < cnt && i > k) ? A[i + int64_t(k) * lda] : 0.0;
    }
    nt = int(threadIdx.x) < cnt ? tau[k1 - int(threadIdx.x)] : 0.0;
  };
  auto store = [&](int b) {
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
      const int e = u * 256 + int(threadIdx.x);
      if (e < CH * m) vb[b][e] = nv[u];
    }
    if (int(threadIdx.x) < CH) tb[b][threadIdx.x] = nt;
  };
  int k1 = m - 3;
  if (k1 >= 0) {
    fetch(k1);
    store(0);
  }
  __syncthreads();
  for (int b = 0; k1 >= 0; k1 -= CH, b ^= 1) {
    if (k1 - CH >= 0) fetch(k1 - CH);
    if (active) {
      const double* vs = vb[b];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const double t = tb[b][c];  // 0 for reflectors past the chunk's end
        double v[YP];
#pragma unroll
        for (int r = 0; r < YP; ++r) {
          const int i = lane + 64 * r;
          v[r] = i < m ? vs[c * m + i] : 0.0;
        }
        double acc = 0.0;
#pragma unroll
        for (int r = 0; r < YP; ++r) acc = fma(v[r], y[r], acc);
        const double sc = t * wave_sum(acc);
#pragma unroll
        for (int r = 0; r < YP; ++r) y[r] = fma(-sc, v[r], y[r]);
      }
    }
    __syncthreads();  // every wave is done with buffer b ^ 1 (read a chunk ago)
    if (k1 - CH >= 0) store(b ^ 1);
    __syncthreads();
  }
  if (active) {
#pragma unroll
    for (int r = 0; r < YP; ++r) {
      const int i = lane + 64 * r;
      if (i < m) yg[i] = y[r];
    }
  }
}

__global__ void k_copy_cols(int m, const double* __restrict__ src, int64_t lds, double* __restrict__ dst,
                            int64_t ldd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < m) dst[i + int64_t(j) * ldd] = src[i + int64_t(j) * lds];
}

void copy_columns(hipStream_t s, int m, int n, const double* src, int64_t lds, double* dst, int64_t ldd) {
  hipLaunchKernelGGL(k_copy_cols, dim3((m + 255) / 256, n), dim3(256), 0, s, m, src, lds, dst, ldd);
}

__global__ void k_zero_i(int* p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}
__global__ void k_zero_d(double* p, int64_t count) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0.0;
}

struct EigenLayout {
  double* d;
  double* e;
  double* tau;
  double* tnorm;
  LU f;
  double* Z;
  double* G;      // want x want (the Loewdin step)
  double* Ytmp;   // m x want
  double* gwork;  // split-K workspace of Y^T Y
  double* td;     // tridiag_multi: V, W (m x kMB), y (m), a scalar
  uint64_t* words;  // k_td_coop: the exchange slots (coop_words(m))
};

EigenLayout eigen_layout(int m, int want, double* work) {
  const int64_t mw = int64_t(m) * want;
  EigenLayout L;
  L.d = work;
  L.e = L.d + m;
  L.tau = L.e + m;
  L.tnorm = L.tau + m;
  L.f.U0i = L.tnorm + 8;
  L.f.U1 = L.f.U0i + mw;
  L.f.U2 = L.f.U1 + mw;
  L.f.Lm = L.f.U2 + mw;
  L.Z = L.f.Lm + mw;
  L.f.piv = reinterpret_cast<unsigned char*>(L.Z + mw);
  L.G = L.Z + mw + (mw + 7) / 8 + 8;
  L.Ytmp = L.G + int64_t(want) * want;
  L.gwork = L.Ytmp + mw;
  L.td = L.gwork + gemm_tn_work(want, want, m);
  L.words = reinterpret_cast<uint64_t*>(L.td + 2 * int64_t(m) * kMB + m + 8);
  return L;
}
}  // namespace

// The attribute is per kernel and process-wide: set per call to the call's
// own size, it would race between host threads that launch the same kernel
// with different sizes (a launch above the current maximum fails). Setting
// the one largest value once avoids that.
void allow_full_lds(const void* fn, const char* name) {
  static std::mutex mu;
  static std::set<const void*> done;
  std::lock_guard<std::mutex> lock(mu);
  if (done.count(fn)) return;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess) {
    (void)hipGetLastError();  // not sticky for the caller's later calls
    fail(BD_ERR_DEVICE, std::string(name) + " LDS attribute");
  }
  done.insert(fn);
}

int64_t sym_eigen_work(int m, int want) {
  const int64_t mw = int64_t(m) * want;
  return 3 * int64_t(m) + 8 + 5 * mw + (mw + 7) / 8 + 8 + int64_t(want) * want + mw + gemm_tn_work(want, want, m) +
         2 * int64_t(m) * kMB + m + 8 + coop_words(m);
}

// Cooperative tridiagonalisation: worker count (0: off; -1: m/10 clamped to
// 8..32, the fastest measured from m = 160 to 576 on MI355X) and the
// subspace sizes it takes [from, to]; the count grows where the columns do
// not fit the workers' LDS.
std::atomic<int> coop_workers{-1};
std::atomic<int> coop_from{128};  // smaller subspaces (the k points') stay in one workgroup
std::atomic<int> coop_to{1025};
std::atomic<int> coop_spins{1 << 21};

int coop_worker_count(int m) {
  int w = coop_workers.load();
  if (w < 0) w = std::min(32, std::max(8, m / 10));
  if (w == 0 || m < 3 || m < coop_from.load() || m > coop_to.load() || m > kCoopRd * kCT) return 0;
  for (int nw = w; nw <= 32; ++nw)
    if (coop_lds(m, nw) <= 160 * 1024) return nw;
  return 0;
}

// Subspaces of at least this size are tridiagonalised across the chip.
std::atomic<int> multi_from{1026};

// One step at a time (k_tridiag): for subspaces whose panel LDS does not fit.
void tridiag_unblocked(hipStream_t s, int m, double* A, int64_t lda, const EigenLayout& L) {
  const size_t tri_lds_lower = size_t(5 * m + 2 * kTW + kTW * m) * 8;
  const bool lower = tri_lds_lower <= 160 * 1024;
  const size_t tri_lds = lower ? tri_lds_lower : size_t(5 * m + 2 * kTW) * 8;
  const void* tri_fn = lower ? reinterpret_cast<const void*>(&k_tridiag<true>)
                             : reinterpret_cast<const void*>(&k_tridiag<false>);
  if (tri_lds > 64 * 1024) allow_full_lds(tri_fn, "k_tridiag");
  if (lower)
    hipLaunchKernelGGL(k_tridiag<true>, dim3(1), dim3(kTT), tri_lds, s, m, A, lda, L.d, L.e, L.tau);
  else
    hipLaunchKernelGGL(k_tridiag<false>, dim3(1), dim3(kTT), tri_lds, s, m, A, lda, L.d, L.e, L.tau);
}

// Across the chip (k_td_column / k_td_symv / k_td_update), for m >= 3; the
// scratch T holds V, W (m x kMB each), y (m) and one scalar.
void tridiag_multi(hipStream_t s, int m, double* A, int64_t lda, const EigenLayout& L, double* scratch) {
  double* V = scratch;
  double* W = V + int64_t(m) * kMB;
  double* y = W + int64_t(m) * kMB;
  double* sc = y + m;
  const size_t col_lds = size_t(2 * m + 4 * kMB + kTW) * 8;
  if (col_lds > 64 * 1024) allow_full_lds(reinterpret_cast<const void*>(&k_td_column), "k_td_column");
  hipLaunchKernelGGL(k_lower_to_upper, dim3((m + 255) / 256, m), dim3(256), 0, s, m, A, lda);
  const int last = m - 2;  // reflector steps k = 0 .. m-3
  for (int p0 = 0; p0 < last; p0 += kMB) {
    const int se = std::min(kMB, last - p0);
    for (int st = 0; st < se; ++st) {
      const int k = p0 + st;
      hipLaunchKernelGGL(k_td_column, dim3(1), dim3(kTT), col_lds, s, m, A, lda, p0, st, 0, V, W, y, sc, L.d, L.e,
                         L.tau);
      hipLaunchKernelGGL(k_td_symv, dim3((m - k - 1 + 3) / 4), dim3(256), 0, s, m, A, lda, k, V + int64_t(st) * m,
                         y);
    }
    hipLaunchKernelGGL(k_td_column, dim3(1), dim3(kTT), col_lds, s, m, A, lda, p0, se, 1, V, W, y, sc, L.d, L.e,
                       L.tau);
    const int p1 = p0 + se, tiles = (m - p1 + 63) / 64;
    hipLaunchKernelGGL(k_td_update, dim3(tiles * (tiles + 1) / 2), dim3(256), 0, s, m, A, lda, p1, se, V, W);
  }
  hipLaunchKernelGGL(k_td_column, dim3(1), dim3(kTT), col_lds, s, m, A, lda, m - 2, 0, 2, V, W, y, sc, L.d, L.e,
                     L.tau);
}

void tridiag_vectors(hipStream_t s, int m, int want, const EigenLayout& L, double* values, double* Ysc, int64_t ldy,
                     int* info);

void sym_eigen_lowest(hipStream_t s, int m, int want, double* A, int64_t lda, double* values, double* Y,
                      int64_t ldy, double* work, int* info, bool blocked) {
  if (m <= 0 || want <= 0 || want > m || lda < m || ldy < m || lda * m >= (int64_t(1) << 31))
    fail(BD_ERR_INVALID_PARAMETER, "sym_eigen_lowest: shape");
  auto bt_lds = [m](int ch) { return size_t((4 + ch) * m + ch) * 8; };
  const int bt_ch = bt_lds(8) <= 160 * 1024 ? 8 : 4;
  if (size_t(3 * m) * 8 > 64 * 1024 || bt_lds(bt_ch) > 160 * 1024)
    fail(BD_ERR_INVALID_PARAMETER, "sym_eigen_lowest: subspace too large for the LDS kernels (m <= 2730)");
  const EigenLayout L = eigen_layout(m, want, work);
  hipLaunchKernelGGL(k_zero_i, dim3(1), dim3(64), 0, s, info, 1);
  const size_t panel_lds = tridiag_panel_lds(m);
  const int nw = blocked ? coop_worker_count(m) : 0;
  if (m >= 3 && m >= multi_from.load()) {
    tridiag_multi(s, m, A, lda, L, L.td);
  } else if (nw > 0) {
    allow_full_lds(reinterpret_cast<const void*>(&k_td_coop), "k_td_coop");
    zero_doubles(s, reinterpret_cast<double*>(L.words), coop_words(m));
    hipLaunchKernelGGL(k_td_coop, dim3(nw), dim3(kCT), coop_lds(m, nw), s, m, nw, A, lda, L.d, L.e, L.tau,
                       L.words, info, (long long)coop_spins.load());
  } else if (blocked && m >= 3 && panel_lds <= 160 * 1024) {
    if (panel_lds > 64 * 1024) allow_full_lds(reinterpret_cast<const void*>(&k_tridiag_panel), "k_tridiag_panel");
    hipLaunchKernelGGL(k_lower_to_upper, dim3((m + 255) / 256, m), dim3(256), 0, s, m, A, lda);
    hipLaunchKernelGGL(k_tridiag_panel, dim3(1), dim3(kTT), panel_lds, s, m, A, lda, L.d, L.e, L.tau);
  } else {
    tridiag_unblocked(s, m, A, lda, L);
  }
  tridiag_vectors(s, m, want, L, values, Y, ldy, info);
  if (m <= 320) {
    hipLaunchKernelGGL(k_back_transform_reg<5>, dim3((want + 3) / 4), dim3(256), bt_reg_lds(m), s, m, want, A, lda,
                       L.tau, L.Ytmp, int64_t(m));
  } else if (m <= 640) {
    if (bt_reg_lds(m) > 64 * 1024)
      allow_full_lds(reinterpret_cast<const void*>(&k_back_transform_reg<10>), "k_back_transform_reg");
    hipLaunchKernelGGL(k_back_transform_reg<10>, dim3((want + 3) / 4), dim3(256), bt_reg_lds(m), s, m, want, A, lda,
                       L.tau, L.Ytmp, int64_t(m));
  } else if (bt_ch == 8) {
    if (bt_lds(8) > 64 * 1024) allow_full_lds(reinterpret_cast<const void*>(&k_back_transform<8>), "k_back_transform");
    hipLaunchKernelGGL(k_back_transform<8>, dim3((want + 3) / 4), dim3(256), bt_lds(8), s, m, want, A, lda, L.tau,
                       L.Ytmp, int64_t(m));
  } else {
    allow_full_lds(reinterpret_cast<const void*>(&k_back_transform<4>), "k_back_transform");
    hipLaunchKernelGGL(k_back_transform<4>, dim3((want + 3) / 4), dim3(256), bt_lds(4), s, m, want, A, lda, L.tau,
                       L.Ytmp, int64_t(m));
  }
  copy_columns(s, m, want, L.Ytmp, m, Y, ldy);
}

// Lowest `want` eigenpairs of the real tridiagonal (L.d, L.e): values, and
// orthonormal eigenvectors in L.Ytmp (m x want, ld m); Ysc (m x want, ld
// ldy) is scratch for the cluster Gram-Schmidt output.
void tridiag_vectors(hipStream_t s, int m, int want, const EigenLayout& L, double* values, double* Ysc, int64_t ldy,
                     int* info) {
  double* Y = Ysc;
  hipLaunchKernelGGL(k_bisect, dim3((want + 3) / 4), dim3(256), size_t(3 * m) * 8, s, m, want, L.d, L.e, values,
                     L.tnorm);
  auto invit_lds = [m](int v) { return size_t(2 * m + 5 * int64_t(m) * v) * 8 + size_t(m) * v; };
  if (invit_lds(8) <= 160 * 1024) {
    if (invit_lds(8) > 64 * 1024) allow_full_lds(reinterpret_cast<const void*>(&k_invit_lds<8>), "k_invit_lds<8>");
    hipLaunchKernelGGL((k_invit_lds<8>), dim3((want + 7) / 8), dim3(8), invit_lds(8), s, m, want, L.d, L.e, values,
                       L.tnorm, L.f, L.Z);
  } else if (invit_lds(4) <= 160 * 1024) {
    allow_full_lds(reinterpret_cast<const void*>(&k_invit_lds<4>), "k_invit_lds<4>");
    hipLaunchKernelGGL((k_invit_lds<4>), dim3((want + 3) / 4), dim3(4), invit_lds(4), s, m, want, L.d, L.e, values,
                       L.tnorm, L.f, L.Z);
  } else {
    hipLaunchKernelGGL(k_invit, dim3((want + 63) / 64), dim3(64), size_t(2 * m) * 8, s, m, want, L.d, L.e, values,
                       L.tnorm, L.f, L.Z);
  }
  hipLaunchKernelGGL(k_cluster_gs, dim3((want + 3) / 4), dim3(256), 0, s, m, want, values, L.tnorm, L.f, L.Z, Y, ldy,
                     info);
  // Loewdin step: Y <- Y (3 I - Y^T Y) / 2.
  gemm_tn(s, want, want, m, 1.0, Y, ldy, Y, ldy, 0.0, L.G, want, L.gwork);
  hipLaunchKernelGGL(k_loewdin_factor, dim3((want + 255) / 256, want), dim3(256), 0, s, want, L.G);
  gemm_nn(s, m, want, want, 1.0, Y, ldy, L.G, want, 0.0, L.Ytmp, m);
}

// Lowest `want` eigenpairs of the complex Hermitian A (m x m, both triangles,
// interleaved (re, im), ld lda complex entries; overwritten): the complex
// tridiagonalisation, the real tridiagonal stage, the complex back-transform.
// Y: m x want complex, ld ldy complex entries (ldy >= m).
void herm_eigen_lowest(hipStream_t s, int m, int want, double* A, int64_t lda, double* values, double* Y,
                       int64_t ldy, double* work, int* info) {
  if (m <= 0 || want <= 0 || want > m || lda < m || ldy < m || lda * m >= (int64_t(1) << 30))
    fail(BD_ERR_INVALID_PARAMETER, "herm_eigen_lowest: shape");
  if ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(Y)) & 15)
    fail(BD_ERR_INVALID_PARAMETER, "herm_eigen_lowest: A and Y must be 16-byte aligned (complex entries)");
  const size_t zt_lds = size_t(2 * m) * 16 + size_t(2 * (kZT / 64)) * 8;
  const size_t zb_lds = size_t(4 * m) * 16;
  if (size_t(3 * m) * 8 > 64 * 1024 || zt_lds > 160 * 1024 || zb_lds > 160 * 1024)
    fail(BD_ERR_INVALID_PARAMETER, "herm_eigen_lowest: subspace too large for the LDS kernels (m <= 2560)");
  const EigenLayout L = eigen_layout(m, want, work);
  double2* Ac = reinterpret_cast<double2*>(A);
  // Complex tau (m), 16-byte aligned in the across-the-chip scratch (2 m kMB + m + 8 doubles).
  double* td = L.td + ((reinterpret_cast<uintptr_t>(L.td) & 15) ? 1 : 0);
  double2* tc = reinterpret_cast<double2*>(td);
  hipLaunchKernelGGL(k_zero_i, dim3(1), dim3(64), 0, s, info, 1);
  if (zt_lds > 64 * 1024) allow_full_lds(reinterpret_cast<const void*>(&k_zhe_tridiag), "k_zhe_tridiag");
  hipLaunchKernelGGL(k_zhe_tridiag, dim3(1), dim3(kZT), zt_lds, s, m, Ac, lda, L.d, L.e, tc);
  // The real stage's scratch is Y itself (m x want doubles fit its 2 ldy x want).
  tridiag_vectors(s, m, want, L, values, Y, int64_t(m), info);
  if (zb_lds > 64 * 1024) allow_full_lds(reinterpret_cast<const void*>(&k_zback_transform), "k_zback_transform");
  hipLaunchKernelGGL(k_zback_transform, dim3((want + 3) / 4), dim3(256), zb_lds, s, m, want, Ac, lda, tc, L.Ytmp,
                     int64_t(m), reinterpret_cast<double2*>(Y), ldy);
}

void zero_ints(hipStream_t s, int* p, int count) {
  if (count > 0) hipLaunchKernelGGL(k_zero_i, dim3((count + 255) / 256), dim3(256), 0, s, p, count);
}
void zero_doubles(hipStream_t s, double* p, int64_t count) {
  if (count > 0) hipLaunchKernelGGL(k_zero_d, dim3(unsigned((count + 255) / 256)), dim3(256), 0, s, p, count);
}

}  // namespace bd

extern "C" bd_status bd_sym_eigen_multi_from(int32_t* inout, int32_t set) {
  BD_GUARD({
    if (!inout) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    if (set) {
      if (*inout < 3) bd::fail(BD_ERR_INVALID_PARAMETER, "threshold must be >= 3");
      bd::multi_from.store(*inout);
    } else {
      *inout = bd::multi_from.load();
    }
  });
}

extern "C" bd_status bd_sym_eigen_coop(int32_t* inout, int32_t set) {
  BD_GUARD({
    if (!inout) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    if (set) {
      if (inout[0] < -1 || inout[0] > 32 || inout[1] < 3 || inout[2] < inout[1])
        bd::fail(BD_ERR_INVALID_PARAMETER, "workers -1..32 and 3 <= from <= to");
      bd::coop_workers.store(inout[0]);
      bd::coop_from.store(inout[1]);
      bd::coop_to.store(inout[2]);
      bd::coop_spins.store(inout[3]);
    } else {
      inout[0] = bd::coop_workers.load();
      inout[1] = bd::coop_from.load();
      inout[2] = bd::coop_to.load();
      inout[3] = bd::coop_spins.load();
    }
  });
}

// Lowest `want` eigenpairs of a complex Hermitian matrix (both triangles,
// interleaved (re, im), ld lda complex entries; overwritten); Y complex
// m x want (ld ldy complex entries); work: bd_sym_eigen_workspace(m, want)
// doubles.
extern "C" bd_status bd_herm_eigen_dev(int32_t m, int32_t want, double* A, int64_t lda, double* values, double* Y,
                                       int64_t ldy, double* work) {
  BD_GUARD({
    if (!A || !values || !Y || !work) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    int* info = reinterpret_cast<int*>(work + bd::sym_eigen_work(m, want));
    bd::herm_eigen_lowest(nullptr, m, want, A, lda, values, Y, ldy, work, info);
    int bad = 0;
    if (hipMemcpy(&bad, info, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      bd::fail(BD_ERR_DEVICE, "eigensolver launch failed");
    if (bad) bd::fail(BD_ERR_APPLICATION, "eigenvector " + std::to_string(bad - 1) + " lost orthogonality");
  });
}

extern "C" int64_t bd_sym_eigen_workspace(int32_t m, int32_t want) {
  if (m <= 0 || want <= 0 || want > m) return 0;
  return bd::sym_eigen_work(m, want) + 1;  // + the status word
}

extern "C" bd_status bd_sym_eigen_dev(int32_t m, int32_t want, double* A, int64_t lda, double* values, double* Y,
                                      int64_t ldy, double* work) {
  BD_GUARD({
    if (!A || !values || !Y || !work) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    int* info = reinterpret_cast<int*>(work + bd::sym_eigen_work(m, want));
    bd::sym_eigen_lowest(nullptr, m, want, A, lda, values, Y, ldy, work, info);
    int bad = 0;
    if (hipMemcpy(&bad, info, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      bd::fail(BD_ERR_DEVICE, "eigensolver launch failed");
    if (bad < 0) bd::fail(BD_ERR_APPLICATION, "cooperative tridiagonalisation abandoned (a worker waited too long)");
    if (bad) bd::fail(BD_ERR_APPLICATION, "eigenvector " + std::to_string(bad - 1) + " lost orthogonality");
  });
}
