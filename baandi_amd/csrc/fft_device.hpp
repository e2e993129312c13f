// LDS-resident mixed-radix (2,3,4,5,8,9) Stockham FFT for gfx950 workgroups.
//
// A workgroup transforms `nlines` lines of length n held in LDS; element
// (line, i) lives at lds[line * lstride + i * estride]. Each pass reads its
// radix-R butterflies into registers, synchronises, and writes the
// self-sorted outputs back in place, so the whole workgroup must hold every
// butterfly of a pass at once: nthreads * pass_mb(R, nthreads) >= nlines * n / R
// (checked by the host before launch, see fft_fits()).
//
// Convention (crates/dft-numerics/src/lib.rs:46-56): forward uses e^{-i},
// inverse e^{+i}; both unscaled.
#pragma once
#include <hip/hip_runtime.h>

namespace bd {

template <typename T>
struct c2 {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ c2<T> cadd(c2<T> a, c2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ __forceinline__ c2<T> csub(c2<T> a, c2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ __forceinline__ c2<T> cmul(c2<T> a, c2<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ __forceinline__ c2<T> cscale(c2<T> a, T s) { return {a.x * s, a.y * s}; }
// Multiply by -i (forward) or +i (inverse).
template <bool INV, typename T>
__host__ __device__ __forceinline__ c2<T> rot_i(c2<T> a) {
  return INV ? c2<T>{-a.y, a.x} : c2<T>{a.y, -a.x};
}

// Radix plan of one line length: radices in pass order, product n.
struct RadixPlan {
  int n;
  int count;
  int radix[12];
};

// Most butterflies one thread holds per pass, by radix and workgroup size
// (at most 24 complex values per thread at 256 threads). Radices 2-5 hold
// the same lines at 256 and 512 threads (half the butterflies per thread at
// 512); radices 8 and 9 keep 16-18 values per thread at 512 threads.
__host__ __device__ constexpr int pass_mb(int R, int NT) {
  return R == 2 ? 12 * 256 / NT
         : R == 3 ? 8 * 256 / NT
         : R == 4 ? 6 * 256 / NT
         : R == 5 ? 4 * 256 / NT
         : R == 8 ? (NT == 256 ? 3 : 2)
                  : 2;  // R == 9
}

// In-register DFT of length R with root e^{-2 pi i/R} (e^{+} for INV).
template <int R, bool INV, typename T>
__host__ __device__ __forceinline__ void dft(c2<T>* v) {
  if constexpr (R == 2) {
    c2<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    const T h = T(0.86602540378443864676372317075294);  // sin(2 pi/3)
    c2<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    c2<T> t = {v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y};
    c2<T> u = rot_i<INV>(cscale(d, h));
    v[0] = cadd(v[0], s);
    v[1] = cadd(t, u);
    v[2] = csub(t, u);
  } else if constexpr (R == 4) {
    c2<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    c2<T> s13 = cadd(v[1], v[3]), d13 = rot_i<INV>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  } else if constexpr (R == 5) {
    const T c1 = T(0.30901699437494742410229341718282);   // cos(2 pi/5)
    const T c2_ = T(-0.80901699437494742410229341718282); // cos(4 pi/5)
    const T s1 = T(0.95105651629515357211643933337938);   // sin(2 pi/5)
    const T s2 = T(0.58778525229247312916870595463907);   // sin(4 pi/5)
    c2<T> s14 = cadd(v[1], v[4]), d14 = csub(v[1], v[4]);
    c2<T> s23 = cadd(v[2], v[3]), d23 = csub(v[2], v[3]);
    c2<T> a1 = {v[0].x + c1 * s14.x + c2_ * s23.x, v[0].y + c1 * s14.y + c2_ * s23.y};
    c2<T> a2 = {v[0].x + c2_ * s14.x + c1 * s23.x, v[0].y + c2_ * s14.y + c1 * s23.y};
    c2<T> b1 = rot_i<INV>(c2<T>{s1 * d14.x + s2 * d23.x, s1 * d14.y + s2 * d23.y});
    c2<T> b2 = rot_i<INV>(c2<T>{s2 * d14.x - s1 * d23.x, s2 * d14.y - s1 * d23.y});
    v[0] = cadd(v[0], cadd(s14, s23));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  } else if constexpr (R == 8) {
    // X[2m] = DFT4(v_j + v_{j+4})[m], X[2m+1] = DFT4(w^j (v_j - v_{j+4}))[m],
    // w = e^{-2 pi i/8} (e^{+} for INV).
    const T h = T(0.70710678118654752440084436210484904);  // 1/sqrt(2)
    c2<T> a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a[j] = cadd(v[j], v[j + 4]);
      b[j] = csub(v[j], v[j + 4]);
    }
    // w^1 = h(1 -+ i), w^2 = -+i, w^3 = h(-1 -+ i).
    b[1] = INV ? c2<T>{h * (b[1].x - b[1].y), h * (b[1].x + b[1].y)}
               : c2<T>{h * (b[1].x + b[1].y), h * (b[1].y - b[1].x)};
    b[2] = rot_i<INV>(b[2]);
    b[3] = INV ? c2<T>{-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y)}
               : c2<T>{h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y)};
    dft<4, INV>(a);
    dft<4, INV>(b);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      v[2 * m] = a[m];
      v[2 * m + 1] = b[m];
    }
  } else if constexpr (R == 9) {
    // 3 x 3: column DFTs over v[n1 + 3 n2], twiddles w^(n1 k1), row DFTs;
    // X[k1 + 3 k2]. w = e^{-2 pi i/9} (e^{+} for INV).
    const T c1 = T(0.76604444311897803520239265055541667);   // cos(2 pi/9)
    const T s1 = T(0.64278760968653932632264340990726344);   // sin(2 pi/9)
    const T c2_ = T(0.17364817766693034885171662676931480);  // cos(4 pi/9)
    const T s2 = T(0.98480775301220805936674302458952301);   // sin(4 pi/9)
    const T c4 = T(-0.93969262078590838405410927732473147);  // cos(8 pi/9)
    const T s4 = T(0.34202014332566873304409961468225958);   // sin(8 pi/9)
    c2<T> col[3][3];
#pragma unroll
    for (int n1 = 0; n1 < 3; ++n1) {
#pragma unroll
      for (int n2 = 0; n2 < 3; ++n2) col[n1][n2] = v[n1 + 3 * n2];
      dft<3, INV>(col[n1]);  // col[n1][k1]
    }
    const T sg = INV ? T(1) : T(-1);  // w^p = cos + sg i sin
    col[1][1] = cmul(col[1][1], c2<T>{c1, sg * s1});
    col[1][2] = cmul(col[1][2], c2<T>{c2_, sg * s2});
    col[2][1] = cmul(col[2][1], c2<T>{c2_, sg * s2});
    col[2][2] = cmul(col[2][2], c2<T>{c4, sg * s4});
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
      c2<T> row[3] = {col[0][k1], col[1][k1], col[2][k1]};
      dft<3, INV>(row);
#pragma unroll
      for (int k2 = 0; k2 < 3; ++k2) v[k1 + 3 * k2] = row[k2];
    }
  }
}

// One in-place Stockham pass (radix R, Ns = product of earlier radices).
// `roots[k]` = e^{-2 pi i k/n} (forward); the inverse conjugates. A workgroup
// of NT threads holds pass_mb(R, NT) butterflies per thread.
template <int R, bool INV, typename T, int NT>
__device__ __forceinline__ void stockham_pass(c2<T>* lds, int nlines, int lstride, int estride,
                                              int n, int Ns, const c2<T>* roots) {
  static_assert(NT == 256 || NT == 512, "workgroups of 256 or 512 threads");
  constexpr int MB = pass_mb(R, NT);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int m = n / R;
  const int total = nlines * m;
  const int ts = n / (Ns * R);
  c2<T> v[MB][R];
#pragma unroll
  for (int k = 0; k < MB; ++k) {
    const int b = tid + k * nthr;
    if (b < total) {
      const int line = b / m, j = b - line * m;
      const int jm = j % Ns;
      const c2<T>* src = lds + line * lstride;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        c2<T> a = src[(j + r * m) * estride];
        if (r > 0) {
          c2<T> w = roots[jm * r * ts];
          if (INV) w.y = -w.y;
          a = cmul(a, w);
        }
        v[k][r] = a;
      }
      dft<R, INV>(v[k]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < MB; ++k) {
    const int b = tid + k * nthr;
    if (b < total) {
      const int line = b / m, j = b - line * m;
      const int jm = j % Ns;
      const int base = (j / Ns) * Ns * R + jm;
      c2<T>* dst = lds + line * lstride;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[(base + r * Ns) * estride] = v[k][r];
    }
  }
  __syncthreads();
}

// Full 1D transform of every line by a workgroup of NT threads (blockDim.x
// == NT). The caller synchronises before the call; the call ends synchronised.
template <bool INV, typename T, int NT = 256>
__device__ void lds_fft(c2<T>* lds, int nlines, int lstride, int estride, const RadixPlan& rp,
                        const c2<T>* roots) {
  int Ns = 1;
  for (int p = 0; p < rp.count; ++p) {
    const int R = rp.radix[p];
    if (R == 8) stockham_pass<8, INV, T, NT>(lds, nlines, lstride, estride, rp.n, Ns, roots);
    else if (R == 9) stockham_pass<9, INV, T, NT>(lds, nlines, lstride, estride, rp.n, Ns, roots);
    else if (R == 4) stockham_pass<4, INV, T, NT>(lds, nlines, lstride, estride, rp.n, Ns, roots);
    else if (R == 2) stockham_pass<2, INV, T, NT>(lds, nlines, lstride, estride, rp.n, Ns, roots);
    else if (R == 3) stockham_pass<3, INV, T, NT>(lds, nlines, lstride, estride, rp.n, Ns, roots);
    else stockham_pass<5, INV, T, NT>(lds, nlines, lstride, estride, rp.n, Ns, roots);
    Ns *= R;
  }
}

}  // namespace bd
