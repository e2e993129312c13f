// LDS-resident mixed-radix (2,3,4,5) Stockham FFT for gfx950 workgroups.
//
// A workgroup transforms `nlines` lines of length n held in LDS; element
// (line, i) lives at lds[line * lstride + i * estride]. Each pass reads its
// radix-R butterflies into registers, synchronises, and writes the
// self-sorted outputs back in place, so the whole workgroup must hold every
// butterfly of a pass at once: nthreads * MAXB(R) >= nlines * n / R (checked
// by the host before launch, see fft_fits()).
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
__device__ __forceinline__ c2<T> cadd(c2<T> a, c2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ c2<T> csub(c2<T> a, c2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ c2<T> cmul(c2<T> a, c2<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ c2<T> cscale(c2<T> a, T s) { return {a.x * s, a.y * s}; }
// Multiply by -i (forward) or +i (inverse).
template <bool INV, typename T>
__device__ __forceinline__ c2<T> rot_i(c2<T> a) {
  return INV ? c2<T>{-a.y, a.x} : c2<T>{a.y, -a.x};
}

// Radix plan of one line length: radices in pass order, product n.
struct RadixPlan {
  int n;
  int count;
  int radix[12];
};

// Most butterflies one thread holds per pass, by radix (<= 24 complex values).
template <int R>
struct MaxB;
template <> struct MaxB<2> { static constexpr int v = 12; };
template <> struct MaxB<3> { static constexpr int v = 8; };
template <> struct MaxB<4> { static constexpr int v = 6; };
template <> struct MaxB<5> { static constexpr int v = 4; };

// In-register DFT of length R with root e^{-2 pi i/R} (e^{+} for INV).
template <int R, bool INV, typename T>
__device__ __forceinline__ void dft(c2<T>* v) {
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
  }
}

// One in-place Stockham pass (radix R, Ns = product of earlier radices).
// `roots[k]` = e^{-2 pi i k/n} (forward); the inverse conjugates. A workgroup
// of NT threads holds MaxB<R>::v * 256 / NT butterflies per thread, so the
// lines a pass can hold (and fft_fits) are the same for 256 and 512 threads
// while the wider workgroup needs half the registers.
template <int R, bool INV, typename T, int NT>
__device__ __forceinline__ void stockham_pass(c2<T>* lds, int nlines, int lstride, int estride,
                                              int n, int Ns, const c2<T>* roots) {
  static_assert(MaxB<R>::v * 256 % NT == 0, "butterflies per thread must be whole");
  constexpr int MB = MaxB<R>::v * 256 / NT;
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
    if (R == 4) stockham_pass<4, INV, T, NT>(lds, nlines, lstride, estride, rp.n, Ns, roots);
    else if (R == 2) stockham_pass<2, INV, T, NT>(lds, nlines, lstride, estride, rp.n, Ns, roots);
    else if (R == 3) stockham_pass<3, INV, T, NT>(lds, nlines, lstride, estride, rp.n, Ns, roots);
    else stockham_pass<5, INV, T, NT>(lds, nlines, lstride, estride, rp.n, Ns, roots);
    Ns *= R;
  }
}

}  // namespace bd
