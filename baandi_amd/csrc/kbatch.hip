// k-point H·X with one workgroup per column: the whole 3D FFT of the column
// stays in LDS (its sticks, then one y-z plane at a time), so a column costs
// one read of X, one read of V (L2-resident, shared by every column) and one
// write of Y. One launch covers every column of every k point of a batch
// (uneven basis sizes included), nonlocal term included. The per-line FFT
// arithmetic, the V multiply and the kinetic
// add are those of the three-stage path (k_kstick_inverse, k_plane,
// k_kstick_forward), so both paths give the same bits for the local part.
//
// Reference: DistributedComplexHamiltonian::apply_block
// (crates/dft-physics/src/distributed_complex.rs:284-386); the reference
// applies one k point at a time (crates/dft-engine/src/refscf/k_map.rs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "op.hpp"

namespace bd {

namespace {

constexpr int kMaxNp2 = 256;  // projector columns (2 p) the fused kernel holds in LDS
// Dynamic LDS of the fused kernel: the 160 KB of a workgroup less q and w.
constexpr size_t kFusedDynMax = 160 * 1024 - 2 * kMaxNp2 * sizeof(double);

template <typename T>
__device__ __forceinline__ void stage_roots_k(c2<T>* dst, const c2<double>* src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = {T(src[i].x), T(src[i].y)};
}

// The k point of work item `item` (works sorted by `first`).
__device__ __forceinline__ int find_work(const KWork* __restrict__ works, int count, int item) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (works[mid].first <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Persistent: workgroup b takes items b, b + grid, ... ; every workgroup
// leaves after its last item.
template <typename F, int NT>
__global__ __launch_bounds__(NT) void k_hpsi_fused(const KWork* __restrict__ works, KWork one, int count, int total,
                                                   int n0,
                                                   int n1, int n2, const c2<double>* __restrict__ roots,
                                                   RadixPlan rx, RadixPlan ry, RadixPlan rz, int xlines, int xplanes,
                                                   int zrows, double inv_n) {
  extern __shared__ char smem[];
  __shared__ double nlq[kMaxNp2], nlw[kMaxNp2];
  c2<F>* rt = reinterpret_cast<c2<F>*>(smem);  // roots x | y | z
  c2<F>* rtx = rt;
  c2<F>* rty = rt + n0;
  c2<F>* rtz = rt + n0 + n1;
  c2<F>* L = rt + n0 + n1 + n2;  // S sticks x n0
  stage_roots_k(rt, roots, n0 + n1 + n2);
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const KWork w = works ? works[find_work(works, count, item)] : one;
    const int j = item - w.first;
    const int S = w.S, Z = w.Z, nk = w.nk;
    c2<F>* P = L + S * n0;            // xplanes x Z x n1
    c2<F>* R = P + xplanes * Z * n1;  // xplanes x zrows x n2
    const double* xc = w.x + int64_t(j) * w.ldx;
    double* yc = w.y + int64_t(j) * w.ldy;
    const F* pot = reinterpret_cast<const F*>(w.potential);
    for (int i = threadIdx.x; i < S * n0; i += NT) L[i] = {F(0), F(0)};
    __syncthreads();
    for (int g = threadIdx.x; g < nk; g += NT) {
      const double2 v = *reinterpret_cast<const double2*>(xc + 2 * int64_t(g));
      L[w.fpos[g]] = {F(v.x), F(v.y)};
    }
    __syncthreads();
    for (int s0 = 0; s0 < S; s0 += xlines) lds_fft<true, F, NT>(L + s0 * n0, min(xlines, S - s0), n0, 1, rx, rtx);
    // y-z planes, `xplanes` at a time: P = [plane][zc][y], R = [plane][row][z].
    for (int x0 = 0; x0 < n0; x0 += xplanes) {
      const int np = min(xplanes, n0 - x0);
      for (int i = threadIdx.x; i < np * Z * n1; i += NT) P[i] = {F(0), F(0)};
      __syncthreads();
      for (int i = threadIdx.x; i < np * S; i += NT) {
        const int pl = i / S, s = i - pl * S;
        const int2 yz = w.yz[s];
        P[(pl * Z + yz.y) * n1 + yz.x] = L[s * n0 + x0 + pl];
      }
      __syncthreads();
      lds_fft<true, F, NT>(P, np * Z, n1, 1, ry, rty);
      for (int y0 = 0; y0 < n1; y0 += zrows) {
        const int c = min(zrows, n1 - y0);
        for (int i = threadIdx.x; i < np * c * n2; i += NT) R[i] = {F(0), F(0)};
        __syncthreads();
        for (int i = threadIdx.x; i < np * c * Z; i += NT) {
          const int pl = i / (c * Z), rem = i - pl * (c * Z);
          const int r = rem / Z, zc = rem - r * Z;
          R[(pl * c + r) * n2 + w.z_of_zc[zc]] = P[(pl * Z + zc) * n1 + y0 + r];
        }
        __syncthreads();
        lds_fft<true, F, NT>(R, np * c, n2, 1, rz, rtz);
        for (int i = threadIdx.x; i < np * c * n2; i += NT) {
          const int pl = i / (c * n2), rem = i - pl * (c * n2);
          R[i] = cscale(R[i], pot[(int64_t(x0 + pl) * n1 + y0) * n2 + rem]);
        }
        __syncthreads();
        lds_fft<false, F, NT>(R, np * c, n2, 1, rz, rtz);
        for (int i = threadIdx.x; i < np * c * Z; i += NT) {
          const int pl = i / (c * Z), rem = i - pl * (c * Z);
          const int r = rem / Z, zc = rem - r * Z;
          P[(pl * Z + zc) * n1 + y0 + r] = R[(pl * c + r) * n2 + w.z_of_zc[zc]];
        }
        __syncthreads();
      }
      lds_fft<false, F, NT>(P, np * Z, n1, 1, ry, rty);
      for (int i = threadIdx.x; i < np * S; i += NT) {
        const int pl = i / S, s = i - pl * S;
        const int2 yz = w.yz[s];
        L[s * n0 + x0 + pl] = P[(pl * Z + yz.y) * n1 + yz.x];
      }
      __syncthreads();
    }
    for (int s0 = 0; s0 < S; s0 += xlines) lds_fft<false, F, NT>(L + s0 * n0, min(xlines, S - s0), n0, 1, rx, rtx);
    // Nonlocal: q = P2^T x (one wave per projector column), w = D q per atom
    // block, then y += P2 w in the gather below (FP64).
    const int np2 = w.np2;
    const int64_t nr = 2 * int64_t(nk);
    if (np2 > 0) {
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      for (int c = wave; c < np2; c += NT / 64) {
        const double* p = w.proj2 + int64_t(c) * nr;
        double acc = 0.0;
        for (int64_t r = lane; r < nr; r += 64) acc += p[r] * xc[r];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (lane == 0) nlq[c] = acc;
      }
      __syncthreads();
      for (int c = threadIdx.x; c < np2; c += NT) {
        const int start = (c / w.nc) * w.nc, a = c - start;
        double v = 0.0;
        for (int b = 0; b < w.nc; ++b) v += w.blockd[a * w.nc + b] * nlq[start + b];
        nlw[c] = v;
      }
      __syncthreads();
    }
    for (int g = threadIdx.x; g < nk; g += NT) {
      const int64_t r = 2 * int64_t(g);
      const c2<F> h = L[w.fpos[g]];
      const double2 v = *reinterpret_cast<const double2*>(xc + r);
      const double k = w.kinetic[r];
      double y0 = double(h.x) * inv_n + k * v.x, y1 = double(h.y) * inv_n + k * v.y;
      if (np2 > 0) {
        double a0 = 0.0, a1 = 0.0;
        for (int c = 0; c < np2; ++c) {
          const double* p = w.proj2 + int64_t(c) * nr + r;
          a0 += p[0] * nlw[c];
          a1 += p[1] * nlw[c];
        }
        y0 += a0;
        y1 += a1;
      }
      *reinterpret_cast<double2*>(yc + r) = make_double2(y0, y1);
    }
    __syncthreads();  // LDS is reused by the next item
  }
}

template <typename F, int NT>
void launch_fused_t(const GammaOp& op, const FusedShape& f, size_t lds, const KWork* works, const KWork& one,
                    int count, int total, hipStream_t s) {
  const StickPlan& p = op.plan;
  // Resident workgroups for this LDS size, looked up once per size (ops may
  // launch from several host threads).
  static std::mutex mu;
  static bool attr = false;
  static size_t cached_lds = 0;
  static int cached_slots = 0;
  int slots = 0;
  {
    std::lock_guard<std::mutex> lock(mu);
    if (!attr) {
      BD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_hpsi_fused<F, NT>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(kFusedDynMax)));
      attr = true;
    }
    if (cached_lds != lds) {
      int dev = 0, cus = 0, per_cu = 0;
      BD_HIP(hipGetDevice(&dev));
      BD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      BD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void*>(&k_hpsi_fused<F, NT>), NT, lds));
      cached_slots = cus * std::max(per_cu, 1);
      cached_lds = lds;
    }
    slots = cached_slots;
  }
  const int grid = std::max(1, std::min(total, slots));
  hipLaunchKernelGGL((k_hpsi_fused<F, NT>), dim3(grid), dim3(NT), lds, s, works, one, count, total, p.n0, p.n1, p.n2,
                     op.d_roots.p, p.rx, p.ry, p.rz, f.xlines, f.xplanes, f.zrows, 1.0 / double(op.N));
  BD_HIP(hipGetLastError());
}
}  // namespace

size_t GammaOp::fused_lds(const FusedShape& f) const {
  const size_t e = fft32() ? sizeof(c2<float>) : sizeof(c2<double>);
  const StickPlan& p = plan;
  return e * size_t(p.n0 + p.n1 + p.n2 + p.S * p.n0 + size_t(f.xplanes) * (p.Z * p.n1 + f.zrows * p.n2));
}

bool GammaOp::fused_valid(const FusedShape& f) const {
  if (!kb) return false;
  if (f.threads != 256 && f.threads != 512) return false;
  if (f.xlines < 1 || f.zrows < 1 || f.xplanes < 1 || f.xplanes > plan.n0) return false;
  if (nproj > kMaxNp2) return false;
  return fft_fits_threads(f.xlines, plan.rx, f.threads) && fft_fits_threads(f.xplanes * plan.Z, plan.ry, f.threads) &&
         fft_fits_threads(f.xplanes * f.zrows, plan.rz, f.threads) && fused_lds(f) <= kFusedDynMax;
}

KWork GammaOp::kwork(const double* x, int64_t ldx, int m, double* y, int64_t ldy, int first) const {
  KWork w{};
  w.x = x;
  w.y = y;
  w.ldx = ldx;
  w.ldy = ldy;
  w.m = m;
  w.first = first;
  w.nk = int(kb->nk);
  w.S = plan.S;
  w.Z = plan.Z;
  w.fpos = d_fpos.p;
  w.yz = d_stick_yz.p;
  w.z_of_zc = d_zc.p;
  w.kinetic = d_kinetic.p;
  w.potential = fft32() ? static_cast<const void*>(d_potential32.p)
                                      : static_cast<const void*>(d_potential.p);
  w.proj2 = d_proj.p;
  w.blockd = d_blockd.p;
  w.np2 = int(nproj);
  w.nc = nc;
  return w;
}

void fused_launch(const GammaOp& op, const FusedShape& f, size_t lds, const KWork* works, const KWork& one, int count,
                  int total, hipStream_t s) {
  if (total <= 0) return;
  if (lds > kFusedDynMax) fail(BD_ERR_INVALID_PARAMETER, "fused kernel LDS exceeds the workgroup limit");
  if (op.fft32()) {
    if (f.threads == 512) launch_fused_t<float, 512>(op, f, lds, works, one, count, total, s);
    else launch_fused_t<float, 256>(op, f, lds, works, one, count, total, s);
  } else {
    if (f.threads == 512) launch_fused_t<double, 512>(op, f, lds, works, one, count, total, s);
    else launch_fused_t<double, 256>(op, f, lds, works, one, count, total, s);
  }
}

// One k point: the work entry travels as a kernel argument.
void GammaOp::fused_apply(const double* x, int64_t ldx, int m, double* y, int64_t ldy) {
  const KWork w = kwork(x, ldx, m, y, ldy, 0);
  fused_launch(*this, fshape, fused_lds(fshape), nullptr, w, 1, m, stream);
}

}  // namespace bd

// ---------------------------------------------------------------------------
// C ABI: batches of k points and the launch-shape autotuner
// ---------------------------------------------------------------------------
struct bd_k_batch {
  std::vector<bd_gamma_op*> ops;
  bd::DevBuf<bd::KWork> works;
  std::vector<bd::KWork> host;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  ~bd_k_batch() {
    if (done) (void)hipEventDestroy(done);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

extern "C" bd_status bd_k_batch_create(bd_gamma_op* const* ops, int32_t count, bd_k_batch** out) {
  BD_GUARD({
    if (!ops || !out || count <= 0) bd::fail(BD_ERR_INVALID_PARAMETER, "empty k batch");
    auto b = std::make_unique<bd_k_batch>();
    const bd::GammaOp& o0 = ops[0]->op;
    for (int i = 0; i < count; ++i) {
      if (!ops[i]) bd::fail(BD_ERR_INVALID_PARAMETER, "null op in k batch");
      const bd::GammaOp& o = ops[i]->op;
      if (!o.kb || !o.fused) bd::fail(BD_ERR_INVALID_PARAMETER, "k batch needs k-point ops on the fused path");
      if (o.sys != o0.sys || o.precision != o0.precision)
        bd::fail(BD_ERR_INVALID_PARAMETER, "k batch ops must share the system and precision");
      b->ops.push_back(ops[i]);
    }
    b->works.alloc(size_t(count));
    b->host.resize(size_t(count));
    BD_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BD_HIP(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    *out = b.release();
  });
}

extern "C" void bd_k_batch_destroy(bd_k_batch* b) { delete b; }

extern "C" void* bd_k_batch_stream(bd_k_batch* b) { return b ? b->stream : nullptr; }

// Y_i = H_i X_i for every op i of the batch: one fused launch for all columns
// of all k points, one nonlocal launch for all k points. Asynchronous on the
// batch stream, ordered after every op's queued work.
extern "C" bd_status bd_k_batch_apply_dev(bd_k_batch* b, const double* const* x, const int64_t* ldx,
                                          const int32_t* m, double* const* y, const int64_t* ldy) {
  BD_GUARD({
    if (!b || !x || !ldx || !m || !y || !ldy) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    const int count = int(b->ops.size());
    int total = 0;
    size_t lds = 0;
    bd::FusedShape f = b->ops[0]->op.fshape;
    for (int i = 0; i < count; ++i) {
      bd::GammaOp& o = b->ops[i]->op;
      if (m[i] < 0 || m[i] > o.capacity) bd::fail(BD_ERR_INVALID_PARAMETER, "column count exceeds capacity");
      if (m[i] > 0 && (!x[i] || !y[i])) bd::fail(BD_ERR_INVALID_PARAMETER, "null column block");
      if (ldx[i] < o.n || ldy[i] < o.n || (ldx[i] & 1) || (ldy[i] & 1) ||
          (reinterpret_cast<uintptr_t>(x[i]) & 15) || (reinterpret_cast<uintptr_t>(y[i]) & 15))
        bd::fail(BD_ERR_INVALID_PARAMETER, "k-point blocks must be 16-byte aligned with even ld >= 2 nk");
      if (!o.has_potential) bd::fail(BD_ERR_INVALID_PARAMETER, "no potential installed");
      // One shape for the launch: the widest x batch and z batch every op accepts.
      if (!o.fused_valid(f)) bd::fail(BD_ERR_INVALID_PARAMETER, "k batch ops need a common fused shape");
      b->host[i] = o.kwork(x[i], ldx[i], m[i], y[i], ldy[i], total);
      lds = std::max(lds, o.fused_lds(f));  // sized for the largest sphere
      total += m[i];
      BD_HIP(hipEventRecord(b->done, o.stream));
      BD_HIP(hipStreamWaitEvent(b->stream, b->done, 0));
    }
    BD_HIP(hipMemcpyAsync(b->works.p, b->host.data(), sizeof(bd::KWork) * count, hipMemcpyHostToDevice, b->stream));
    const bd::KWork none{};
    bd::fused_launch(b->ops[0]->op, f, lds, b->works.p, none, count, total, b->stream);
    // The host copy of the work table must outlive the transfer.
    BD_HIP(hipStreamSynchronize(b->stream));
  });
}

// Time every valid fused shape (threads 256/512, x batches, z batches,
// planes at a time) on m columns of the op's own scratch, reps launches each,
// and keep the fastest. out5 (optional): threads, xlines, zrows, xplanes,
// microseconds per launch.
//
// On the stick/plane path (Gamma ops, or k ops whose column does not fit one
// workgroup) it times k_plane's shapes instead: threads 256/512 and the y
// rows per z-FFT chunk; out5 = threads, chunk, 0, 0, microseconds per apply.
extern "C" bd_status bd_gamma_op_autotune(bd_gamma_op* h, int32_t m, int32_t reps, double* out5) {
  BD_GUARD({
    if (!h) bd::fail(BD_ERR_INVALID_PARAMETER, "null op");
    bd::GammaOp& o = h->op;
    if (!(o.kb && o.fused)) {
      if (m <= 0 || m > o.capacity || reps <= 0) bd::fail(BD_ERR_INVALID_PARAMETER, "autotune columns/reps");
      if (!o.has_potential) bd::fail(BD_ERR_INVALID_PARAMETER, "no potential installed");
      const bd::StickPlan& p = o.plan;
      std::vector<std::pair<int, int>> cands;
      for (int t : {256, 512})
        for (int c = 1; c <= p.n1; ++c) {
          const bool useful = c == p.n1 || (c % 4 == 0) || c == 1;
          if (useful && o.plane_shape_valid(t, c)) cands.push_back({t, c});
        }
      if (cands.empty()) bd::fail(BD_ERR_APPLICATION, "no valid plane shape");
      BD_HIP(hipMemsetAsync(o.d_x.p, 0, size_t(o.n) * m * 8, o.stream));
      hipEvent_t a, z;
      BD_HIP(hipEventCreate(&a));
      BD_HIP(hipEventCreate(&z));
      double best = 1e300;
      std::pair<int, int> pick{o.plane_threads, p.chunk};
      const bool prof = o.profiling;
      o.profiling = false;
      for (const auto& c : cands) {
        o.plane_threads = c.first;
        o.plan.chunk = c.second;
        o.apply(o.d_x.p, o.n, m, o.d_y.p, o.n);  // warm-up
        BD_HIP(hipEventRecord(a, o.stream));
        for (int r = 0; r < reps; ++r) o.apply(o.d_x.p, o.n, m, o.d_y.p, o.n);
        BD_HIP(hipEventRecord(z, o.stream));
        BD_HIP(hipEventSynchronize(z));
        float ms = 0.f;
        BD_HIP(hipEventElapsedTime(&ms, a, z));
        const double us = 1e3 * ms / reps;
        if (us < best) {
          best = us;
          pick = c;
        }
      }
      o.profiling = prof;
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(z);
      o.plane_threads = pick.first;
      o.plan.chunk = pick.second;
      if (out5) {
        out5[0] = pick.first;
        out5[1] = pick.second;
        out5[2] = out5[3] = 0;
        out5[4] = best;
      }
      return BD_OK;
    }
    if (m <= 0 || m > o.capacity || reps <= 0) bd::fail(BD_ERR_INVALID_PARAMETER, "autotune columns/reps");
    if (!o.has_potential) bd::fail(BD_ERR_INVALID_PARAMETER, "no potential installed");
    std::vector<bd::FusedShape> cands;
    for (int t : {256, 512}) {
      std::vector<int> xs, zs;
      for (int v = o.plan.S; v >= 1; v = v > 1 ? (v + 1) / 2 : 0) {
        xs.push_back(v);
        if (v == 1) break;
      }
      for (int v = o.plan.n1; v >= 1; v = v > 1 ? (v + 1) / 2 : 0) {
        zs.push_back(v);
        if (v == 1) break;
      }
      for (int xl : xs)
        for (int zr : zs)
          for (int xp : {1, 2, 3, 4, 5, 6, 8}) {
            bd::FusedShape f{t, xl, zr, xp};
            if (o.fused_valid(f) && cands.size() < 96) cands.push_back(f);
          }
    }
    if (cands.empty()) bd::fail(BD_ERR_APPLICATION, "no valid fused shape");
    // Inputs: the op's staging buffers (contents do not affect the timing).
    BD_HIP(hipMemsetAsync(o.d_x.p, 0, size_t(o.n) * m * 8, o.stream));
    const bd::KWork w = o.kwork(o.d_x.p, o.n, m, o.d_y.p, o.n, 0);
    hipEvent_t a, z;
    BD_HIP(hipEventCreate(&a));
    BD_HIP(hipEventCreate(&z));
    double best = 1e300;
    bd::FusedShape pick = o.fshape;
    for (const auto& f : cands) {
      const size_t lds = o.fused_lds(f);
      bd::fused_launch(o, f, lds, nullptr, w, 1, m, o.stream);  // warm-up
      BD_HIP(hipEventRecord(a, o.stream));
      for (int r = 0; r < reps; ++r) bd::fused_launch(o, f, lds, nullptr, w, 1, m, o.stream);
      BD_HIP(hipEventRecord(z, o.stream));
      BD_HIP(hipEventSynchronize(z));
      float ms = 0.f;
      BD_HIP(hipEventElapsedTime(&ms, a, z));
      const double us = 1e3 * ms / reps;
      if (us < best) {
        best = us;
        pick = f;
      }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(z);
    o.fshape = pick;
    if (out5) {
      out5[0] = pick.threads;
      out5[1] = pick.xlines;
      out5[2] = pick.zrows;
      out5[3] = pick.xplanes;
      out5[4] = best;
    }
  });
}

// Read (set = 0) or set (set = 1) the plane kernel's launch shape on the
// stick/plane path: inout2 = (threads, y rows per z-FFT chunk). Set accepts
// exactly the shapes the autotuner may pick.
extern "C" bd_status bd_gamma_op_plane_shape(bd_gamma_op* h, int32_t* inout2, int32_t set) {
  BD_GUARD({
    if (!h || !inout2) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    bd::GammaOp& o = h->op;
    if (o.kb && o.fused) bd::fail(BD_ERR_INVALID_PARAMETER, "op runs the fused k-point kernel");
    const bd::StickPlan& p = o.plan;
    if (set) {
      const int t = inout2[0], c = inout2[1];
      if (!o.plane_shape_valid(t, c)) bd::fail(BD_ERR_INVALID_PARAMETER, "invalid plane shape");
      o.plane_threads = t;
      o.plan.chunk = c;
    } else {
      inout2[0] = o.plane_threads;
      inout2[1] = p.chunk;
    }
  });
}

// Read (set = 0) or set the plane pass: 0 = the fused one-workgroup plane
// kernel (only where the plane fits), zb > 0 = the split pass with zb
// z-lines per y-transform batch. Setting resets the shape to 256 threads.
extern "C" bd_status bd_gamma_op_plane_split(bd_gamma_op* h, int32_t* inout, int32_t set) {
  BD_GUARD({
    if (!h || !inout) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    bd::GammaOp& o = h->op;
    if (o.kb && o.fused) bd::fail(BD_ERR_INVALID_PARAMETER, "op runs the fused k-point kernel");
    if (!set) {
      inout[0] = o.plan.zb;
      return BD_OK;
    }
    if (inout[0] < 0) bd::fail(BD_ERR_INVALID_PARAMETER, "z-line batch must be >= 0");
    if (inout[0] == 0) {
      const int zb = o.plan.zb;
      o.plan.zb = 0;
      int c = std::min(16, o.plan.n1);
      while (c > 1 && !o.plane_shape_valid(256, c)) --c;
      if (!o.plane_shape_valid(256, c)) {
        o.plan.zb = zb;
        bd::fail(BD_ERR_INVALID_PARAMETER, "the plane does not fit one workgroup: the split pass is required");
      }
      o.plane_threads = 256;
      o.plan.chunk = c;
    } else {
      o.enable_split(inout[0]);
    }
  });
}

// Read or set the fused shape (threads, xlines, zrows, xplanes); set validates it.
// Setting threads = 0 switches the op to the three-stage path; reading gives
// threads = 0 on that path.
extern "C" bd_status bd_gamma_op_fused_shape(bd_gamma_op* h, int32_t* inout4, int32_t set) {
  BD_GUARD({
    if (!h || !inout4) bd::fail(BD_ERR_INVALID_PARAMETER, "null argument");
    bd::GammaOp& o = h->op;
    if (!o.kb) bd::fail(BD_ERR_INVALID_PARAMETER, "not a k-point op");
    if (set && inout4[0] == 0) {
      o.fused = false;
    } else if (set) {
      bd::FusedShape f{inout4[0], inout4[1], inout4[2], inout4[3]};
      if (!o.fused_valid(f)) bd::fail(BD_ERR_INVALID_PARAMETER, "invalid fused shape");
      o.fshape = f;
      o.fused = true;
    } else if (!o.fused) {
      inout4[0] = inout4[1] = inout4[2] = inout4[3] = 0;
    } else {
      inout4[0] = o.fshape.threads;
      inout4[1] = o.fshape.xlines;
      inout4[2] = o.fshape.zrows;
      inout4[3] = o.fshape.xplanes;
    }
  });
}
