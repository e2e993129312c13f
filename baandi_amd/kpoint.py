"""Host-side mirror of the reference k-point path.

``KPointHamiltonian`` is DistributedComplexHamiltonian
(crates/dft-physics/src/distributed_complex.rs:88-386) at one k point of a
``GammaSystem`` cell, built natively (bd_k_op_create). Complex coefficients
live as interleaved (Re, Im) float64 rows, 2 nk per column: a CUDA complex128
tensor of shape (m, nk) is the same memory as ``torch.view_as_real(x)``
reshaped to (m, 2 nk), which is what the kernels read.

``solve`` is ref_davidson_complex::solve_with_initial_hx
(crates/dft-solvers/src/ref_davidson_complex.rs:641-1113): the GPU complex
Hermitian Davidson on complex columns (Hermitian Grams, the complex reduced
eigenproblem, each Ritz value once). ``solve(..., realified=True)`` keeps the
earlier form, the real Davidson on the realified operator from the pairs
(x_j, i x_j): the real search space stays closed under multiplication by i,
its Ritz values are the complex ones each twice, and the complex eigenvectors
are read back from the Ritz pairs by a complex Gram-Schmidt.

``monkhorst_pack`` and ``fermi_dirac`` follow crates/dft-geometry/src/kpoints.rs
and crates/dft-physics/src/occupations.rs:57-232; ``run_scf_k`` is the
sequential-k map (crates/dft-engine/src/refscf/k_map.rs) over the Gamma
driver's density field and mixer.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np

from . import lib as L
from . import ref_davidson
from .hamiltonian import DistributedRealHamiltonian, GammaSystem


def wrap_fractional(k):
    """kpoints.rs:21-32: each component into [-1/2, 1/2)."""
    out = []
    for x in k:
        x = float(x)
        if not math.isfinite(x):
            raise ValueError("k point must be finite")
        out.append(x if -0.5 <= x < 0.5 else x - math.floor(x + 0.5))
    return tuple(out)


def monkhorst_pack(mesh, shift=(0.0, 0.0, 0.0), time_reversal=True):
    """Unshifted Monkhorst-Pack points (i + 1/2 + shift)/n - 1/2 per axis,
    wrapped, weights 1/(n0 n1 n2); k and -k merged when time_reversal (the
    first of the two, in mesh order, keeps the summed weight)."""
    pts = []
    for i in range(mesh[0]):
        for j in range(mesh[1]):
            for l in range(mesh[2]):
                pts.append(wrap_fractional([(c + s + 0.5) / n - 0.5 for c, s, n in zip((i, j, l), shift, mesh)]))
    weight = 1.0 / len(pts)
    merged = {}
    for p in pts:
        key = tuple(round(x * 1e12) for x in p)
        if time_reversal:
            neg = tuple(round(x * 1e12) for x in wrap_fractional([-x for x in p]))
            if neg in merged:
                merged[neg][1] += weight
                continue
        if key in merged:
            merged[key][1] += weight
        else:
            merged[key] = [p, weight]
    return [(tuple(v[0]), v[1]) for v in merged.values()]


def _fraction_entropy(x):
    """occupations.rs:57-66: fraction and entropy from x = (e - mu)/theta,
    both tails through exp(-|x|)."""
    a = np.abs(x)
    t = np.exp(-a)
    q = t / (1.0 + t)
    p = np.where(x >= 0.0, q, 1.0 / (1.0 + t))
    s = np.log1p(t) + np.where(np.isinf(a), 0.0, a * q)
    return p, s


def fermi_dirac(energies, weights, electrons, theta, capacity=2.0, tolerance=1e-13, max_iterations=256):
    """Occupations (capacity per level), chemical potential and TS (Ha) of
    energies (nk, nb) with k weights: bracket mu by doubling, bisect to a
    relative electron-count tolerance (occupations.rs:71-232)."""
    e = np.asarray(energies, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64)[:, None]

    def number(mu):
        return float((capacity * w * _fraction_entropy((e - mu) / theta)[0]).sum())

    width = theta
    for _ in range(max_iterations):
        lo, hi = float(e.min()) - width, float(e.max()) + width
        if number(lo) <= electrons <= number(hi):
            break
        width *= 2.0
    else:
        raise L.BaandiError(L.BD_ERR_APPLICATION, "unable to bracket the chemical potential")
    mu = None
    for _ in range(max_iterations):
        mid = 0.5 * lo + 0.5 * hi
        count = number(mid)
        if abs(count - electrons) <= tolerance * electrons:
            mu = mid
            break
        if mid == lo or mid == hi:
            break
        if count < electrons:
            lo = mid
        else:
            hi = mid
    if mu is None:
        raise L.BaandiError(L.BD_ERR_APPLICATION, "electron count did not converge")
    p, s = _fraction_entropy((e - mu) / theta)
    return capacity * p, mu, theta * float((capacity * w * s).sum())


class KPointHamiltonian(DistributedRealHamiltonian):
    """H_k on one GPU. Blocks are CUDA tensors of shape (m, 2 nk) float64
    (realified) or (m, nk) complex128 through the *_complex helpers."""

    def __init__(self, system: GammaSystem, k_frac, capacity, precision=L.BD_FP64):
        lib = L.load()
        self.system = system
        self.capacity = int(capacity)
        k = np.ascontiguousarray(np.asarray(k_frac, dtype=np.float64).reshape(3))
        h = C.c_void_p()
        L.check(lib.bd_k_op_create(system.handle, L.ptr(k), self.capacity, int(precision), C.byref(h)))
        self.handle = h
        self.epoch = 0
        nk = C.c_int64(0)
        kf, kc = np.zeros(3), np.zeros(3)
        L.check(lib.bd_k_op_info(h, C.byref(nk), L.ptr(kf), L.ptr(kc)))
        self.nk = int(nk.value)
        self.k_frac = tuple(float(v) for v in kf)
        self.k_cart = kc

    def global_len(self):
        return 2 * self.nk

    def get(self, name):
        p = self.system.nproj
        sizes = {"miller": (3 * self.nk, np.int32), "g2": (self.nk, np.float64), "q2": (self.nk, np.float64),
                 "fft_index": (self.nk, np.int64), "kinetic": (2 * self.nk, np.float64),
                 "diagonal_static": (2 * self.nk, np.float64),
                 "projectors": (4 * self.nk * p, np.float64)}
        count, dtype = sizes[name]
        out = np.empty(count, dtype=dtype)
        L.check(L.load().bd_k_op_get(self.handle, name.encode(), out.ctypes.data_as(C.c_void_p), count))
        if name == "miller":
            out = out.reshape(-1, 3)
        if name == "projectors":
            out = out.reshape(2 * p, 2 * self.nk).T
        return out

    def projectors_complex(self):
        """The complex n x p projector table P (columns at q = k+G)."""
        p2 = self.get("projectors")
        p = self.system.nproj
        return p2[0::2, :p] + 1j * p2[1::2, :p]

    def fused_shape(self):
        """(threads, xlines, zrows, xplanes) of the fused one-workgroup-per-column
        kernel, or (0, 0, 0, 0) on the three-stage path."""
        v = np.zeros(4, dtype=np.int32)
        L.check(L.load().bd_gamma_op_fused_shape(self.handle, L.ptr(v), 0))
        return tuple(int(a) for a in v)

    def set_fused_shape(self, shape):
        """Select a fused shape, or (0, 0, 0, 0) for the three-stage path."""
        v = np.zeros(4, dtype=np.int32)
        v[:len(shape)] = shape
        L.check(L.load().bd_gamma_op_fused_shape(self.handle, L.ptr(v), 1))

    def autotune(self, columns, reps=5):
        """Time every valid fused shape on `columns` columns and keep the
        fastest; returns (threads, xlines, zrows, xplanes, microseconds per launch)."""
        out = super().autotune(columns, reps)
        return int(out[0]), int(out[1]), int(out[2]), int(out[3]), float(out[4])

    def apply_complex(self, x):
        """H x for a CUDA complex128 tensor (m, nk)."""
        import torch
        xr = torch.view_as_real(x.contiguous()).reshape(x.shape[0], -1)
        y = torch.empty_like(xr)
        self.apply_block(xr, y)
        return torch.view_as_complex(y.view(x.shape[0], self.nk, 2))


class KPointBatch:
    """Every column of every k point of `ops` through one fused H·X launch
    (plus one nonlocal launch): bd_k_batch_apply_dev."""

    def __init__(self, ops):
        lib = L.load()
        self.ops = list(ops)
        arr = (C.c_void_p * len(self.ops))(*[op.handle.value for op in self.ops])
        h = C.c_void_p()
        L.check(lib.bd_k_batch_create(arr, len(self.ops), C.byref(h)))
        self.handle = h

    def apply(self, xs, ys):
        """ys[i] = H_i xs[i] for CUDA float64 blocks (m_i, 2 nk_i)."""
        import torch
        n = len(self.ops)
        assert len(xs) == n and len(ys) == n
        xp = (C.c_void_p * n)(*[x.data_ptr() for x in xs])
        yp = (C.c_void_p * n)(*[y.data_ptr() for y in ys])
        ldx = np.array([x.shape[1] for x in xs], dtype=np.int64)
        ldy = np.array([y.shape[1] for y in ys], dtype=np.int64)
        m = np.array([x.shape[0] for x in xs], dtype=np.int32)
        for op, x, y in zip(self.ops, xs, ys):
            assert x.is_contiguous() and y.is_contiguous() and x.shape == y.shape
            assert x.shape[1] == op.global_len()
            op.enter()
        torch.cuda.current_stream().synchronize()
        L.check(L.load().bd_k_batch_apply_dev(self.handle, xp, L.ptr(ldx), L.ptr(m), yp, L.ptr(ldy)))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                L.load().bd_k_batch_destroy(h)
            except (TypeError, AttributeError):
                pass
            self.handle = None


def set_potential_all(ops, potential):
    """The same V(r) (CUDA float64 grid) into every k-point op, one call,
    ordered after torch's current stream (which then waits for the copies)."""
    import torch
    v = potential.reshape(-1).contiguous()
    epoch = max(op.epoch for op in ops) + 1
    arr = (C.c_void_p * len(ops))(*[op.handle.value for op in ops])
    L.check(L.load().bd_k_set_potential_dev(arr, len(ops), L.ptr(v), epoch,
                                            C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    for op in ops:
        op.epoch = epoch


def band_density_all(ops, xs, weights, out):
    """out(grid) = sum over k of the band density of xs[k] ((m, 2 nk) CUDA
    float64, complex columns) with weights[k], summed in k order (one call,
    ordered after torch's current stream; synchronous)."""
    import torch
    count = len(ops)
    arr = lambda vals: (C.c_void_p * count)(*vals)
    ws = [np.ascontiguousarray(np.asarray(w, dtype=np.float64)) for w in weights]
    ld = np.array([x.shape[1] for x in xs], dtype=np.int64)
    m = np.array([x.shape[0] for x in xs], dtype=np.int32)
    L.check(L.load().bd_k_band_density_dev(arr([op.handle.value for op in ops]), count,
                                           arr([x.data_ptr() for x in xs]), L.ptr(ld), L.ptr(m),
                                           arr([w.ctypes.data for w in ws]), L.ptr(out),
                                           C.c_void_p(torch.cuda.current_stream().cuda_stream)))


class KDavidson:
    """The complex Davidson of every k point in lock step (bd_k_davidson_*):
    each k point keeps its own Davidson (ref_davidson_complex.rs:641-1113),
    every stage of an iteration is one launch over all k points and H of all
    new columns one fused launch. Blocks are CUDA float64 tensors (bands,
    2 nk_i), complex columns as interleaved (re, im) rows."""

    def __init__(self, ops, bands, capacity=0):
        self.ops, self.bands, self.capacity = list(ops), int(bands), int(capacity)
        arr = (C.c_void_p * len(self.ops))(*[op.handle.value for op in self.ops])
        h = C.c_void_p()
        L.check(L.load().bd_k_davidson_create(arr, len(self.ops), self.bands, self.capacity, C.byref(h)))
        self.handle = h

    def solve(self, starts, diagonals, options: ref_davidson.Options, outs, initial_hx=None):
        """Solve every k point from ``starts`` into ``outs`` ((x, hx or None)
        per k point); returns one ref_davidson.Solution per k point."""
        import torch
        count = len(self.ops)
        assert len(starts) == count and len(diagonals) == count and len(outs) == count
        assert options.bands == self.bands
        arr = lambda vals: (C.c_void_p * count)(*vals)
        for x, op in zip(starts, self.ops):
            assert x.is_contiguous() and x.shape == (self.bands, op.global_len())
        x0 = arr([x.data_ptr() for x in starts])
        hx0 = arr([h.data_ptr() for h in initial_hx]) if initial_hx is not None else None
        dg = arr([d.contiguous().data_ptr() for d in diagonals])
        xo = arr([o[0].data_ptr() for o in outs])
        hxo = arr([(o[1].data_ptr() if o[1] is not None else None) for o in outs]) if options.return_hx else None
        ld0 = np.array([x.shape[1] for x in starts], dtype=np.int64)
        ldx = np.array([o[0].shape[1] for o in outs], dtype=np.int64)
        opts = L.DavidsonOptions(bands=self.bands, threshold=float(options.threshold), occupied=int(options.occupied),
                                 first_step=int(bool(options.first_step)), capacity=int(options.capacity),
                                 max_iterations=int(options.max_iterations), return_hx=int(bool(options.return_hx)),
                                 preconditioner=int(options.preconditioner), complex_columns=1)
        eig = np.zeros((count, self.bands))
        conv = np.zeros((count, self.bands), dtype=np.int32)
        recs = (L.DavidsonRecord * count)()
        statuses = np.zeros(count, dtype=np.int32)
        # The inputs are complete once torch's stream is; the solver orders
        # itself after the ops' own streams.
        torch.cuda.current_stream().synchronize()
        st = L.load().bd_k_davidson_solve_dev(self.handle, x0, L.ptr(ld0), hx0, dg, C.byref(opts), L.ptr(eig), xo,
                                              hxo, L.ptr(ldx), L.ptr(conv), recs, L.ptr(statuses))
        L.check(st, allow=(L.BD_NOT_CONVERGED,))
        sols = []
        for k in range(count):
            record = {f: getattr(recs[k], f) for f, _ in L.DavidsonRecord._fields_}
            sols.append(ref_davidson.Solution(eigenvalues=eig[k].copy(), eigenvectors=outs[k][0], hx=outs[k][1],
                                              converged=conv[k].astype(bool),
                                              band_thresholds_ha=[options.band_threshold(b)
                                                                  for b in range(self.bands)],
                                              record=record))
        return sols

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                L.load().bd_k_davidson_destroy(h)
            except (TypeError, AttributeError):
                pass
            self.handle = None


class KBlocks:
    """The block iterations of all k points of a map stepped together
    (bd_kblocks_*): per-k work on each op's stream, one fused H·W launch over
    every column of every k point, the iteration captured once as a HIP graph.
    Blocks are CUDA float64 tensors (bands, 2 nk_i) per k point."""

    def __init__(self, ops, bands, preconditioner=L.BD_PRECOND_TPA):
        lib = L.load()
        self.ops = list(ops)
        self.bands = int(bands)
        arr = (C.c_void_p * len(self.ops))(*[op.handle.value for op in self.ops])
        h = C.c_void_p()
        L.check(lib.bd_kblocks_create(arr, len(self.ops), self.bands, int(preconditioner), C.byref(h)))
        self.handle = h
        self.values = np.zeros((len(self.ops), self.bands))

    def _ptrs(self, ts):
        return (C.c_void_p * len(ts))(*[t.data_ptr() for t in ts])

    def _enter(self):
        for op in self.ops:
            op.enter()

    def set_diagonals(self, diags):
        self._enter()
        L.check(L.load().bd_kblocks_set_diagonals_dev(self.handle, self._ptrs([d.contiguous() for d in diags])))

    def start(self, xs):
        """Rayleigh-Ritz of every start block; returns the values (count, bands)."""
        xs = [x.contiguous() for x in xs]
        for op, x in zip(self.ops, xs):
            assert x.shape == (self.bands, op.global_len())
        ld = np.array([x.shape[1] for x in xs], dtype=np.int64)
        self._enter()
        L.check(L.load().bd_kblocks_start_dev(self.handle, self._ptrs(xs), L.ptr(ld), L.ptr(self.values)))
        return self.values.copy()

    def step(self, steps=1, use_graph=True):
        L.check(L.load().bd_kblocks_step(self.handle, int(steps), 1 if use_graph else 0, L.ptr(self.values)))
        return self.values.copy()

    def vectors(self, outs):
        """Copy each k point's Ritz block X into outs[k] (bands, 2 nk)."""
        for op, o in zip(self.ops, outs):
            assert o.is_contiguous() and o.shape == (self.bands, op.global_len())
        ld = np.array([o.shape[1] for o in outs], dtype=np.int64)
        L.check(L.load().bd_kblocks_vectors_dev(self.handle, self._ptrs(outs), None, L.ptr(ld)))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                L.load().bd_kblocks_destroy(h)
            except (TypeError, AttributeError):
                pass
            self.handle = None


def as_real(x):
    """(m, nk) complex128 -> (m, 2 nk) float64 view of the same memory."""
    import torch
    return torch.view_as_real(x.contiguous()).reshape(x.shape[0], -1)


def as_complex(x):
    import torch
    return torch.view_as_complex(x.contiguous().view(x.shape[0], -1, 2))


def realified_start(x):
    """Rows (x_j, i x_j) of a complex start block (bands, nk): the real start of
    the realified Davidson (2 bands, 2 nk)."""
    import torch
    b, nk = x.shape
    pairs = torch.stack([x, 1j * x], dim=1).reshape(2 * b, nk)
    return as_real(pairs).contiguous()


def complex_ritz(values, vectors, bands, keep=0.5):
    """Complex eigenpairs from the realified Ritz block: vectors (2 bands, 2 nk)
    in ascending eigenvalue order are read as complex rows and orthonormalised
    by a two-pass complex Gram-Schmidt; a row whose remainder has norm <= keep
    lies in the complex span of the rows kept before it (its partner i x) and
    is dropped. Returns (values (bands,), vectors (bands, nk) complex128)."""
    import torch
    z = as_complex(vectors)
    kept, vals = [], []
    for j in range(z.shape[0]):
        v = z[j].clone()
        if kept:
            q = torch.stack(kept)
            for _ in range(2):
                v = v - (q.conj() @ v) @ q
        nv = float(torch.linalg.vector_norm(v))
        if nv <= keep:
            continue
        kept.append(v / nv)
        vals.append(float(values[j]))
        if len(kept) == bands:
            break
    if len(kept) != bands:
        raise L.BaandiError(L.BD_ERR_APPLICATION,
                            f"realified Ritz block spans {len(kept)} complex vectors, expected {bands}")
    return np.array(vals), torch.stack(kept)


@dataclass
class KSolution:
    eigenvalues: np.ndarray   # (bands,) complex-problem eigenvalues, Ha
    eigenvectors: object      # (bands, nk) complex128 CUDA
    realified: object         # the underlying ref_davidson.Solution (2 bands if realified)
    converged: np.ndarray
    record: dict


def solve(operator: KPointHamiltonian, start, hamiltonian_diagonal, options: ref_davidson.Options,
          realified=False, initial_hx=None):
    """Lowest ``options.bands`` eigenpairs of H_k from a complex start block
    (bands, nk). ``hamiltonian_diagonal``: (2 nk,) per row. The options are the
    complex problem's. With ``realified`` the real Davidson runs on the
    realified operator with twice the bands, occupied count and capacity."""
    b = options.bands
    if not realified:
        x0 = as_real(start).contiguous()
        hx0 = as_real(initial_hx).contiguous() if initial_hx is not None else None
        sol = ref_davidson.solve_with_initial_hx(operator, x0, hamiltonian_diagonal.contiguous(), options,
                                                 initial_hx=hx0, complex_columns=True)
        return KSolution(sol.eigenvalues, as_complex(sol.eigenvectors), sol, sol.converged, sol.record)
    ro = ref_davidson.Options(bands=2 * b, threshold=options.threshold, occupied=2 * options.occupied,
                              first_step=options.first_step, capacity=2 * options.capacity,
                              max_iterations=options.max_iterations, return_hx=options.return_hx,
                              preconditioner=options.preconditioner, mlp=options.mlp, mlp_vbar=options.mlp_vbar)
    sol = ref_davidson.solve(operator, realified_start(start), hamiltonian_diagonal.contiguous(), ro)
    vals, vecs = complex_ritz(sol.eigenvalues, sol.eigenvectors, b)
    conv = np.logical_and(sol.converged[0::2], sol.converged[1::2])
    return KSolution(vals, vecs, sol, conv, sol.record)


@dataclass
class KScfOptions:
    bands: int                      # complex bands per k point
    theta: float = 0.01             # Fermi-Dirac smearing, Ha
    threshold_ry: float = 1e-8      # density residual (refscf/profile.rs:17-26)
    beta: float = 0.7
    history: int = 8
    max_iterations: int = 100
    davidson_max_iterations: int = 100
    preconditioner: int = L.BD_PRECOND_TPA
    precision: int = L.BD_FP64
    seed: int = 1
    threads: int = 8                # k points solved concurrently (host threads, one stream each)
    autotune: bool = False          # tune the fused kernel's launch shape at set-up (bit-neutral)
    fused_shape: tuple = None       # (threads, xlines, zrows, xplanes): fix the fused shape instead
    preconditioning: str = "plain"  # density mixing: "plain" or "thomas_fermi"
    # "davidson": the reference Davidson per k point, `threads` at a time;
    # "block": the block iterations of all k points stepped together (one
    # fused H·W launch for every column of every k point, the iteration
    # replayed from a captured HIP graph) until the largest Ritz value change
    # is below the map's threshold (realified columns only).
    solver: str = "davidson"
    # "complex": the complex Hermitian Davidson on `bands` complex columns
    # (ref_davidson_complex.rs); "realified": the real Davidson on 2 bands
    # realified columns (the pairs (x, i x)).
    columns: str = "complex"
    # Complex columns: the Davidsons of all k points in lock step (one launch
    # per stage for all k points, KDavidson) where the ops allow it (fused
    # path, bands <= 24); False: one solve per k point, `threads` at a time.
    lockstep: bool = True


def _complex_start(nk, bands, seed):
    """Complex N(0,1) columns, each normalised (seeded, on the host)."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    x = torch.complex(torch.randn(bands, nk, generator=g, dtype=torch.float64),
                      torch.randn(bands, nk, generator=g, dtype=torch.float64))
    return x / torch.linalg.vector_norm(x, dim=1, keepdim=True)


def run_scf_k(system: GammaSystem, kpoints, options: KScfOptions, log=None):
    """Sequential-k SCF (refscf/k_map.rs) to dr2 < threshold_ry with
    Fermi-Dirac occupations; returns the Mermin free energy terms (Ha),
    eigenvalues (nk, bands), occupations, mu and timings.

    Each k point's Davidson is the complex one on `bands` complex columns
    (levels of capacity 2), or with ``columns="realified"`` the realified one
    (2 bands real columns, the pairs (x, i x) of the complex start): a real
    Ritz level then holds one electron (its pair partner the other), so the
    occupations, the density weights and the entropy are taken over the real
    levels with capacity 1."""
    import torch
    from .field import DensityField, Mixer, gaussian_start
    from .scf import next_ethr_ry
    dev = torch.device("cuda")
    t_setup = time.perf_counter()
    b = int(options.bands)
    electrons = system.electrons
    weights = np.array([w for _, w in kpoints], dtype=np.float64)
    field = DensityField(system)
    vion_g = torch.from_numpy(system.get("ionic_density_re") + 1j * system.get("ionic_density_im")).to(dev)
    vion = field.to_grid(vion_g).real.contiguous()
    if options.solver not in ("davidson", "block"):
        raise ValueError(f"unknown k-point solver {options.solver!r}")
    if options.columns not in ("complex", "realified"):
        raise ValueError(f"unknown k-point columns {options.columns!r}")
    cplx = options.columns == "complex" and options.solver == "davidson"
    if options.solver == "block" and options.columns != "realified":
        raise ValueError("the block k-point solver runs on realified columns (columns='realified')")
    nb = b if cplx else 2 * b          # Davidson columns per k point
    level_capacity = 2.0 if cplx else 1.0
    ops = [KPointHamiltonian(system, k, capacity=nb, precision=options.precision) for k, _ in kpoints]
    if options.fused_shape is not None:
        for op in ops:
            op.set_fused_shape(tuple(options.fused_shape))
    elif options.autotune and ops[0].fused_shape()[0] > 0:
        # Tuned on the first k point with the ionic potential, then shared by
        # every k point that accepts it (bases differ by a few coefficients).
        ops[0].set_owned_potential(vion.reshape(-1).contiguous())
        shape = ops[0].autotune(nb, reps=3)[:4]
        for op in ops[1:]:
            try:
                op.set_fused_shape(shape)
            except L.BaandiError:
                pass
    # The static diagonals of all k points in one tensor (one add of <V> per map).
    diag_flat = torch.from_numpy(np.concatenate([op.get("diagonal_static") for op in ops])).to(dev)
    diag_now = torch.empty_like(diag_flat)
    diags, at = [], 0
    for op in ops:
        diags.append(diag_now[at:at + 2 * op.nk])
        at += 2 * op.nk
    blocks = KBlocks(ops, 2 * b, options.preconditioner) if options.solver == "block" else None
    lockstep = None
    if cplx and options.lockstep and b <= 24 and all(op.fused_shape()[0] > 0 for op in ops) \
            and 2 * b <= min(op.nk for op in ops):
        try:
            lockstep = KDavidson(ops, b)
        except L.BaandiError:
            lockstep = None  # no common fused shape: one solve per k point
    kinetic = [torch.from_numpy(op.get("kinetic")).to(dev) for op in ops]
    ewald = float(system.get("ewald")[0])
    x = torch.from_numpy(gaussian_start(system)).to(dev)
    mixer = Mixer(field, options.beta, options.history, options.preconditioning)
    start = as_real if cplx else realified_start
    orbitals = [start(_complex_start(op.nk, b, options.seed + i).to(dev)).contiguous() for i, op in enumerate(ops)]
    outs = [torch.empty_like(o) for o in orbitals]
    rho_dev = torch.empty(field.size, dtype=torch.float64, device=dev)
    potential = torch.empty_like(rho_dev)
    fres = torch.empty_like(x)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup

    t0 = time.perf_counter()
    ethr = 1e-2
    last_dr2 = math.inf
    history = []
    totals = dict(h_columns=0, davidson_iterations=0, dense_s=0.0, apply_s=0.0, davidson_s=0.0, fields_s=0.0,
                  eigen_retries=0)
    result = None
    for iteration in range(1, options.max_iterations + 1):
        if iteration > 1:
            ethr = next_ethr_ry(1e-2 if iteration == 2 else ethr, last_dr2, electrons)
        redo = False
        hcol0, dav0 = totals["h_columns"], totals["davidson_iterations"]
        while True:
            tf = time.perf_counter()
            vmean = field.potential(x, vion, potential)  # V_ion + V_H + v_xc of the input density
            set_potential_all(ops, potential)
            totals["fields_s"] += time.perf_counter() - tf
            td = time.perf_counter()
            opts = ref_davidson.Options(bands=nb, threshold=ethr / 2.0, occupied=nb,
                                        first_step=(iteration == 1 and not redo),
                                        max_iterations=options.davidson_max_iterations,
                                        preconditioner=options.preconditioner)
            torch.add(diag_flat, vmean, out=diag_now)
            if blocks is None:
                if lockstep is not None:
                    sols = lockstep.solve(orbitals, diags, opts, [(o, None) for o in outs])
                else:
                    sols = ref_davidson.solve_many(ops, orbitals, diags, [opts] * len(ops),
                                                   [(o, None) for o in outs], threads=options.threads,
                                                   complex_columns=cplx)
                for sol in sols:
                    totals["h_columns"] += sol.record["h_columns"]
                    totals["eigen_retries"] += sol.record["eigen_retries"]
                    totals["davidson_iterations"] += sol.record["iterations"]
                    totals["dense_s"] += sol.record["seconds_dense"]
                    totals["apply_s"] += sol.record["seconds_apply"]
                levels = np.stack([s.eigenvalues for s in sols])
                unconverged = sum(int((~s.converged).sum()) for s in sols)
            else:
                blocks.set_diagonals(diags)
                prev = blocks.start(orbitals)
                steps, change = 0, math.inf
                while steps < options.davidson_max_iterations and change >= ethr / 2.0:
                    levels = blocks.step(1)
                    change = float(np.abs(levels - prev).max())
                    prev = levels
                    steps += 1
                levels = prev
                blocks.vectors(outs)
                unconverged = 0 if change < ethr / 2.0 else int(levels.size)
                totals["h_columns"] += len(ops) * 2 * b * (steps + 1)
                totals["davidson_iterations"] += steps
            totals["davidson_s"] += time.perf_counter() - td
            tf = time.perf_counter()
            occ, mu, ts = fermi_dirac(levels, weights, electrons, options.theta, capacity=level_capacity)
            band_density_all(ops, outs, [weights[i] * occ[i] for i in range(len(ops))], rho_dev)
            rho_out = rho_dev.view(field.dims)
            dr2 = field.residual(rho_out, x, fres)
            totals["fields_s"] += time.perf_counter() - tf
            if iteration == 1 and not redo and dr2 < ethr * max(electrons, 1.0):
                ethr = next_ethr_ry(ethr, dr2, electrons)
                redo = True
                orbitals, outs = outs, orbitals
                continue
            break
        orbitals, outs = outs, orbitals
        last_dr2 = dr2
        history.append(dict(iteration=iteration, dr2_ry=dr2, ethr_ry=ethr, mu=mu,
                            davidson_iterations=totals["davidson_iterations"] - dav0,
                            h_columns=totals["h_columns"] - hcol0, unconverged_levels=unconverged))
        if log:
            log(f"kscf {iteration:3d} dr2={dr2:.3e} ethr={ethr:.2e} mu={mu:.8f} "
                f"hcols={history[-1]['h_columns']} unconverged={unconverged}")
        if dr2 < options.threshold_ry:
            result = (levels, occ, mu, ts, rho_out)
            break
        x = mixer.update(x, fres)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if result is None:
        raise RuntimeError(f"k SCF did not reach dr2 < {options.threshold_ry} Ry in {options.max_iterations} maps")
    levels, occ, mu, ts, rho_out = result
    kin = nl = 0.0
    for i, op in enumerate(ops):
        w = weights[i] * occ[i]
        xo = orbitals[i]
        kin += float(((xo * xo * kinetic[i][None, :]).sum(dim=1).cpu().numpy() * w).sum())
        nl += float(np.dot(op.nonlocal_band_energies(xo), w))
    hartree, xc, local = field.energies(rho_out, vion)
    total = kin + nl + local + hartree + xc + ewald - ts
    energy = {"kinetic": kin, "nonlocal": nl, "local": local, "hartree": hartree, "xc": xc, "ewald": ewald,
              "entropy_term": ts, "total": total}
    if cplx:
        eig, occupations = levels, occ
    else:
        eig, occupations = 0.5 * (levels[:, 0::2] + levels[:, 1::2]), occ[:, 0::2] + occ[:, 1::2]
    return dict(energy=energy, eigenvalues=eig, levels=levels, occupations=occupations, mu=mu,
                history=history, iterations=len(history), wall_s=wall, setup_s=t_setup, totals=totals,
                kpoints=kpoints, ops=ops, orbitals=orbitals)
