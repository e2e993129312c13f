"""Host-side mirror of ``ref_davidson`` (crates/dft-solvers/src/ref_davidson.rs):
``Options`` (:38-78), ``Solution`` (:80-97) and ``solve_with_initial_hx``
(:117-124) with the same argument meaning, over the GPU operator.

ORC blocks are CUDA float64 tensors of shape (bands, n) (each row one ORC
column). The solve itself runs in libbaandi_amd.so (bd_davidson_solve_dev).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import lib as L


@dataclass
class Options:
    bands: int
    threshold: float          # occupied Ritz-change threshold, Ha (strict)
    occupied: int
    first_step: bool = False
    capacity: int = 0         # 0: min(2*bands, n)
    max_iterations: int = 20
    return_hx: bool = False
    preconditioner: int = L.BD_PRECOND_QE  # BD_PRECOND_QE, BD_PRECOND_TPA or BD_PRECOND_MLP
    mlp: object = None        # BD_PRECOND_MLP: a baandi_amd.learned.Network
    mlp_vbar: float = 0.0     # BD_PRECOND_MLP: mean local potential (Ha)
    profile: bool = False     # record the per-iteration Ritz history (ref_davidson.rs:52-54)

    def band_threshold(self, band):
        """ref_davidson.rs:70-77."""
        if self.first_step or band < self.occupied:
            return self.threshold
        return max(5.0 * self.threshold, 5e-6)


@dataclass
class Solution:
    eigenvalues: np.ndarray
    eigenvectors: object      # CUDA tensor (bands, n)
    hx: object                # CUDA tensor or None
    converged: np.ndarray
    band_thresholds_ha: list
    record: dict = field(default_factory=dict)
    # Per correction (iteration) and band: |Ritz change| (Ha) and the
    # (previous, current) Ritz values; empty unless Options.profile.
    ritz_changes_ha: list = field(default_factory=list)
    ritz_values_ha: list = field(default_factory=list)


class NotConverged(L.BaandiError):
    def __init__(self, solution, message):
        super().__init__(L.BD_NOT_CONVERGED, message)
        self.solution = solution


def solve_with_initial_hx(operator, start, hamiltonian_diagonal, options: Options, initial_hx=None,
                          out=None, raise_not_converged=False, complex_columns=False):
    """Solve H X = X E for the lowest ``options.bands`` roots.

    ``start`` (bands, n) and ``hamiltonian_diagonal`` (n,) are CUDA float64
    tensors. Returns a Solution; an exhausted iteration limit is reported in
    ``converged`` (and raised when ``raise_not_converged``), never hidden.
    ``complex_columns``: each row of ``start`` is a complex vector of n / 2
    interleaved (re, im) pairs of a k-point operator, solved by the complex
    Hermitian Davidson (ref_davidson_complex.rs:641-1113)."""
    import torch
    nb, n = start.shape
    assert nb == options.bands and n == operator.global_len()
    start = start.contiguous()
    diag = hamiltonian_diagonal.contiguous()
    x = torch.empty_like(start) if out is None else out[0]
    hx = torch.empty_like(start) if options.return_hx else None
    if options.return_hx and out is not None:
        hx = out[1]
    hist_c = hist_v = None
    if options.profile:
        hist_c = np.zeros((options.max_iterations, nb))
        hist_v = np.zeros((options.max_iterations, nb, 2))
    o = L.DavidsonOptions(bands=nb, threshold=float(options.threshold), occupied=int(options.occupied),
                          first_step=int(bool(options.first_step)), capacity=int(options.capacity),
                          max_iterations=int(options.max_iterations), return_hx=int(bool(options.return_hx)),
                          preconditioner=int(options.preconditioner),
                          mlp_blob=(L.ptr(options.mlp.blob) if options.mlp is not None else None),
                          mlp_vbar=float(options.mlp_vbar), profile=int(bool(options.profile)),
                          ritz_changes_ha=(hist_c.ctypes.data if hist_c is not None else None),
                          ritz_values_ha=(hist_v.ctypes.data if hist_v is not None else None),
                          complex_columns=int(bool(complex_columns)))
    if options.preconditioner == L.BD_PRECOND_MLP and options.mlp is None:
        raise ValueError("BD_PRECOND_MLP needs a trained network (Options.mlp)")
    eig = np.zeros(nb)
    conv = np.zeros(nb, dtype=np.int32)
    rec = L.DavidsonRecord()
    operator.enter()
    st = L.load().bd_davidson_solve_dev(
        operator.handle, L.ptr(start), n, L.ptr(initial_hx) if initial_hx is not None else None,
        L.ptr(diag), C.byref(o), L.ptr(eig), L.ptr(x), L.ptr(hx) if hx is not None else None, n,
        L.ptr(conv), C.byref(rec))
    operator.leave()
    L.check(st, allow=(L.BD_NOT_CONVERGED,))
    record = {f: getattr(rec, f) for f, _ in L.DavidsonRecord._fields_}
    sol = Solution(eigenvalues=eig, eigenvectors=x, hx=hx, converged=conv.astype(bool),
                   band_thresholds_ha=[options.band_threshold(i) for i in range(nb)], record=record)
    if options.profile:
        its = record["iterations"]
        sol.ritz_changes_ha = [hist_c[i].copy() for i in range(its)]
        sol.ritz_values_ha = [hist_v[i].copy() for i in range(its)]
    if st == L.BD_NOT_CONVERGED and raise_not_converged:
        raise NotConverged(sol, "reference Davidson reached the iteration limit")
    return sol


def solve(operator, start, hamiltonian_diagonal, options: Options, **kw):
    """ref_davidson::solve (:107-116): always applies H to the start."""
    return solve_with_initial_hx(operator, start, hamiltonian_diagonal, options, None, **kw)


def solve_many(operators, starts, diagonals, options, outs, threads=8, complex_columns=False):
    """Independent solves on distinct operators (the k points of one map), up
    to ``threads`` at a time on the GPU, each on its operator's stream
    (bd_davidson_solve_many_dev). ``options``: one Options per solve;
    ``outs``: (x, hx or None) per solve; ``complex_columns`` as for
    solve_with_initial_hx. Returns the Solutions."""
    import torch
    count = len(operators)
    assert len(starts) == count and len(diagonals) == count and len(options) == count and len(outs) == count
    arr = lambda vals: (C.c_void_p * count)(*vals)
    ops = arr([op.handle.value for op in operators])
    x0 = arr([s.data_ptr() for s in starts])
    dg = arr([d.data_ptr() for d in diagonals])
    xo = arr([o[0].data_ptr() for o in outs])
    hx_any = any(o[1] is not None for o in outs)
    hxo = arr([(o[1].data_ptr() if o[1] is not None else None) for o in outs]) if hx_any else None
    ld0 = np.array([s.shape[1] for s in starts], dtype=np.int64)
    ldx = np.array([o[0].shape[1] for o in outs], dtype=np.int64)
    opts = (L.DavidsonOptions * count)()
    eigs = [np.zeros(o.bands) for o in options]
    convs = [np.zeros(o.bands, dtype=np.int32) for o in options]
    for i, o in enumerate(options):
        assert starts[i].shape[0] == o.bands and starts[i].is_contiguous() and diagonals[i].is_contiguous()
        if o.preconditioner == L.BD_PRECOND_MLP and o.mlp is None:
            raise ValueError("BD_PRECOND_MLP needs a trained network (Options.mlp)")
        opts[i] = L.DavidsonOptions(bands=o.bands, threshold=float(o.threshold), occupied=int(o.occupied),
                                    first_step=int(bool(o.first_step)), capacity=int(o.capacity),
                                    max_iterations=int(o.max_iterations), return_hx=int(bool(o.return_hx)),
                                    preconditioner=int(o.preconditioner),
                                    mlp_blob=(L.ptr(o.mlp.blob).value if o.mlp is not None else None),
                                    mlp_vbar=float(o.mlp_vbar), complex_columns=int(bool(complex_columns)))
    ev = arr([e.ctypes.data for e in eigs])
    cv = arr([c.ctypes.data for c in convs])
    recs = (L.DavidsonRecord * count)()
    statuses = np.zeros(count, dtype=np.int32)
    for op in operators:
        op.enter()
    torch.cuda.current_stream().synchronize()
    st = L.load().bd_davidson_solve_many_dev(count, ops, x0, L.ptr(ld0), dg, opts, ev, xo, hxo, L.ptr(ldx), cv,
                                            recs, L.ptr(statuses), int(threads))
    for op in operators:
        op.leave()
    L.check(st, allow=(L.BD_NOT_CONVERGED,))
    sols = []
    for i, o in enumerate(options):
        record = {f: getattr(recs[i], f) for f, _ in L.DavidsonRecord._fields_}
        sols.append(Solution(eigenvalues=eigs[i], eigenvectors=outs[i][0], hx=outs[i][1],
                             converged=convs[i].astype(bool),
                             band_thresholds_ha=[o.band_threshold(b) for b in range(o.bands)], record=record))
    return sols


class BlockIteration:
    """Fixed-shape block iteration on one operator (bd_block_*): every band
    active, Rayleigh-Ritz on [X, W] each step, restart to the Ritz block. With
    ``graph`` the step is captured once as a HIP graph and replayed."""

    def __init__(self, operator, bands, preconditioner=L.BD_PRECOND_TPA, graph=True):
        self.operator, self.bands, self.graph = operator, int(bands), bool(graph)
        h = C.c_void_p()
        L.check(L.load().bd_block_create(operator.handle, self.bands, int(preconditioner), C.byref(h)))
        self.handle = h
        self.values = np.zeros(self.bands)

    def set_diagonal(self, diag):
        self.operator.enter()
        L.check(L.load().bd_block_set_diagonal_dev(self.handle, L.ptr(diag.contiguous())))

    def start(self, x0):
        """Rayleigh-Ritz of the (bands, n) start block; returns the values."""
        x0 = x0.contiguous()
        self.operator.enter()
        L.check(L.load().bd_block_start_dev(self.handle, L.ptr(x0), x0.shape[1], L.ptr(self.values)))
        return self.values.copy()

    def step(self, steps=1):
        self.operator.enter()
        L.check(L.load().bd_block_step(self.handle, int(steps), int(self.graph), L.ptr(self.values)))
        return self.values.copy()

    def step_async(self, steps=1):
        self.operator.enter()
        L.check(L.load().bd_block_step_async(self.handle, int(steps), int(self.graph)))

    def finish(self):
        L.check(L.load().bd_block_finish(self.handle, L.ptr(self.values)))
        return self.values.copy()

    def vectors(self, x, hx=None):
        """Copy X (and HX) into (bands, n) CUDA tensors."""
        L.check(L.load().bd_block_vectors_dev(self.handle, L.ptr(x), L.ptr(hx) if hx is not None else None,
                                              x.shape[1]))
        self.operator.leave()

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                L.load().bd_block_destroy(h)
            except (TypeError, AttributeError):
                pass
            self.handle = None


def solve_block(operator, start, hamiltonian_diagonal, bands, threshold, max_steps=200,
                preconditioner=L.BD_PRECOND_TPA, graph=True, check_every=1, converge_bands=None):
    """Block iterations until each of the lowest ``converge_bands`` Ritz
    values (all by default; the others are a buffer) changes by less than
    ``threshold`` (Ha) in one step; returns (values, X, steps)."""
    import torch
    it = BlockIteration(operator, bands, preconditioner, graph)
    it.set_diagonal(hamiltonian_diagonal)
    old = it.start(start)
    steps = 0
    while steps < max_steps:
        try:
            new = it.step(check_every)
        except L.BaandiError as e:
            raise L.BaandiError(e.status, f"{e} (after {steps} steps, values {old})") from None
        steps += check_every
        k = bands if converge_bands is None else int(converge_bands)
        if np.abs(new[:k] - old[:k]).max() < threshold:
            break
        old = new
    x = torch.empty_like(start)
    it.vectors(x)
    return it.values.copy(), x, steps
