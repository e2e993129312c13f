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


class NotConverged(L.BaandiError):
    def __init__(self, solution, message):
        super().__init__(L.BD_NOT_CONVERGED, message)
        self.solution = solution


def solve_with_initial_hx(operator, start, hamiltonian_diagonal, options: Options, initial_hx=None,
                          out=None, raise_not_converged=False):
    """Solve H X = X E for the lowest ``options.bands`` roots.

    ``start`` (bands, n) and ``hamiltonian_diagonal`` (n,) are CUDA float64
    tensors. Returns a Solution; an exhausted iteration limit is reported in
    ``converged`` (and raised when ``raise_not_converged``), never hidden."""
    import torch
    nb, n = start.shape
    assert nb == options.bands and n == operator.global_len()
    start = start.contiguous()
    diag = hamiltonian_diagonal.contiguous()
    x = torch.empty_like(start) if out is None else out[0]
    hx = torch.empty_like(start) if options.return_hx else None
    if options.return_hx and out is not None:
        hx = out[1]
    o = L.DavidsonOptions(bands=nb, threshold=float(options.threshold), occupied=int(options.occupied),
                          first_step=int(bool(options.first_step)), capacity=int(options.capacity),
                          max_iterations=int(options.max_iterations), return_hx=int(bool(options.return_hx)),
                          preconditioner=int(options.preconditioner),
                          mlp_blob=(L.ptr(options.mlp.blob) if options.mlp is not None else None),
                          mlp_vbar=float(options.mlp_vbar))
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
    if st == L.BD_NOT_CONVERGED and raise_not_converged:
        raise NotConverged(sol, "reference Davidson reached the iteration limit")
    return sol


def solve(operator, start, hamiltonian_diagonal, options: Options, **kw):
    """ref_davidson::solve (:107-116): always applies H to the start."""
    return solve_with_initial_hx(operator, start, hamiltonian_diagonal, options, None, **kw)
