"""Oracle: the reference block Davidson on ORC columns.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows crates/dft-solvers/src/ref_davidson.rs:117-366 with the workspace of
ref_davidson/workspace.rs, the streaming correction of ref_davidson/stream.rs
and the bordered inverse Cholesky of lean_davidson.rs:164-444. Contraction
(workspace.rs:772-935) takes the ideal reset (S = I, H_r = diag(theta)) when
the restarted columns have unit norms to 1e-12, a measured refresh every
fourth contraction, and measured refreshes for the rest of the solve once a
guard trips (ref_davidson/ideal_reset.rs:7-8).

A complex start block runs the same iteration on complex columns, which is
ref_davidson_complex::solve_with_initial_hx
(crates/dft-solvers/src/ref_davidson_complex.rs:641-1113): Hermitian Grams
(V^H B), the complex Hermitian reduced problem, |x|^2 weights. On real input
every conjugation is the identity and the arithmetic is the real solver's.

Parity: converged eigenvalues against a dense LAPACK eigensolve of the dense
ORC H (<=1e-11 Ha, the tolerance of crates/dft-solvers/tests/ref_davidson.rs:223)
and of the dense H_k for complex columns, and the denominator identity
(stream.rs:16-22 vs ref_davidson.rs:418-422).
"""
from __future__ import annotations

import math

import numpy as np

EPS = np.finfo(np.float64).eps
ORTHOGONALITY_GATE = 1e-12  # lean_davidson.rs:62
PIVOT_RATIO_MIN = EPS / ORTHOGONALITY_GATE  # lean_davidson.rs:64
DEPENDENCE_TOLERANCE = 1e-12  # crates/dft-numerics/src/linalg.rs:12
ORTHOGONALITY_LIMIT = 1e-12  # ref_davidson/ideal_reset.rs:7
REFRESH_INTERVAL = 4  # ref_davidson/ideal_reset.rs:8


def sqrt_denominator(h, e):
    """QE smooth denominator in Ha (stream.rs:16-22)."""
    x = 2.0 * (h - e)
    z = np.abs(x - 1.0)
    return 0.25 * (2.0 * np.maximum(x, 1.0) + 1.0 / (np.sqrt(1.0 + z * z) + z))


def preconditioner_denominator(h, e):
    """Hypot form of the same denominator (ref_davidson.rs:418-422)."""
    x = 2.0 * (h - e)
    z = np.abs(x - 1.0)
    return 0.25 * (np.where(x >= 1.0, 2.0 * x, 2.0) + 1.0 / (np.hypot(z, 1.0) + z))


def kinetic_scale(weights, kinetic):
    """Weighted mean kinetic energy with the 0.1 Ha floor
    (crates/dft-solvers/src/lib.rs:178-193)."""
    s = float(np.dot(weights, kinetic) / np.sum(weights))
    return max(s, 0.1)


def tpa_weight(kinetic, scale):
    """Teter-Payne-Allan K(x), x = T_G/<T> (lib.rs:197-207)."""
    kinetic = np.asarray(kinetic, dtype=np.float64)
    x = kinetic / scale
    out = np.empty_like(x)
    lo = x <= 1.0
    xl = x[lo]
    p = 27.0 + xl * (18.0 + xl * (12.0 + xl * 8.0))
    out[lo] = p / (p + 16.0 * xl ** 4)
    y = scale / kinetic[~lo]
    p = y * (8.0 + y * (12.0 + y * (18.0 + y * 27.0)))
    out[~lo] = p / (p + 16.0)
    return out


class NotConverged(Exception):
    pass


def _h(a):
    """Conjugate transpose (the transpose on real arrays)."""
    return a.conj().T


class Davidson:
    def __init__(self, apply, diag, bands, threshold, occupied, first_step=False,
                 capacity=0, max_iterations=20, preconditioner="qe", kinetic=None, ideal_reset=True):
        self.apply = apply
        self.diag = np.asarray(diag, dtype=np.float64)
        self.n = bands
        self.threshold = threshold
        self.occupied = occupied
        self.first_step = first_step
        dim = len(self.diag)
        self.cap = min(2 * bands if capacity == 0 else capacity, dim)
        self.max_iterations = max_iterations
        self.preconditioner = preconditioner
        self.kinetic = kinetic
        self.h_calls = 0
        self.h_columns = 0
        self.repairs = 0
        self.restarts = 0
        self.ideal_reset = ideal_reset
        self.ideal_resets = 0
        self.periodic_refreshes = 0
        self.guard_fallbacks = 0

    def band_threshold(self, i):
        """Options::band_threshold (ref_davidson.rs:70-77)."""
        if self.first_step or i < self.occupied:
            return self.threshold
        return max(5.0 * self.threshold, 5e-6)

    # -- algebra helpers ---------------------------------------------------
    def _h(self, x):
        self.h_calls += 1
        self.h_columns += x.shape[1]
        y = self.apply(x)
        if not np.all(np.isfinite(y)):
            raise FloatingPointError("reference H application failed")
        return y

    def _factor(self, S, m_from, m_to, T):
        """Bordered inverse Cholesky of S[:m_to,:m_to] (extend_factor,
        lean_davidson.rs:167-331); returns (ok, T) with T = L^{-1}."""
        Sm = S[:m_to, :m_to]
        try:
            L = np.linalg.cholesky(Sm)
        except np.linalg.LinAlgError:
            return False, T
        piv = np.diag(L).real ** 2 / np.diag(Sm).real
        if np.any(~np.isfinite(piv[m_from:])) or np.any(piv[m_from:] < PIVOT_RATIO_MIN):
            return False, T
        Tn = np.linalg.inv(L)
        Tn = np.tril(Tn)
        if np.iscomplexobj(Sm):
            # The complex factor checks the pivots and finiteness only
            # (crates/dft-numerics/src/linalg.rs:277-340): no row growth guard.
            return bool(np.all(np.isfinite(Tn))), Tn
        growth = np.abs(Tn[m_from:m_to, :]) @ np.sqrt(np.diag(Sm).real)
        if np.any(~np.isfinite(growth)) or np.any(EPS * growth * growth > math.sqrt(EPS)):
            return False, T
        return True, Tn

    def _eigen(self, m):
        """A = T H_r T^T, EVD, C = T^T Y[:, :n] (workspace.rs:435-499,
        reduced_border/ritz_coefficients lean_davidson.rs:335-444)."""
        T = self.T[:m, :m]
        A = T @ self.Hr[:m, :m] @ _h(T)
        A = 0.5 * (A + _h(A))
        w, Y = np.linalg.eigh(A)
        self.values = w
        self.C = _h(T) @ Y[:, : self.n]

    def _border(self, V, B, m_from, m_to):
        """Border columns V^T B[:, from:to] mirrored (workspace.rs:300-381)."""
        return _h(V[:, :m_to]) @ B[:, m_from:m_to]

    def _refresh(self, m):
        V, HV = self.V[:, :m], self.HV[:, :m]
        self.S[:m, :m] = _h(V) @ V
        self.Hr[:m, :m] = _h(V) @ HV
        self.Hr[:m, :m] = np.triu(self.Hr[:m, :m]) + _h(np.triu(self.Hr[:m, :m], 1))
        self.S[:m, :m] = np.triu(self.S[:m, :m]) + _h(np.triu(self.S[:m, :m], 1))
        if np.iscomplexobj(self.S):
            for M in (self.S, self.Hr):
                M[np.arange(m), np.arange(m)] = M[np.arange(m), np.arange(m)].real
        ok, self.T = self._factor(self.S, 0, m, self.T)
        if not ok:
            raise FloatingPointError("reference contracted basis is dependent")

    def _set_border(self, M, vals, m_from, m_to):
        for jj, j in enumerate(range(m_from, m_to)):
            for i in range(j):
                M[i, j] = vals[i, jj]
                M[j, i] = np.conj(vals[i, jj])
            M[j, j] = vals[j, jj].real if np.iscomplexobj(vals) else vals[j, jj]

    # -- corrections --------------------------------------------------------
    def _corrections(self, X, HX, active):
        E = np.empty((X.shape[0], len(active)), dtype=X.dtype)
        for j, band in enumerate(active):
            theta = self.values[band]
            r = HX[:, j] - theta * X[:, j]
            if self.preconditioner == "tpa":
                w = np.abs(X[:, j]) ** 2
                scale = kinetic_scale(w, self.kinetic)
                E[:, j] = -r * tpa_weight(self.kinetic, scale)
            elif self.preconditioner == "qe":
                E[:, j] = r / sqrt_denominator(self.diag, theta)
            else:  # callable learned preconditioner (residual, x, theta) -> correction
                E[:, j] = self.preconditioner(r, X[:, j], theta, band)
        norms = np.sqrt(np.sum(np.abs(E) ** 2, axis=0))
        if not np.all(np.isfinite(norms)):
            raise FloatingPointError("nonfinite reference correction")
        keep = norms > 32.0 * EPS
        return E[:, keep] / norms[keep]

    def _repair(self, m, E):
        """Rank repair (workspace.rs:586-771): whiten the old basis, refresh,
        two full projections of E, then MGS with the dependence tolerance."""
        self.repairs += 1
        T = self.T[:m, :m]
        P = self.S[:m, :m] @ self.C
        Y = T @ P
        self.V[:, :m] = self.V[:, :m] @ _h(T)
        self.HV[:, :m] = self.HV[:, :m] @ _h(T)
        self.C = Y
        self._refresh(m)
        V = self.V[:, :m]
        for _ in range(2):
            E = E - V @ (_h(V) @ E)
        kept = []
        for j in range(E.shape[1]):
            e = E[:, j].copy()
            for _ in range(2):
                if kept:
                    K = np.stack(kept, 1)
                    e -= K @ (_h(K) @ e)
            nrm = math.sqrt(float(np.sum(np.abs(e) ** 2)))
            if nrm <= DEPENDENCE_TOLERANCE or m + len(kept) == self.cap:
                continue
            kept.append(e / nrm)
        k = len(kept)
        if k:
            E = np.stack(kept, 1)
            self.V[:, m:m + k] = E
            self._set_border(self.S, self._border(self.V, self.V, m, m + k), m, m + k)
            ok, self.T = self._factor(self.S, m, m + k, self.T)
            if not ok:
                raise FloatingPointError("reference overlap remains dependent after repair")
        return k

    def _contract(self, m):
        """Restart to the Ritz block (workspace.rs:772-935)."""
        n = self.n
        self.restarts += 1
        self.V[:, :n] = self.V[:, :m] @ self.C
        self.HV[:, :n] = self.HV[:, :m] @ self.C
        self.contractions += 1
        ideal = False
        if self.reset_disabled:
            pass
        elif self.contractions % REFRESH_INTERVAL == 0:
            self.periodic_refreshes += 1
        elif np.all(np.abs(np.sum(np.abs(self.V[:, :n]) ** 2, axis=0) - 1.0) <= ORTHOGONALITY_LIMIT):
            ideal = True  # reset_norm_guard, workspace.rs:895-918 (NaN fails the test)
            self.ideal_resets += 1
        else:
            self.reset_disabled = True
            self.guard_fallbacks += 1
        self.C = np.eye(n)
        if ideal:
            self.S[:n, :n] = np.eye(n)
            self.Hr[:n, :n] = np.diag(self.values[:n])
            self.T[:n, :n] = np.eye(n)
        else:
            self._refresh(n)
            if not self.reset_disabled:
                defect = np.abs(self.S[:n, :n] - np.eye(n)).max()
                if not defect <= ORTHOGONALITY_LIMIT:
                    self.reset_disabled = True
                    self.guard_fallbacks += 1

    # -- driver -------------------------------------------------------------
    def solve(self, x0, hx0=None):
        rows, n = x0.shape
        cap = self.cap
        if n != self.n or n > rows or self.occupied > n or not np.all(np.isfinite(x0)):
            raise ValueError("reference Davidson dimensions/threshold/start")
        dt = np.complex128 if np.iscomplexobj(x0) else np.float64
        self.V = np.zeros((rows, cap), dtype=dt)
        self.HV = np.zeros((rows, cap), dtype=dt)
        self.S = np.zeros((cap, cap), dtype=dt)
        self.Hr = np.zeros((cap, cap), dtype=dt)
        self.T = np.zeros((cap, cap), dtype=dt)
        self.contractions = 0
        self.reset_disabled = not self.ideal_reset
        self.V[:, :n] = x0
        self.HV[:, :n] = self._h(x0) if hx0 is None else hx0
        self._set_border(self.S, self._border(self.V, self.V, 0, n), 0, n)
        ok, self.T = self._factor(self.S, 0, n, self.T)
        if not ok:
            raise ValueError("reference Davidson start is numerically dependent")
        self._set_border(self.Hr, self._border(self.V, self.HV, 0, n), 0, n)
        self._eigen(n)
        m = n
        active = list(range(n))
        converged = [False] * n
        contracted = False
        self.iterations = 0
        for step in range(self.max_iterations):
            if m + len(active) > cap and m > n:
                self._contract(m)
                m = n
                contracted = True
            if contracted:
                X = self.V[:, active]
                HX = self.HV[:, active]
            else:
                Ca = self.C[:, active]
                X = self.V[:, :m] @ Ca
                HX = self.HV[:, :m] @ Ca
            contracted = False
            E = self._corrections(X, HX, active)
            k = E.shape[1]
            if m + k > cap or k == 0:
                k = self._repair(m, E)
            else:
                self.V[:, m:m + k] = E
                self._set_border(self.S, self._border(self.V, self.V, m, m + k), m, m + k)
                ok, T = self._factor(self.S, m, m + k, self.T)
                if ok:
                    self.T = T
                else:
                    k = self._repair(m, E)
            old = self.values[:n].copy()
            if k > 0:
                self.HV[:, m:m + k] = self._h(self.V[:, m:m + k])
                self._set_border(self.Hr, self._border(self.V, self.HV, m, m + k), m, m + k)
                self._eigen(m + k)
                m += k
            active = []
            for i in range(n):
                converged[i] = abs(self.values[i] - old[i]) < self.band_threshold(i)
                if not converged[i]:
                    active.append(i)
            self.iterations = step + 1
            if not active or self.iterations == self.max_iterations:
                break
            if m + len(active) > cap:
                self._contract(m)
                m = n
                contracted = True
        X = self.V[:, :m] @ self.C
        HX = self.HV[:, :m] @ self.C
        return dict(eigenvalues=self.values[:n].copy(), x=X, hx=HX, converged=list(converged),
                    iterations=self.iterations, h_calls=self.h_calls, h_columns=self.h_columns,
                    repairs=self.repairs, restarts=self.restarts, ideal_resets=self.ideal_resets,
                    periodic_refreshes=self.periodic_refreshes, guard_fallbacks=self.guard_fallbacks)


def random_orc_start(rows, bands, seed):
    """Complex N(0,1) coefficients on the half sphere, Hermitian-completed:
    in ORC these are real Gaussian rows (G=0 real), each column normalised."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((rows, bands))
    return x / np.sqrt(np.sum(x * x, axis=0))
