"""Oracle: Gamma SCF driver around the reference Davidson.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows crates/dft-engine/src/refscf/driver.rs:479-757 (ethr schedule,
iteration-1 redo, dr2 stop), map.rs:200-384 (fields, Davidson call, output
density, energy audit), broyden.rs:317-445 (difference-history Anderson update
with beta=0.7, history 8) and metric.rs (half-sphere Hartree metric in Ry).
The start density is a superposition of the synthetic Gaussian atomic charges
(the PP_RHOATOM shape of crates/dft-engine/tests/common/mod.rs:190-196).

Parity: KAT-1 total energy -4.684561184277879 Ha and eigenvalues; KAT-2
-0.8646074276373095 Ha (tests/test_oracle_scf.py).
"""
from __future__ import annotations

import math

import numpy as np

from . import fields as F
from .davidson import Davidson, random_orc_start
from .hamiltonian import apply_orc
from .physics import ewald, map_miller, flat_index, structure_factor


def next_ethr_ry(previous, dr2_ry, electrons):
    """refscf/profile.rs:49-51."""
    return max(min(previous, 0.1 * dr2_ry / max(electrons, 1.0)), 1e-13)


def davidson_diagonal(system):
    """T_G + sum_ij D_ij Re(beta_i beta_j^*) per ORC row, shared by the two
    rows of a representative (refstart.rs:605-650)."""
    n = len(system.basis.g2)
    full = 0.5 * system.basis.g2.copy()
    p = system.projectors
    for start, nc, d in system.coupling:
        b = p[:, start:start + nc]
        full += np.real(np.einsum("gi,ij,gj->g", b, d.T, np.conj(b)))
    out = np.empty(n)
    out[0] = full[0]
    out[1::2] = full[system.half.plus[1:]]
    out[2::2] = full[system.half.plus[1:]]
    return out


class DensityPlan:
    """Density-sphere coefficients <-> grid; the Hartree metric."""

    def __init__(self, system):
        self.s = system
        dens = system.density
        idx = map_miller(dens.miller, system.dims)
        self.flat = flat_index(idx, system.dims)
        self.g2 = dens.g2
        self.zero = np.all(dens.miller == 0, axis=1)
        # half-sphere weights 8 pi Omega/G^2 (Ry) on positive representatives
        pos = np.array([tuple(m) > (0, 0, 0) for m in dens.miller.tolist()])
        w = np.zeros(len(self.g2))
        w[pos] = 8.0 * math.pi * system.cell.volume / self.g2[pos]
        self.weights = w

    def forward(self, rho_r):
        return (np.fft.fftn(rho_r) / rho_r.size).reshape(-1)[self.flat]

    def inverse(self, coeffs):
        dims = self.s.dims
        g = np.zeros(int(np.prod(dims)), dtype=np.complex128)
        g[self.flat] = coeffs
        return (np.fft.ifftn(g.reshape(dims)) * g.size).real

    def dot(self, a, b):
        return float(np.sum(self.weights * np.real(np.conj(a) * b)))


class Broyden:
    """Unnormalised difference-history Anderson update (broyden.rs:317-445)."""

    def __init__(self, plan, beta=0.7, history=8):
        self.plan, self.beta, self.history = plan, beta, history
        self.it = 0
        self.df, self.dx = [], []
        self.prev_x = self.prev_f = None

    def update(self, x, f):
        self.it += 1
        if self.it > 1:
            slot = (self.it - 2) % self.history
            dfv, dxv = self.prev_f - f, self.prev_x - x
            if slot < len(self.df):
                self.df[slot], self.dx[slot] = dfv, dxv
            else:
                self.df.append(dfv)
                self.dx.append(dxv)
        self.prev_x, self.prev_f = x.copy(), f.copy()
        xx, ff = x.copy(), f.copy()
        while self.df:
            p = len(self.df)
            a = np.array([[self.plan.dot(self.df[i], self.df[j]) for j in range(p)] for i in range(p)])
            c = np.array([self.plan.dot(self.df[i], f) for i in range(p)])
            lam, q = np.linalg.eigh(a)
            if lam[-1] == 0.0 or lam[0] <= math.sqrt(p * np.finfo(float).eps) * lam[-1]:
                # Conditioning opt-in of broyden.rs:384-404: drop the oldest
                # pair instead of failing on a rank-deficient history.
                oldest = (self.it - 1) % self.history if len(self.df) == self.history else 0
                del self.df[oldest], self.dx[oldest]
                self.it = len(self.df) + 1
                continue
            gamma = q @ ((q.T @ c) / lam)
            for i in range(p):
                xx -= gamma[i] * self.dx[i]
                ff -= gamma[i] * self.df[i]
            break
        return xx + self.beta * ff


def initial_density_coefficients(system):
    """Superposed Gaussian atomic charges Z (alpha/pi)^{3/2} e^{-alpha r^2},
    alpha = 1/rc^2: rho(G) = (Z/Omega) e^{-G^2 rc^2/4} S(G)."""
    dens = system.density
    out = np.zeros(len(dens.g2), dtype=np.complex128)
    for sp in system.species:
        pp = sp.pseudo
        out += pp.z / system.cell.volume * np.exp(-dens.g2 * pp.rc ** 2 / 4.0) * structure_factor(dens.cart, sp.cart)
    return out


def scf(system, bands, seed=1, threshold_ry=1e-8, max_iterations=60, beta=0.7, history=8,
        preconditioner="qe", davidson_max_iterations=20, log=None, final_tight=None):
    """Reference-loop SCF at Gamma; returns dict with energy audit, eigenvalues,
    X and the history. ``final_tight``: tighter Davidson threshold (Ha) for
    one verification map at the end (driver.rs:727-757 fixed-point check)."""
    plan = DensityPlan(system)
    electrons = system.electrons
    occ = np.zeros(bands)
    occ[: int(round(electrons)) // 2] = 2.0
    occupied = int(round(electrons)) // 2
    diag_static = davidson_diagonal(system)
    ew = ewald(system.cell, [sp.pseudo.z for sp in system.species for _ in sp.frac],
               np.concatenate([sp.cart for sp in system.species]))["energy"]
    x = initial_density_coefficients(system)
    x[plan.zero] = electrons / system.cell.volume
    mixer = Broyden(plan, beta, history)
    orbitals = random_orc_start(len(system.basis.g2), bands, seed)
    ethr = 1e-2
    last_dr2 = math.inf
    history_ = []
    total_h_columns = 0
    result = None
    for iteration in range(1, max_iterations + 1):
        if iteration > 1:
            ethr = next_ethr_ry(1e-2 if iteration == 2 else ethr, last_dr2, electrons)
        redo = False
        while True:
            rho_in = plan.inverse(x)
            fin = F.fields(system.cell, rho_in)
            potential = system.ionic_grid + fin["hxc"]
            v0 = float(np.mean(potential))
            diag = diag_static + v0
            dav = Davidson(lambda v: apply_orc(system, v, potential), diag, bands, ethr / 2.0, occupied,
                           first_step=(iteration == 1 and not redo), max_iterations=davidson_max_iterations,
                           preconditioner=preconditioner, kinetic=system.kinetic_orc)
            sol = dav.solve(orbitals)
            total_h_columns += dav.h_columns
            rho_out = F.orc_density(system, sol["x"], occ)
            y = plan.forward(rho_out)
            y[plan.zero] = electrons / system.cell.volume
            f = y - x
            dr2 = plan.dot(f, f)
            if iteration == 1 and not redo and dr2 < ethr * max(electrons, 1.0):
                ethr = next_ethr_ry(ethr, dr2, electrons)
                redo = True
                orbitals = sol["x"]
                continue
            break
        orbitals = sol["x"]
        last_dr2 = dr2
        history_.append(dict(iteration=iteration, dr2_ry=dr2, ethr_ry=ethr,
                             davidson_iterations=sol["iterations"], h_columns=dav.h_columns))
        if log:
            log(f"scf {iteration}: dr2={dr2:.3e} ethr={ethr:.2e} dav_it={sol['iterations']}")
        if dr2 < threshold_ry:
            result = (sol, rho_in, fin, potential)
            break
        x = mixer.update(x, f)
    if result is None:
        raise RuntimeError("oracle SCF did not converge")
    sol, rho_in, fin, potential = result
    if final_tight is not None:
        diag = diag_static + float(np.mean(potential))
        dav = Davidson(lambda v: apply_orc(system, v, potential), diag, bands, final_tight, occupied,
                       first_step=True, max_iterations=200, preconditioner=preconditioner,
                       kinetic=system.kinetic_orc)
        sol = dav.solve(sol["x"])
    audit = F.energy_audit(system, sol["x"], occ, sol["eigenvalues"], rho_in, fin, ew)
    return dict(energy=audit, eigenvalues=sol["eigenvalues"], x=sol["x"], history=history_,
                occupations=occ, potential=potential, h_columns=total_h_columns)
