"""Gamma SCF driver around the GPU H·X operator and Davidson (the caller of
the hot path: crates/dft-engine/src/refscf/driver.rs:479-757 and map.rs:200-384).

Per map: Hartree + LDA-PZ fields of the input density, V = V_ion + V_H + V_xc
installed on the operator, the reference Davidson on the GPU with the ethr
schedule (refscf/profile.rs:49-51) and the iteration-1 redo, the output
density from the GPU band-density transform, the half-sphere Hartree metric
dr2 (refscf/metric.rs) and difference-history Anderson mixing with beta 0.7,
history 8 (refscf/broyden.rs:317-445). Stops when dr2 < threshold_ry.

The density-sphere field glue (grid transforms of the density and the Hartree
potential, pointwise XC, the mixer's metric products) is the caller's side of
the boundary; it runs on the GPU with torch tensors (torch.fft for the 4 Ecut
density grid), so that nothing but scalars and the small mixing matrix
crosses to the host per map.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from . import lib as L
from .hamiltonian import DistributedRealHamiltonian, GammaSystem
from . import ref_davidson


def lda_pz(n):
    """Slater exchange + PZ81 correlation per point (crates/dft-physics/src/xc.rs:48-73):
    returns (eps_xc, v_xc); zero at or below 1e-10."""
    eps = np.zeros_like(n)
    vxc = np.zeros_like(n)
    live = n > 1e-10
    d = n[live]
    cd = np.cbrt(d)
    rs = np.cbrt(3.0 / (4.0 * math.pi)) / cd
    ex = -0.75 * np.cbrt(3.0 / math.pi) * cd
    ec = np.empty_like(d)
    dec = np.empty_like(d)
    low = rs >= 1.0
    sr = np.sqrt(rs[low])
    den = 1.0 + 1.0529 * sr + 0.3334 * rs[low]
    ec[low] = -0.1423 / den
    dec[low] = 0.1423 * (1.0529 / (2.0 * sr) + 0.3334) / (den * den)
    rh = rs[~low]
    lg = np.log(rh)
    ec[~low] = 0.0311 * lg - 0.048 + 0.0020 * rh * lg - 0.0116 * rh
    dec[~low] = 0.0311 / rh + 0.0020 * (lg + 1.0) - 0.0116
    eps[live] = ex + ec
    vxc[live] = 4.0 * ex / 3.0 + ec - rs * dec / 3.0
    return eps, vxc


def lda_pz_torch(n):
    """lda_pz on a CUDA tensor (same formulas, same 1e-10 cut)."""
    import torch
    live = n > 1e-10
    d = torch.where(live, n, torch.ones_like(n))
    cd = torch.pow(d, 1.0 / 3.0)
    rs = (3.0 / (4.0 * math.pi)) ** (1.0 / 3.0) / cd
    ex = -0.75 * (3.0 / math.pi) ** (1.0 / 3.0) * cd
    low = rs >= 1.0
    sr = torch.sqrt(rs)
    den = 1.0 + 1.0529 * sr + 0.3334 * rs
    lg = torch.log(rs)
    ec = torch.where(low, -0.1423 / den, 0.0311 * lg - 0.048 + 0.0020 * rs * lg - 0.0116 * rs)
    dec = torch.where(low, 0.1423 * (1.0529 / (2.0 * sr) + 0.3334) / (den * den),
                      0.0311 / rs + 0.0020 * (lg + 1.0) - 0.0116)
    zero = torch.zeros_like(n)
    eps = torch.where(live, ex + ec, zero)
    vxc = torch.where(live, 4.0 * ex / 3.0 + ec - rs * dec / 3.0, zero)
    return eps, vxc


@dataclass
class ScfOptions:
    bands: int
    threshold_ry: float = 1e-8       # refscf/profile.rs:17-26
    beta: float = 0.7
    history: int = 8
    max_iterations: int = 100
    davidson_max_iterations: int = 20
    preconditioner: int = L.BD_PRECOND_QE
    seed: int = 1
    # BD_FP64: every map in FP64. BD_MIXED: maps whose Davidson threshold is at
    # least mixed_floor_ha run on the mixed-tier operator (FP32 FFT H·X, BF16x3
    # GEMMs, FP64 factor/eigensolve); the later maps, and the final answer, run
    # in FP64, so the fixed point is the FP64 one.
    precision: int = L.BD_FP64
    mixed_floor_ha: float = 5e-5
    mlp: object = None  # a baandi_amd.learned.Network when preconditioner == BD_PRECOND_MLP
    allow_unconverged: bool = False  # return the last map instead of raising at max_iterations
    preconditioning: str = "plain"   # density mixing: "plain" or "thomas_fermi" (refscf/profile.rs:4-7)
    capacity: int = 0                # Davidson subspace capacity (0: 2 bands, ref_davidson.rs:38-78)
    autotune: bool = False           # tune the H·X launch shapes at set-up (bit-neutral)
    # Optional fixed-point verification map (driver.rs:727-757): one more
    # Davidson at the converged potential with every band at this threshold (Ha).
    final_threshold: float = 0.0
    # Fermi-Dirac smearing width (Ha) for cells whose gap is too small for
    # fixed occupations (occupations.rs:71-232); 0 keeps the lowest
    # electrons/2 bands doubly occupied. With smearing every band is held to
    # the Davidson threshold and the energy is the free energy E - TS.
    theta: float = 0.0


class DensityField:
    """Density-sphere coefficients (4 Ecut) and their grid transforms, on the GPU."""

    def __init__(self, system: GammaSystem):
        import torch
        dev = torch.device("cuda")
        self.dims = tuple(system.dims)
        self.size = int(np.prod(self.dims))
        self.flat = torch.from_numpy(system.get("density_index").astype(np.int64)).to(dev)
        g2 = system.get("density_g2")
        miller = system.get("density_miller")
        zero = np.all(miller == 0, axis=1)
        # Positive representatives carry the half-sphere Hartree weight
        # 8 pi Omega / G^2 in Ry (refscf/metric.rs:1-5).
        first = np.where(miller[:, 0] != 0, miller[:, 0], np.where(miller[:, 1] != 0, miller[:, 1], miller[:, 2]))
        positive = first > 0
        metric = np.zeros(len(g2))
        metric[positive] = 8.0 * math.pi * system.volume / g2[positive]
        coulomb = np.zeros(len(g2))
        coulomb[~zero] = 4.0 * math.pi / g2[~zero]
        self.g2 = g2
        self.zero_np = zero
        self.zero = torch.from_numpy(zero).to(dev)
        self.metric = torch.from_numpy(metric).to(dev)
        self.coulomb = torch.from_numpy(coulomb).to(dev)

    def to_grid(self, coeffs):
        import torch
        g = torch.zeros(self.size, dtype=torch.complex128, device=coeffs.device)
        g[self.flat] = coeffs
        return torch.fft.ifftn(g.view(self.dims)) * self.size

    def from_grid(self, field):
        import torch
        return (torch.fft.fftn(field.view(self.dims).to(torch.complex128)) / self.size).reshape(-1)[self.flat]

    def dot(self, a, b):
        return float((self.metric * (a.real * b.real + a.imag * b.imag)).sum())

    def gram(self, rows, f):
        """Metric products <r_i, r_j> and <r_i, f> of stacked coefficients
        (host arrays). Elementwise products and row sums: a library GEMM with
        an 8 x 8 result and a 50k inner dimension is far slower here."""
        w = self.metric
        re, im = rows.real, rows.imag
        wre, wim = re * w, im * w
        a = (wre[:, None, :] * re[None, :, :]).sum(dim=-1) + (wim[:, None, :] * im[None, :, :]).sum(dim=-1)
        c = (wre * f.real).sum(dim=-1) + (wim * f.imag).sum(dim=-1)
        return a.cpu().numpy(), c.cpu().numpy()


class Mixer:
    """Difference-history Anderson update with the Hartree metric (broyden.rs).

    ``preconditioning`` "thomas_fermi" scales the history-reduced residual by
    |G|^2 / (|G|^2 + a), a = (12/pi)^(2/3) / r_s (broyden.rs:106-113, applied
    to F~ only, :439); "plain" (the reference default, profile.rs:17-26)
    leaves it unscaled."""

    def __init__(self, field: DensityField, beta, history, preconditioning="plain", volume=None, electrons=None):
        import torch
        self.field, self.beta, self.history = field, beta, history
        self.pairs = []  # (dx, dF), newest last
        self.prev = None
        self.screening = None
        if preconditioning == "thomas_fermi":
            rs = (3.0 * volume / (4.0 * math.pi * electrons)) ** (1.0 / 3.0)
            a = (12.0 / math.pi) ** (2.0 / 3.0) / rs
            g2 = torch.from_numpy(field.g2).to(field.metric.device)
            self.screening = g2 / (g2 + a)
        elif preconditioning != "plain":
            raise ValueError(f"unknown preconditioning {preconditioning!r}")

    def update(self, x, f):
        import torch
        if self.prev is not None:
            px, pf = self.prev
            self.pairs.append((px - x, pf - f))
            if len(self.pairs) > self.history:
                self.pairs.pop(0)
        self.prev = (x.clone(), f.clone())
        xx, ff = x.clone(), f.clone()
        while self.pairs:
            p = len(self.pairs)
            a, c = self.field.gram(torch.stack([df for _, df in self.pairs]), f)
            a = 0.5 * (a + a.T)
            lam, q = np.linalg.eigh(a)
            if lam[-1] <= 0.0 or lam[0] <= math.sqrt(p * np.finfo(float).eps) * lam[-1]:
                self.pairs.pop(0)  # rank-deficient history: drop the oldest pair
                continue
            gamma = torch.from_numpy(q @ ((q.T @ c) / lam)).to(x.device)
            dX = torch.stack([dx for dx, _ in self.pairs])
            dF = torch.stack([df for _, df in self.pairs])
            xx = xx - (gamma[:, None] * dX).sum(dim=0)
            ff = ff - (gamma[:, None] * dF).sum(dim=0)
            break
        if self.screening is not None:
            ff = ff * self.screening
        return xx + self.beta * ff


def next_ethr_ry(previous, dr2_ry, electrons):
    """refscf/profile.rs:49-51."""
    return max(min(previous, 0.1 * dr2_ry / max(electrons, 1.0)), 1e-13)


def run_scf(system: GammaSystem, options: ScfOptions, log=None, start=None, profile_kernels=False):
    """SCF to dr2 < threshold_ry; returns a dict with energies (Ha), eigenvalues,
    per-iteration history and timings. ``start``: optional (bands, n) CUDA tensor."""
    import torch
    dev = torch.device("cuda")
    t_setup = time.perf_counter()
    bands = options.bands
    electrons = system.electrons
    occupied = int(round(electrons)) // 2
    occ = np.zeros(bands)
    occ[:occupied] = 2.0
    smeared = options.theta > 0.0
    converge = bands if smeared else occupied
    mu, ts = None, 0.0
    if smeared:
        from .kpoint import fermi_dirac
        if 2 * bands <= electrons:
            raise L.BaandiError(L.BD_ERR_INVALID_PARAMETER, "smearing needs more bands than electrons/2")
    field = DensityField(system)
    vion_g = torch.from_numpy(system.get("ionic_density_re") + 1j * system.get("ionic_density_im")).to(dev)
    vion = field.to_grid(vion_g).real.contiguous()
    op = DistributedRealHamiltonian(system, capacity=bands, precision=L.BD_FP64)
    op_mixed = (DistributedRealHamiltonian(system, capacity=bands, precision=L.BD_MIXED)
                if options.precision == L.BD_MIXED else None)
    if options.autotune:  # on the ionic potential (the shapes do not depend on V)
        for o in (op, op_mixed):
            if o is not None:
                o.set_owned_potential(vion.reshape(-1).contiguous())
                o.autotune(bands)
    diag_static = torch.from_numpy(system.get("diagonal_static")).to(dev)
    kinetic = system.get("kinetic_orc")
    ewald = float(system.get("ewald")[0])
    # Start density: superposed Gaussian atomic charges Z (alpha/pi)^1.5 e^{-alpha r^2}, alpha = 1/rc^2.
    atoms = system.get("atoms_cart")
    rows = np.array(system.lattice_rows)
    b = 2.0 * math.pi * np.linalg.inv(rows).T  # reciprocal vectors as rows
    dm = system.get("density_miller").astype(np.float64)
    dcart = dm @ b
    sfac = np.exp(-1j * (dcart @ atoms.T)).sum(axis=1)
    x = system.z_valence / system.volume * np.exp(-field.g2 * system.rc ** 2 / 4.0) * sfac
    x[field.zero_np] = electrons / system.volume
    x = torch.from_numpy(x).to(dev)
    mixer = Mixer(field, options.beta, options.history, options.preconditioning, system.volume, electrons)
    if start is None:
        g = torch.Generator(device="cpu").manual_seed(options.seed)
        start = torch.randn(bands, system.n, generator=g, dtype=torch.float64)
        start = (start / start.norm(dim=1, keepdim=True)).to(dev)
    orbitals = start.contiguous()
    xout = torch.empty_like(orbitals)
    hxout = torch.empty_like(orbitals)
    rho_dev = torch.empty(field.size, dtype=torch.float64, device=dev)
    if profile_kernels:
        op.set_profiling(True)
        if op_mixed is not None:
            op_mixed.set_profiling(True)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup

    t0 = time.perf_counter()
    ethr = 1e-2
    last_dr2 = math.inf
    history = []
    totals = dict(h_columns=0, davidson_iterations=0, dense_s=0.0, apply_s=0.0, davidson_s=0.0, fields_s=0.0,
                  mixed_maps=0, mlp_fallbacks=0, eigen_retries=0)
    result = None
    for iteration in range(1, options.max_iterations + 1):
        if iteration > 1:
            ethr = next_ethr_ry(1e-2 if iteration == 2 else ethr, last_dr2, electrons)
        redo = False
        while True:
            tf = time.perf_counter()
            vh_g = field.coulomb * x
            both = field.to_grid(x + 1j * vh_g)  # rho_in + i V_H (both real fields)
            rho_in, vh = both.real, both.imag
            eps, vxc = lda_pz_torch(rho_in.abs())
            potential = vion + vh + vxc
            pot_dev = potential.reshape(-1).contiguous()
            vmean = float(pot_dev.mean())
            use_mixed = op_mixed is not None and ethr / 2.0 >= options.mixed_floor_ha
            cur = op_mixed if use_mixed else op
            cur.set_owned_potential(pot_dev)
            diag = diag_static + vmean
            totals["fields_s"] += time.perf_counter() - tf
            td = time.perf_counter()
            opts = ref_davidson.Options(bands=bands, threshold=ethr / 2.0, occupied=converge,
                                        first_step=(iteration == 1 and not redo),
                                        max_iterations=options.davidson_max_iterations, return_hx=True,
                                        preconditioner=options.preconditioner, mlp=options.mlp,
                                        mlp_vbar=vmean, capacity=options.capacity)
            sol = ref_davidson.solve_with_initial_hx(cur, orbitals, diag, opts, out=(xout, hxout))
            totals["davidson_s"] += time.perf_counter() - td
            totals["mixed_maps"] += int(use_mixed)
            totals["mlp_fallbacks"] += int(sol.record.get("mlp_fallback", -1) >= 0)
            totals["h_columns"] += sol.record["h_columns"]
            totals["eigen_retries"] += sol.record["eigen_retries"]
            totals["davidson_iterations"] += sol.record["iterations"]
            totals["dense_s"] += sol.record["seconds_dense"]
            totals["apply_s"] += sol.record["seconds_apply"]
            tf = time.perf_counter()
            if smeared:
                occ, mu, ts = fermi_dirac(sol.eigenvalues[None, :], np.ones(1), electrons, options.theta)
                occ = occ[0]
            cur.band_density_sum(xout, occ, rho_dev)
            rho_out = rho_dev.view(field.dims)
            y = field.from_grid(rho_out)
            y[field.zero] = electrons / system.volume
            f = y - x
            dr2 = field.dot(f, f)
            totals["fields_s"] += time.perf_counter() - tf
            if iteration == 1 and not redo and dr2 < ethr * max(electrons, 1.0):
                ethr = next_ethr_ry(ethr, dr2, electrons)
                redo = True
                orbitals, xout = xout, orbitals
                continue
            break
        orbitals, xout = xout, orbitals
        last_dr2 = dr2
        history.append(dict(iteration=iteration, dr2_ry=dr2, ethr_ry=ethr, mixed=use_mixed,
                            davidson_iterations=sol.record["iterations"], h_columns=sol.record["h_columns"],
                            converged=bool(sol.converged.all())))
        if log:
            log(f"scf {iteration:3d} dr2={dr2:.3e} ethr={ethr:.2e} dav_it={sol.record['iterations']} "
                f"hcols={sol.record['h_columns']}")
        if dr2 < options.threshold_ry or (options.allow_unconverged and iteration == options.max_iterations):
            result = (sol, rho_in, potential, rho_out)
            break
        x = mixer.update(x, f)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if result is None:
        raise RuntimeError(f"SCF did not reach dr2 < {options.threshold_ry} Ry in {options.max_iterations} maps")
    sol, rho_in, potential, rho_out = result
    if options.final_threshold > 0.0:
        vmean = float(potential.mean())
        diag = diag_static + vmean
        opts = ref_davidson.Options(bands=bands, threshold=options.final_threshold, occupied=converge,
                                    first_step=True, max_iterations=200, return_hx=True,
                                    preconditioner=options.preconditioner, mlp=options.mlp,
                                    mlp_vbar=vmean)
        sol = ref_davidson.solve_with_initial_hx(op, orbitals, diag, opts, out=(xout, hxout))
        orbitals, xout = xout, orbitals
        if smeared:
            occ, mu, ts = fermi_dirac(sol.eigenvalues[None, :], np.ones(1), electrons, options.theta)
            occ = occ[0]
        op.band_density_sum(orbitals, occ, rho_dev)
        rho_out = rho_dev.view(field.dims)
    # Energy audit of the converged map (refscf/map.rs:466-540).
    xo = orbitals.cpu().numpy()
    kin = float(np.sum(occ[:, None] * xo ** 2 * kinetic[None, :]))
    nonlocal_ = float(np.dot(op.nonlocal_band_energies(orbitals), occ))
    dv = system.volume / field.size
    yo = field.from_grid(rho_out)
    vh_out = field.to_grid(field.coulomb * yo).real
    hartree = 0.5 * dv * float((rho_out * vh_out).sum())
    eps_o, _ = lda_pz_torch(rho_out.abs())
    xc = dv * float((rho_out.abs() * eps_o).sum())
    local = dv * float((rho_out * vion).sum())
    total = kin + local + nonlocal_ + hartree + xc + ewald - ts
    energy = {"kinetic": kin, "local": local, "nonlocal": nonlocal_, "hartree": hartree, "xc": xc,
              "ewald": ewald, "entropy_term": ts, "total": total, "band_sum": float(np.dot(sol.eigenvalues, occ))}
    out = dict(energy=energy, eigenvalues=sol.eigenvalues.copy(), occupations=occ.copy(), mu=mu,
               history=history, wall_s=wall,
               setup_s=t_setup, iterations=len(history), totals=totals, orbitals=orbitals, op=op,
               plan=op.plan_info(), potential=potential.cpu().numpy())
    if profile_kernels:
        out["kernel_totals"] = op.stage_totals()
        if op_mixed is not None:
            out["kernel_totals_mixed"] = op_mixed.stage_totals()
    return out
