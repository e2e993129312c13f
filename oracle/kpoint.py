"""Oracle: the complex (k != 0) Hamiltonian, Monkhorst-Pack meshes, Fermi-Dirac
occupations and a k-point SCF.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows:
* crates/dft-basis/src/lib.rs:234-269 (PlaneWaveBasis::at_k: |k+G|^2/2 <= Ecut,
  order by |G|^2 then Miller; momenta k+G) -- ``physics.plane_wave_basis``;
* crates/dft-physics/src/distributed_complex.rs:29-54 (static diagonal),
  284-386 (apply_block: FFT of the whole vector, + T_{k+G} x, + P D P^H x) and
  630-720 (the explicit dense H_k);
* crates/dft-physics/src/hamiltonian.rs:885-1035 (projector columns at q = k+G)
  -- ``physics.nonlocal_table``;
* crates/dft-physics/src/occupations.rs:57-232 (Fermi-Dirac: bracket, bisect to
  a 1e-13 relative electron count, entropy from x = (e - mu)/theta);
* crates/dft-geometry/src/kpoints.rs:21-32 (k wrapped into [-1/2, 1/2)).

Pinned by: the k = 0 spectrum equals the Gamma (ORC) spectrum of the same cell;
with V = 0 and no projectors the spectrum is |k+G|^2/2; H_k is Hermitian and
equals its FFT apply; E(k) = E(-k) (time reversal); Fermi-Dirac against the
logistic/entropy closed forms; and the Gamma-only k SCF reproduces the KAT-1
certified energy (tests/test_oracle_kpoint.py).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import fields as F
from .physics import (Cell, Species, plane_wave_basis, grid_dims, map_miller, flat_index, nonlocal_table,
                      positions_cart, ionic_coefficients, to_grid, ewald)


def wrap_fractional(k):
    """kpoints.rs:21-32: each component into [-1/2, 1/2)."""
    return tuple(x if -0.5 <= x < 0.5 else x - math.floor(x + 0.5) for x in k)


@dataclass
class KSystem:
    cell: Cell
    ecut: float
    species: list
    k_frac: tuple
    basis: object
    dims: tuple
    density: object
    fft_index: np.ndarray
    projectors: np.ndarray  # (n, p) complex, columns at q = k+G
    coupling: list
    kinetic: np.ndarray     # |k+G|^2 / 2
    ionic_grid: np.ndarray
    electrons: float


def build_k_system(cell_rows, atoms_frac, pseudo, ecut, k_frac, grid_override=None):
    cell = Cell.from_rows(cell_rows)
    sp = Species(pseudo, np.asarray(atoms_frac, dtype=np.float64))
    sp.cart = positions_cart(cell, sp.frac)
    k = wrap_fractional(k_frac)
    basis = plane_wave_basis(cell, ecut, k)
    dims, dens = grid_dims(cell, ecut)
    if grid_override is not None:
        dims = tuple(grid_override)
    fft_index = flat_index(map_miller(basis.miller, dims), dims)
    proj, coupling = nonlocal_table(basis, cell.volume, [sp])
    vion = to_grid(ionic_coefficients(dens, cell.volume, [sp]), dens.miller, dims).real
    return KSystem(cell=cell, ecut=ecut, species=[sp], k_frac=k, basis=basis, dims=dims, density=dens,
                   fft_index=fft_index, projectors=proj, coupling=coupling, kinetic=0.5 * basis.q2,
                   ionic_grid=vion, electrons=pseudo.z * len(sp.frac))


def nonlocal_apply(ks, x):
    """P D P^H x (distributed_complex.rs:284-338, 368-384)."""
    if ks.projectors.shape[1] == 0:
        return np.zeros_like(x)
    q = ks.projectors.conj().T @ x
    w = np.zeros_like(q)
    for start, nc, d in ks.coupling:
        w[start:start + nc] = d @ q[start:start + nc]
    return ks.projectors @ w


def apply_k(ks, x, potential):
    """H_k x for complex columns x (n, m): one FFT per column
    (distributed_complex.rs:340-386)."""
    x = np.asarray(x, dtype=np.complex128)
    if x.ndim == 1:
        x = x[:, None]
    n, m = x.shape
    N = int(np.prod(ks.dims))
    out = np.empty_like(x)
    for j in range(m):
        g = np.zeros(N, dtype=np.complex128)
        g[ks.fft_index] = x[:, j]
        r = np.fft.ifftn(g.reshape(ks.dims)) * N
        r *= potential
        out[:, j] = (np.fft.fftn(r) / N).reshape(-1)[ks.fft_index]
    out += ks.kinetic[:, None] * x
    out += nonlocal_apply(ks, x)
    return out


def dense_k(ks, potential):
    """Explicit H_k (distributed_complex.rs:630-720): T delta + V(G-G') + P D P^H,
    V(h) = N^-1 sum_r V(r) exp(-i h.r), Hermitised."""
    n = len(ks.kinetic)
    dims = ks.dims
    N = int(np.prod(dims))
    vg = np.fft.fftn(potential) / N
    m = ks.basis.miller
    diff = m[:, None, :] - m[None, :, :]
    idx = [np.mod(diff[..., a], dims[a]) for a in range(3)]
    h = vg[idx[0], idx[1], idx[2]].astype(np.complex128)
    h[np.diag_indices(n)] += ks.kinetic
    if ks.projectors.shape[1]:
        pd = np.zeros_like(ks.projectors)
        for start, nc, d in ks.coupling:
            pd[:, start:start + nc] = ks.projectors[:, start:start + nc] @ d.T
        h += pd @ ks.projectors.conj().T
    return 0.5 * (h + h.conj().T)


def static_diagonal(ks):
    """T + sum_ij D_ij Re(P_gi conj(P_gj)) per row (distributed_complex.rs:21-54)."""
    diag = ks.kinetic.copy()
    p = ks.projectors
    for start, nc, d in ks.coupling:
        b = p[:, start:start + nc]
        diag += np.real(np.einsum("gi,ij,gj->g", b, d, np.conj(b)))
    return diag


def monkhorst_pack(mesh, shift=(0.0, 0.0, 0.0), time_reversal=True):
    """Fractional points ((i + shift)/n - 1/2 + 1/(2n) per axis, wrapped) with
    weights summing to 1; k and -k merged when time_reversal."""
    pts = []
    for i in range(mesh[0]):
        for j in range(mesh[1]):
            for l in range(mesh[2]):
                f = [(c + s + 0.5) / n - 0.5 for c, s, n in zip((i, j, l), shift, mesh)]
                pts.append(wrap_fractional(f))
    weight = 1.0 / len(pts)
    out = {}
    for p in pts:
        key = tuple(round(x * 1e12) for x in p)
        if time_reversal:
            neg = tuple(round(x * 1e12) for x in wrap_fractional([-x for x in p]))
            if neg in out:
                out[neg][1] += weight
                continue
        if key in out:
            out[key][1] += weight
        else:
            out[key] = [p, weight]
    return [(tuple(v[0]), v[1]) for v in out.values()]


def _fraction_entropy(x):
    """occupations.rs:57-66: occupation fraction and entropy from x, both tails
    through exp(-|x|)."""
    a = abs(x)
    t = math.exp(-a)
    q = t / (1.0 + t)
    p = q if x >= 0.0 else 1.0 / (1.0 + t)
    s = math.log1p(t) + (0.0 if math.isinf(a) else a * q)
    return p, s


def fermi_dirac(energies, weights, electrons, theta, tolerance=1e-13, max_iterations=256):
    """Fermi-Dirac occupations (capacity 2) of energies (nk, nb) with k weights:
    returns occupations, mu and the entropy term TS (Ha) (occupations.rs:71-232)."""
    e = np.asarray(energies, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64)

    def number(mu):
        return sum(2.0 * w[k] * _fraction_entropy((e[k, i] - mu) / theta)[0]
                   for k in range(e.shape[0]) for i in range(e.shape[1]))

    width = theta
    lo = hi = None
    for _ in range(max_iterations):
        lo, hi = e.min() - width, e.max() + width
        if number(lo) <= electrons <= number(hi):
            break
        width *= 2.0
    else:
        raise RuntimeError("unable to bracket the chemical potential")
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
        raise RuntimeError("electron count did not converge")
    occ = np.empty_like(e)
    entropy = 0.0
    for k in range(e.shape[0]):
        for i in range(e.shape[1]):
            p, s = _fraction_entropy((e[k, i] - mu) / theta)
            occ[k, i] = 2.0 * p
            entropy += w[k] * 2.0 * s
    return occ, mu, theta * entropy


def band_density(ks, x, occ, weight):
    """weight * sum_n f_n |psi_n(r)|^2 / Omega on the grid."""
    N = int(np.prod(ks.dims))
    rho = np.zeros(ks.dims)
    for j in range(x.shape[1]):
        if occ[j] == 0.0:
            continue
        g = np.zeros(N, dtype=np.complex128)
        g[ks.fft_index] = x[:, j]
        r = np.fft.ifftn(g.reshape(ks.dims)) * N
        rho += weight * occ[j] * (r.real ** 2 + r.imag ** 2)
    return rho / ks.cell.volume


def scf_k(cell_rows, atoms_frac, pseudo, ecut, mesh, bands, theta, shift=(0.0, 0.0, 0.0), threshold_ry=1e-8,
          max_iterations=80, beta=0.7, history=8, kpoints=None, log=None):
    """k-point SCF with dense H_k solves, Fermi-Dirac occupations and the Gamma
    driver's mixing (oracle/scf.py): returns the Mermin free energy terms."""
    from .scf import DensityPlan, Broyden, initial_density_coefficients
    kp = kpoints if kpoints is not None else monkhorst_pack(mesh, shift)
    systems = [build_k_system(cell_rows, atoms_frac, pseudo, ecut, k) for k, _ in kp]
    weights = np.array([w for _, w in kp])
    s0 = systems[0]
    plan = DensityPlan(s0)
    electrons = s0.electrons
    ew = ewald(s0.cell, [sp.pseudo.z for sp in s0.species for _ in sp.frac],
               np.concatenate([sp.cart for sp in s0.species]))["energy"]
    x = initial_density_coefficients(s0)
    x[plan.zero] = electrons / s0.cell.volume
    mixer = Broyden(plan, beta, history)
    result = None
    for it in range(1, max_iterations + 1):
        rho_in = plan.inverse(x)
        fin = F.fields(s0.cell, rho_in)
        potential = s0.ionic_grid + fin["hxc"]
        eigs, vecs = [], []
        for ks in systems:
            w_, v_ = np.linalg.eigh(dense_k(ks, potential))
            eigs.append(w_[:bands])
            vecs.append(v_[:, :bands])
        occ, mu, ts = fermi_dirac(np.array(eigs), weights, electrons, theta)
        rho_out = sum(band_density(ks, v, occ[k], weights[k]) for k, (ks, v) in enumerate(zip(systems, vecs)))
        y = plan.forward(rho_out)
        y[plan.zero] = electrons / s0.cell.volume
        f = y - x
        dr2 = plan.dot(f, f)
        if log:
            log(f"kscf {it}: dr2={dr2:.3e} mu={mu:.8f}")
        if dr2 < threshold_ry:
            result = (eigs, vecs, occ, mu, ts, rho_out)
            break
        x = mixer.update(x, f)
    if result is None:
        raise RuntimeError("oracle k SCF did not converge")
    eigs, vecs, occ, mu, ts, rho_out = result
    kin = nl = 0.0
    for k, (ks, v) in enumerate(zip(systems, vecs)):
        for j in range(bands):
            if occ[k, j] == 0.0:
                continue
            c = weights[k] * occ[k, j]
            kin += c * float(np.real(np.vdot(v[:, j], ks.kinetic * v[:, j])))
            nl += c * float(np.real(np.vdot(v[:, j], nonlocal_apply(ks, v[:, j:j + 1])[:, 0])))
    out_fields = F.fields(s0.cell, rho_out)
    dv = s0.cell.volume / rho_out.size
    local = dv * float(np.sum(rho_out * s0.ionic_grid))
    total = kin + nl + local + out_fields["hartree"] + out_fields["xc"] + ew - ts
    return dict(energy={"kinetic": kin, "nonlocal": nl, "local": local, "hartree": out_fields["hartree"],
                        "xc": out_fields["xc"], "ewald": ew, "entropy_term": ts, "total": total},
                eigenvalues=np.array(eigs), occupations=occ, mu=mu, kpoints=kp)
