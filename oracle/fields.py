"""Oracle: density, Hartree, LDA-PZ exchange-correlation and the energy audit.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Parity: KAT-1 term by term (local, Hartree, XC, kinetic, nonlocal energies of
the stored converged orbitals agree with the stored audit to <=1e-10 Ha);
LDA-PZ is additionally checked against its closed-form limits.
"""
from __future__ import annotations

import math

import numpy as np

from .physics import map_miller

DENSITY_THRESHOLD = 1e-10  # crates/dft-physics/src/xc.rs:7


def lda_pz(n):
    """Slater exchange and PZ81 correlation, unpolarised (xc.rs:48-73).
    Returns (eps_xc, v_xc) per point; zero at/below the density threshold."""
    n = np.asarray(n, dtype=np.float64)
    eps = np.zeros_like(n)
    v = np.zeros_like(n)
    on = n > DENSITY_THRESHOLD
    nn = n[on]
    rs = np.cbrt(3.0 / (4.0 * math.pi)) / np.cbrt(nn)
    eps_x = -0.75 * np.cbrt(3.0 / math.pi) * np.cbrt(nn)
    hi = rs >= 1.0
    eps_c = np.empty_like(rs)
    der = np.empty_like(rs)
    rh = rs[hi]
    den = 1.0 + 1.0529 * np.sqrt(rh) + 0.3334 * rh
    eps_c[hi] = -0.1423 / den
    der[hi] = 0.1423 * (1.0529 / (2.0 * np.sqrt(rh)) + 0.3334) / den ** 2
    rl = rs[~hi]
    eps_c[~hi] = 0.0311 * np.log(rl) - 0.048 + 0.0020 * rl * np.log(rl) - 0.0116 * rl
    der[~hi] = 0.0311 / rl + 0.0020 * (np.log(rl) + 1.0) - 0.0116
    eps[on] = eps_x + eps_c
    v[on] = 4.0 * eps_x / 3.0 + eps_c - rs * der / 3.0
    return eps, v


def reciprocal_g2(cell, dims):
    """|G|^2 of every full-grid point in canonical order, index > (n-1)/2
    wraps negative (refscf/map.rs:100-112)."""
    b = np.array(cell.reciprocal())
    ms = []
    for a in range(3):
        i = np.arange(dims[a])
        ms.append(np.where(i <= (dims[a] - 1) // 2, i, i - dims[a]).astype(np.float64))
    m0, m1, m2 = np.meshgrid(*ms, indexing="ij")
    g = [b[i, 0] * m0 + b[i, 1] * m1 + b[i, 2] * m2 for i in range(3)]
    return g[0] ** 2 + g[1] ** 2 + g[2] ** 2


def hartree_potential(cell, density):
    """Full-box Poisson solve V_H(G) = 4 pi rho(G)/G^2, zero at G=0
    (refscf/map.rs:97-114)."""
    dims = density.shape
    rho_g = np.fft.fftn(density) / density.size
    g2 = reciprocal_g2(cell, dims)
    g2[0, 0, 0] = 1.0
    vg = rho_g * (4.0 * math.pi / g2)
    vg[0, 0, 0] = 0.0
    return (np.fft.ifftn(vg) * density.size).real


def fields(cell, density):
    """Hartree + LDA-PZ potential and energies of a density on the grid
    (refscf/map.rs:41-198): returns dict with vh, vxc, hxc, hartree, xc,
    xc_integral (Ha). XC is evaluated on |rho| (map.rs:132-141)."""
    dv = cell.volume / density.size
    vh = hartree_potential(cell, density)
    absn = np.abs(density)
    eps, vxc = lda_pz(absn)
    eh = 0.5 * dv * np.sum(density * vh)
    ex = dv * np.sum(absn * eps)
    vx = dv * np.sum(density * vxc)
    return dict(vh=vh, vxc=vxc, hxc=vh + vxc, hartree=eh, xc=ex, xc_integral=vx)


def orc_band_values(system, x):
    """Real-space values psi_b(r) of ORC columns (unscaled inverse transform
    of the Hermitian completion; distributed_real_gamma.rs:770-800)."""
    from .physics import orc_unpack
    dims = system.dims
    full = orc_unpack(system.half, x)
    out = np.empty((x.shape[1],) + tuple(dims))
    ntot = dims[0] * dims[1] * dims[2]
    for b in range(x.shape[1]):
        g = np.zeros(ntot, dtype=np.complex128)
        g[system.fft_index] = full[:, b]
        out[b] = (np.fft.ifftn(g.reshape(dims)) * ntot).real
    return out


def orc_density(system, x, occupations):
    """sum_b f_b psi_b(r)^2 / Omega (band_densities, distributed_real_gamma.rs:691-735)."""
    rho = np.zeros(system.dims)
    vals = orc_band_values(system, x[:, : len(occupations)])
    for b, f in enumerate(occupations):
        if f != 0.0:
            rho += f * vals[b] ** 2 / system.cell.volume
    return rho


def complex_density(system, c, occupations):
    """sum_b f_b |psi_b(r)|^2 / Omega for complex full-sphere coefficients."""
    dims = system.dims
    ntot = dims[0] * dims[1] * dims[2]
    rho = np.zeros(dims)
    for b, f in enumerate(occupations):
        if f == 0.0:
            continue
        g = np.zeros(ntot, dtype=np.complex128)
        g[system.fft_index] = c[:, b]
        psi = np.fft.ifftn(g.reshape(dims)) * ntot
        rho += f * (psi.real ** 2 + psi.imag ** 2) / system.cell.volume
    return rho


def nonlocal_energies_complex(system, c):
    """c^H beta D beta^H c per column (nonlocal_band_energies, complex form)."""
    q = system.projectors.conj().T @ c
    e = np.zeros(c.shape[1])
    for start, nc, d in system.coupling:
        qq = q[start:start + nc]
        e += np.real(np.sum(qq.conj() * (d @ qq), axis=0))
    return e


def nonlocal_energies_orc(system, x):
    """x^T P D P^T x per column (distributed_real_gamma.rs:878-928)."""
    q = system.orc_p.T @ x
    e = np.zeros(x.shape[1])
    for start, nc, d in system.coupling:
        qq = q[start:start + nc]
        e += np.sum(qq * (d @ qq), axis=0)
    return e


def energy_audit(system, x, occupations, eigenvalues, rho_in, fields_in, ewald_energy):
    """Full energy audit of one map (refscf/map.rs:466-540) for ORC orbitals."""
    vol = system.cell.volume
    dv = vol / np.prod(system.dims)
    occ = np.asarray(occupations)
    xo = x[:, : len(occ)]
    kinetic = float(np.sum(occ[None, :] * xo ** 2 * system.kinetic_orc[:, None]))
    nonlocal_ = float(np.dot(nonlocal_energies_orc(system, xo), occ))
    rho_out = orc_density(system, xo, occ)
    f_out = fields(system.cell, rho_out)
    local = dv * float(np.sum(rho_out * system.ionic_grid))
    band_sum = float(np.dot(eigenvalues[: len(occ)], occ))
    total = kinetic + local + nonlocal_ + f_out["hartree"] + f_out["xc"] + ewald_energy
    return {"kinetic": kinetic, "local": local, "nonlocal": nonlocal_, "hartree": f_out["hartree"],
            "xc": f_out["xc"], "ewald": ewald_energy, "band_sum": band_sum, "total": total,
            "local_zero_mode": system.electrons * float(np.mean(system.ionic_grid)),
            "rho_out": rho_out}
