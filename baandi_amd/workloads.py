"""The synthetic cells the path is measured and tested on (seeded inputs).

Si cells use the reference's synthetic pseudopotential family
(common::potential(4, 1, true, 801): Z=4, rc=1 Bohr, s,s,p Gaussian betas,
D = [[2,.4,0],[.4,4,0],[0,0,6]] Ry; crates/dft-engine/tests/common/mod.rs:146-209).
"""
from __future__ import annotations

CONVENTIONAL = [[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0],
                [0.25, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.25, 0.75], [0.75, 0.75, 0.25]]


def supercell(reps):
    """Fractional positions of the 8-atom diamond cell repeated reps=(a,b,c)."""
    a, b, c = reps
    return [[(f[0] + i) / a, (f[1] + j) / b, (f[2] + k) / c]
            for i in range(a) for j in range(b) for k in range(c) for f in CONVENTIONAL]


def si2(a=10.26, ecut_ha=10.0):
    """Si2 diamond primitive cell (fcc a=10.26 Bohr, atoms 0 and a/4(1,1,1)), 20 Ry."""
    rows = [[0.0, a / 2, a / 2], [a / 2, 0.0, a / 2], [a / 2, a / 2, 0.0]]
    return dict(name="si2_gamma_20ry", rows=rows, atoms=[[0, 0, 0], [0.25] * 3], z=4.0, ecut_ha=ecut_ha,
                bands=8)


def si64(a=20.52, ecut_ha=15.0):
    """Si64 cubic supercell (2x2x2 of the 8-atom cell, a=20.52 Bohr), Gamma,
    30 Ry, 160 bands (128 occupied)."""
    rows = [[a, 0.0, 0.0], [0.0, a, 0.0], [0.0, 0.0, a]]
    return dict(name="si64_gamma_30ry", rows=rows, atoms=supercell((2, 2, 2)), z=4.0, ecut_ha=ecut_ha,
                bands=160)


def si128(a=10.26, ecut_ha=12.5, sigma=0.05, seed=128):
    """Si128 cell (4x2x2 conventional cells), Gamma, 25 Ry, 288 bands, atoms
    displaced by N(0, sigma) Bohr (seeded): the learned-preconditioner cell."""
    import numpy as np
    rows = [[4 * a, 0.0, 0.0], [0.0, 2 * a, 0.0], [0.0, 0.0, 2 * a]]
    atoms = np.array(supercell((4, 2, 2)), dtype=np.float64)
    rng = np.random.default_rng(seed)
    atoms = atoms + rng.normal(0.0, sigma, atoms.shape) / np.array([4 * a, 2 * a, 2 * a])
    return dict(name="si128_gamma_25ry_displaced", rows=rows, atoms=atoms.tolist(), z=4.0, ecut_ha=ecut_ha,
                bands=288)


def si512(a=10.26, ecut_ha=15.0, sigma=0.03, seed=512):
    """Si512 cell (4x4x4 conventional cells, 41.04 Bohr cube), Gamma, 30 Ry,
    1126 bands (1024 occupied), atoms displaced by N(0, sigma) Bohr (seeded):
    the combined-at-scale cell (n = 191,711, 144^3 grid, 2560 projectors)."""
    import numpy as np
    rows = [[4 * a, 0.0, 0.0], [0.0, 4 * a, 0.0], [0.0, 0.0, 4 * a]]
    atoms = np.array(supercell((4, 4, 4)), dtype=np.float64)
    rng = np.random.default_rng(seed)
    atoms = atoms + rng.normal(0.0, sigma, atoms.shape) / (4 * a)
    return dict(name="si512_gamma_30ry_displaced", rows=rows, atoms=atoms.tolist(), z=4.0, ecut_ha=ecut_ha,
                bands=1126)


def si8(a=10.26, ecut_ha=15.0):
    rows = [[a, 0.0, 0.0], [0.0, a, 0.0], [0.0, 0.0, a]]
    return dict(name="si8_gamma_30ry", rows=rows, atoms=CONVENTIONAL, z=4.0, ecut_ha=ecut_ha, bands=20)


# The aluminium variant of the synthetic family: one l=0 and one l=1 Gaussian
# beta (widths 1 and 2 Bohr^2) with a diagonal D in Ry. No reference file holds
# an Al potential; these values are this project's choice.
AL_BETAS = ((0, 1.0), (1, 2.0))
AL_DIJ_RY = ((2.0, 0.0), (0.0, 3.0))
FCC = [[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]


def al4(a=7.65, ecut_ha=12.5, mesh=(6, 6, 6)):
    """Al fcc 4-atom cubic cell (a=7.65 Bohr), Z=3, 25 Ry, 16 bands per k,
    Monkhorst-Pack mesh, Fermi-Dirac 0.01 Ha."""
    rows = [[a, 0.0, 0.0], [0.0, a, 0.0], [0.0, 0.0, a]]
    return dict(name="al4_fcc_%dx%dx%d_25ry" % tuple(mesh), rows=rows, atoms=FCC, z=3.0, ecut_ha=ecut_ha,
                bands=16, mesh=tuple(mesh), theta=0.01, betas=AL_BETAS, dij_ry=AL_DIJ_RY)


def build(workload, device=False):
    """The native set-up of a workload; ``device`` runs its atom loops on the GPU."""
    from .hamiltonian import GammaSystem, SYNTHETIC_BETAS, SYNTHETIC_DIJ_RY
    return GammaSystem(workload["rows"], workload["atoms"], workload["z"], workload["ecut_ha"],
                       betas=workload.get("betas", SYNTHETIC_BETAS), dij_ry=workload.get("dij_ry", SYNTHETIC_DIJ_RY),
                       device=device)
