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


def si8(a=10.26, ecut_ha=15.0):
    rows = [[a, 0.0, 0.0], [0.0, a, 0.0], [0.0, 0.0, a]]
    return dict(name="si8_gamma_30ry", rows=rows, atoms=CONVENTIONAL, z=4.0, ecut_ha=ecut_ha, bands=20)


def build(workload):
    from .hamiltonian import GammaSystem
    return GammaSystem(workload["rows"], workload["atoms"], workload["z"], workload["ecut_ha"])
