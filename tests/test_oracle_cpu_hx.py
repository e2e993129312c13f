"""The CPU baseline's C++ FP64 H·X (oracle/cpu_hx.cpp: OpenMP over band
pairs, pruned stick/plane FFTs) against the oracle's NumPy apply_orc on the
KAT-1, Si2 and KAT-2 cells, odd and even column counts, 1 and 4 threads.
CPU only."""
import numpy as np
import pytest

from conftest import fcc_rows

CELLS = {
    "kat1": (fcc_rows(10.261212856717933), [[0, 0, 0], [0.25] * 3], 4.0, 5.0, 801),
    "si2_20ry": (fcc_rows(10.26), [[0, 0, 0], [0.25] * 3], 4.0, 10.0, 801),
    "kat2_skew": ([[5.0, 0, 0], [0.2, 5.0, 0], [0, 0.1, 5.0]], [[0.13, 0.21, 0.32]], 2.0, 1.0, 161),
}


@pytest.mark.parametrize("cell", list(CELLS))
def test_cpu_hx_matches_oracle(cell):
    from baandi_amd.hamiltonian import GammaSystem
    from oracle import physics as P, cpu_hx
    from oracle.hamiltonian import apply_orc
    rows, atoms, z, ecut, pts = CELLS[cell]
    s = GammaSystem(rows, atoms, z, ecut, mesh_points=pts)
    o = P.build_gamma_system(rows, atoms, P.synthetic_pseudo(z, 1.0, pts), ecut)
    h = cpu_hx.CpuHamiltonian(s)
    rng = np.random.default_rng(5)
    pot = o.ionic_grid + 0.1 * rng.standard_normal(o.dims)
    h.set_potential(pot)
    for threads in (1, 4):
        h.set_threads(threads)
        for m in (1, 6, 7):
            x = rng.standard_normal((s.n, m))
            ref = apply_orc(o, x, pot)
            err = np.abs(h.apply(x) - ref).max() / np.abs(ref).max()
            assert err < 1e-12, (threads, m, err)


def test_cpu_hx_plan_is_pruned():
    """Only the occupied sticks and planes are transformed: the Si2 sphere
    at 20 Ry on the 24^3 grid has 73 sticks in 11 planes (the GPU plan's S, Z)."""
    from baandi_amd.hamiltonian import GammaSystem
    from oracle import cpu_hx
    rows, atoms, z, ecut, pts = CELLS["si2_20ry"]
    s = GammaSystem(rows, atoms, z, ecut, mesh_points=pts)
    assert cpu_hx.CpuHamiltonian(s, threads=1).plan() == {"sticks": 73, "z_planes": 11}
