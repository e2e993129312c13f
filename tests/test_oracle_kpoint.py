"""The k-point oracle (oracle/kpoint.py) pinned by what the mathematics fixes:
the k = 0 spectrum of the complex H equals the Gamma (ORC) spectrum, free
electrons at k, the FFT apply equals the explicit H_k, time reversal, the
Fermi-Dirac closed forms, and the Gamma-only k SCF reproduces KAT-1."""
import math

import numpy as np
import pytest

from conftest import fcc_rows
from oracle import kpoint as K
from oracle import physics as P
from oracle.hamiltonian import dense_orc


@pytest.fixture(scope="module")
def si2():
    return fcc_rows(10.26), [[0, 0, 0], [0.25] * 3], P.synthetic_pseudo(4.0, 1.0, 801)


def test_k0_spectrum_equals_gamma_orc(si2):
    rows, atoms, pp = si2
    g = P.build_gamma_system(rows, atoms, pp, 6.0)
    ks = K.build_k_system(rows, atoms, pp, 6.0, (0.0, 0.0, 0.0))
    pot = g.ionic_grid
    wg = np.linalg.eigvalsh(dense_orc(g, pot))
    wk = np.linalg.eigvalsh(K.dense_k(ks, pot))
    assert len(wg) == len(wk)
    assert np.abs(wg - wk).max() < 1e-10


def test_free_electrons_at_k(si2):
    rows, atoms, _ = si2
    pp = P.synthetic_pseudo(4.0, 1.0, 801, nonlocal_=False)
    ks = K.build_k_system(rows, atoms, pp, 4.0, (0.25, -0.125, 0.375))
    w = np.linalg.eigvalsh(K.dense_k(ks, np.zeros(ks.dims)))
    assert np.abs(w - np.sort(0.5 * ks.basis.q2)).max() < 1e-12
    assert np.all(0.5 * ks.basis.q2 <= 4.0 * (1 + 1e-12))


def test_fft_apply_equals_dense_and_time_reversal(si2):
    rows, atoms, pp = si2
    k = (0.2, 0.1, -0.3)
    ks = K.build_k_system(rows, atoms, pp, 5.0, k)
    kn = K.build_k_system(rows, atoms, pp, 5.0, tuple(-x for x in k))
    rng = np.random.default_rng(0)
    pot = ks.ionic_grid + 0.05 * rng.standard_normal(ks.dims)
    h = K.dense_k(ks, pot)
    x = rng.standard_normal((len(ks.kinetic), 3)) + 1j * rng.standard_normal((len(ks.kinetic), 3))
    assert np.abs(K.apply_k(ks, x, pot) - h @ x).max() < 1e-11 * np.abs(h).max()
    assert np.allclose(np.diag(h).real - K.static_diagonal(ks) - np.mean(pot), 0.0, atol=1e-10)
    w1 = np.linalg.eigvalsh(h)
    w2 = np.linalg.eigvalsh(K.dense_k(kn, pot))
    assert np.abs(w1 - w2).max() < 1e-10


def test_monkhorst_pack_weights():
    kp = K.monkhorst_pack((2, 2, 2))
    assert abs(sum(w for _, w in kp) - 1.0) < 1e-15
    assert len(kp) == 4  # 8 points, pairs k/-k
    kp6 = K.monkhorst_pack((6, 6, 6))
    assert abs(sum(w for _, w in kp6) - 1.0) < 1e-14
    assert len(kp6) == 108  # 216 points, none its own inverse on an even unshifted-centre mesh
    assert K.wrap_fractional((0.5, -0.5, 1.25)) == (-0.5, -0.5, 0.25)


def test_fermi_dirac_closed_forms():
    theta = 0.01
    a = 0.003
    occ, mu, ts = K.fermi_dirac(np.array([[-a, a]]), [1.0], 2.0, theta)
    f = 1.0 / (1.0 + math.exp(-a / theta))
    assert abs(mu) < 1e-12 and abs(occ[0, 0] - 2.0 * f) < 1e-12 and abs(occ[0, 1] - 2.0 * (1 - f)) < 1e-12
    s = -f * math.log(f) - (1 - f) * math.log(1 - f)
    assert abs(ts - theta * 2.0 * 2.0 * s) < 1e-14
    # degenerate level half filled: entropy 2 ln 2 per state pair
    occ, mu, ts = K.fermi_dirac(np.array([[0.1, 0.1]]), [1.0], 2.0, theta)
    assert np.allclose(occ, 1.0) and abs(ts - theta * 4.0 * math.log(2.0)) < 1e-14
    # weighted k points, exact count
    e = np.array([[-0.2, 0.01, 0.3], [-0.1, 0.02, 0.25]])
    occ, mu, ts = K.fermi_dirac(e, [0.25, 0.75], 3.0, theta)
    assert abs((occ * np.array([[0.25], [0.75]])).sum() - 3.0) < 1e-12


def test_gamma_only_k_scf_reproduces_kat1(kat1):
    """A Gamma-only 'mesh' on the KAT-1 cell gives the certified Gamma energy.
    Its gap at Gamma is 8 mHa, so the smearing is 1e-5 Ha: every fraction is
    then 0 or 1 to e^-400 and the entropy term vanishes."""
    pp = P.synthetic_pseudo(4.0, 1.0, 801)
    r = K.scf_k(fcc_rows(kat1["cell_a_bohr"]), [[0, 0, 0], [0.25] * 3], pp, kat1["ecut_ha"], None, 8, 1e-5,
                threshold_ry=1e-12, kpoints=[((0.0, 0.0, 0.0), 1.0)])
    assert abs(r["energy"]["entropy_term"]) < 1e-12
    assert abs(r["energy"]["total"] - kat1["certified_energy_ha"]) < 1e-8
    assert np.abs(r["eigenvalues"][0] - np.array(kat1["eigenvalues_ha"])).max() < 5e-8


@pytest.mark.parametrize("pc", ["qe", "tpa"])
def test_complex_davidson_matches_dense_h_k(si2, pc):
    """The oracle Davidson on complex columns (ref_davidson_complex.rs:641-1113)
    converges to the dense H_k spectrum; its Ritz vectors are orthonormal in
    the Hermitian inner product; with a capacity of twice the bands it
    contracts (ideal resets) on the way."""
    from oracle import davidson as OD
    rows, atoms, pp = si2
    ks = K.build_k_system(rows, atoms, pp, 5.0, (0.2, 0.1, -0.3))
    pot = ks.ionic_grid + 0.05 * np.random.default_rng(0).standard_normal(ks.dims)
    h = K.dense_k(ks, pot)
    w = np.linalg.eigvalsh(h)
    bands = 8
    rng = np.random.default_rng(5)
    x0 = rng.standard_normal((len(ks.kinetic), bands)) + 1j * rng.standard_normal((len(ks.kinetic), bands))
    x0 /= np.linalg.norm(x0, axis=0)
    diag = K.static_diagonal(ks) + pot.mean()
    d = OD.Davidson(lambda x: K.apply_k(ks, x, pot), diag, bands, 1e-12, bands, max_iterations=300,
                    preconditioner=pc, kinetic=ks.kinetic)
    r = d.solve(x0)
    x = r["x"]
    print(f"[complex oracle davidson] {pc}: nk={len(ks.kinetic)} iterations={r['iterations']} "
          f"h_columns={r['h_columns']} restarts={r['restarts']} ideal={r['ideal_resets']} "
          f"error={np.abs(r['eigenvalues'] - w[:bands]).max():.1e}")
    assert all(r["converged"])
    assert np.abs(r["eigenvalues"] - w[:bands]).max() < 1e-10
    assert np.abs(x.conj().T @ x - np.eye(bands)).max() < 1e-10
    assert np.linalg.norm(h @ x - x * r["eigenvalues"][None, :], axis=0).max() < 1e-5
    assert r["restarts"] >= 1 and r["ideal_resets"] >= 1


def test_complex_davidson_rank_repair(si2):
    """More bands than half the plane waves: the capacity is the dimension,
    the first corrections do not fit and the rank repair runs; the solve
    still reaches the dense spectrum."""
    from oracle import davidson as OD
    rows, atoms, pp = si2
    ks = K.build_k_system(rows, atoms, pp, 1.5, (0.25, -0.125, 0.375))
    nk = len(ks.kinetic)
    pot = ks.ionic_grid
    w = np.linalg.eigvalsh(K.dense_k(ks, pot))
    bands = nk // 2 + 3
    rng = np.random.default_rng(2)
    x0 = rng.standard_normal((nk, bands)) + 1j * rng.standard_normal((nk, bands))
    x0 /= np.linalg.norm(x0, axis=0)
    d = OD.Davidson(lambda x: K.apply_k(ks, x, pot), K.static_diagonal(ks) + pot.mean(), bands, 1e-12, bands,
                    max_iterations=100)
    r = d.solve(x0)
    print(f"[complex oracle repair] nk={nk} bands={bands} repairs={r['repairs']} iterations={r['iterations']}")
    assert r["repairs"] >= 1 and all(r["converged"])
    assert np.abs(r["eigenvalues"] - w[:bands]).max() < 1e-10
