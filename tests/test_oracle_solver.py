"""Pins of the oracle's H·X, Davidson and SCF: dense brute force, the
free-electron spectrum, LAPACK eigenvalues and the reference goldens. CPU only."""
import math

import numpy as np
import pytest

from conftest import fcc_rows
from oracle import physics as P
from oracle import hamiltonian as H
from oracle import davidson as D
from oracle.scf import scf


@pytest.fixture(scope="module")
def si2_20ry():
    pp = P.synthetic_pseudo(4.0, 1.0, 801)
    return P.build_gamma_system(fcc_rows(10.26), [[0, 0, 0], [0.25] * 3], pp, 10.0)


def test_fft_apply_matches_dense(si2_20ry):
    """FFT H·X vs V(G-G') + T + V_NL assembled densely (hamiltonian.rs:288-417: 1e-11)."""
    s = si2_20ry
    hd = H.dense_orc(s, s.ionic_grid)
    assert np.abs(hd - hd.T).max() < 1e-12
    x = np.random.default_rng(2).standard_normal((len(s.basis.g2), 9))  # odd: one unpaired column
    y = H.apply_orc(s, x, s.ionic_grid)
    assert np.abs(y - hd @ x).max() < 1e-11 * np.abs(y).max()


def test_free_electron_spectrum():
    """Constant V0 and no projectors: eigenvalues |G|^2/2 + V0 (hamiltonian.rs:395-405)."""
    pp = P.synthetic_pseudo(4.0, 1.0, 401, nonlocal_=False)
    s = P.build_gamma_system([[6.0, 0, 0], [0.3, 7.0, 0], [0, 0.2, 8.0]], [[0.1, 0.2, 0.3]], pp, 4.0)
    v0 = -0.37
    pot = np.full(s.dims, v0)
    hd = H.dense_orc(s, pot)
    w = np.linalg.eigvalsh(hd)
    assert np.abs(w - (np.sort(0.5 * s.basis.g2) + v0)).max() < 1e-12


def test_pair_packing_equals_single_columns(si2_20ry):
    """The pair trick (two real bands in one complex FFT) equals transforming
    each Hermitian-completed column alone."""
    s = si2_20ry
    rng = np.random.default_rng(9)
    xa, xb = rng.standard_normal(len(s.basis.g2)), rng.standard_normal(len(s.basis.g2))
    h = H.local_apply_full(s, H.pack_pair(s, xa, xb), s.ionic_grid)
    ya, yb = H.unpack_pair(s, h)
    for x, y in ((xa, ya), (xb, yb)):
        c = P.orc_unpack(s.half, x)
        ref = P.orc_pack(s.half, H.local_apply_full(s, c, s.ionic_grid))
        assert np.abs(ref - y).max() < 1e-13


def test_denominator_identity():
    """sqrt and hypot forms of the QE denominator agree (ref_davidson.rs:245-255)."""
    rng = np.random.default_rng(4)
    h = rng.uniform(-5, 50, 20000)
    e = rng.uniform(-3, 5, 20000)
    a = D.sqrt_denominator(h, e)
    b = D.preconditioner_denominator(h, e)
    assert np.all(a > 0)
    assert np.abs(a - b).max() <= 8 * np.finfo(float).eps * np.abs(b).max()


def test_denominator_reference_ry_form():
    """The reference's independent Ry expression (crates/dft-solvers/tests/
    ref_davidson.rs:238-256): with x = 2(h - e), d_Ry = (1 + x + sqrt(1 +
    (x-1)^2))/2 equals 2 d_Ha; d depends on h - e only; at e = 0 the direct
    form 0.25 (1 + 2d + sqrt(1 + (2d - 1)^2)) holds from -1e8 to 100. A slip
    of the Ha -> Ry factor in x fails the first check."""
    for h, e in [(-0.7, -0.2), (0.35, -0.15), (2.1, 0.6)]:
        x = 2.0 * (h - e)
        d_ry = (1.0 + x + math.sqrt(1.0 + (x - 1.0) ** 2)) / 2.0
        for f in (D.sqrt_denominator, D.preconditioner_denominator):
            a = float(f(np.array(h), np.array(e)))
            assert abs(2.0 * a - d_ry) < 2e-15
            assert abs(a - float(f(np.array(h + 0.25), np.array(e + 0.25)))) < 2e-15
    for d in [-1e8, -10.0, -1.0, 0.0, 0.5, 1.0, 100.0]:
        direct = 0.25 * (1.0 + 2.0 * d + math.sqrt(1.0 + (2.0 * d - 1.0) ** 2))
        for f in (D.sqrt_denominator, D.preconditioner_denominator):
            p = float(f(np.array(d), np.array(0.0)))
            assert np.isfinite(p) and p > 0.0
            assert abs(p - direct) < 1e-8


def test_tpa_weight_closed_form():
    """Teter-Payne-Allan (Phys. Rev. B 40, 12255 (1989)): K(x) = p/(p + 16 x^4),
    p = 27 + 18x + 12x^2 + 8x^3, x = T_G/<T>, evaluated here in exact rational
    arithmetic. The oracle's two branches (x <= 1 direct, x > 1 in y = 1/x;
    crates/dft-solvers/src/lib.rs:197-207) match it to 2 ulp, with K(0) = 1,
    K(1) = 65/81, K ~ 1/(2x) for large x, and the kinetic scale's 0.1 Ha floor
    (lib.rs:178-193)."""
    from fractions import Fraction as Fr

    def exact(x):
        x = Fr(x)
        p = 27 + 18 * x + 12 * x ** 2 + 8 * x ** 3
        return p / (p + 16 * x ** 4)

    xs = [0.0, 1e-3, 0.25, 0.5, 0.999, 1.0, 1.001, 1.5, 2.0, 7.25, 40.0, 1e3]
    scale = 0.375
    got = D.tpa_weight(np.array(xs) * scale, scale)
    for x, k in zip(xs, got):
        ref = float(exact(Fr(x * scale) / Fr(scale)))
        assert abs(k - ref) <= 2 * np.finfo(float).eps * ref, (x, k, ref)
    assert D.tpa_weight(np.array([0.0]), 1.0)[0] == 1.0
    assert abs(D.tpa_weight(np.array([1.0]), 1.0)[0] - 65.0 / 81.0) < 1e-16
    big = D.tpa_weight(np.array([1e6]), 1.0)[0]
    assert abs(big * 2e6 - 1.0) < 1e-5
    w = np.array([1.0, 3.0])
    assert D.kinetic_scale(w, np.array([0.2, 1.0])) == pytest.approx((0.2 + 3.0) / 4.0, rel=1e-15)
    assert D.kinetic_scale(w, np.array([0.01, 0.02])) == 0.1


@pytest.mark.parametrize("pc", ["qe", "tpa"])
def test_davidson_matches_dense_eigensolve(si2_20ry, pc):
    """Reference Davidson vs dense EVD, 1e-11 Ha (crates/dft-solvers/tests/ref_davidson.rs:223)."""
    s = si2_20ry
    hd = H.dense_orc(s, s.ionic_grid)
    w = np.linalg.eigvalsh(hd)
    dav = D.Davidson(lambda x: H.apply_orc(s, x, s.ionic_grid), np.diag(hd), 8, 1e-13, 4,
                     first_step=True, max_iterations=200, preconditioner=pc, kinetic=s.kinetic_orc)
    r = dav.solve(D.random_orc_start(len(s.basis.g2), 8, 1))
    assert np.abs(r["eigenvalues"] - w[:8]).max() < 1e-11
    res = r["hx"] - r["x"] * r["eigenvalues"][None, :]
    assert np.abs(r["x"].T @ r["x"] - np.eye(8)).max() < 1e-12
    assert np.sqrt((res ** 2).sum(0)).max() < 1e-5


def test_ideal_reset_keeps_the_fixed_point(si2_20ry):
    """The ideal reset (S = I, H_r = diag(theta) after a restart whose columns
    have unit norms; ref_davidson/ideal_reset.rs) against measured refreshes
    at every restart: the same converged eigenvalues (1e-11 Ha vs dense), with
    ideal resets taken, every fourth contraction measured and no guard trip."""
    s = si2_20ry
    hd = H.dense_orc(s, s.ionic_grid)
    w = np.linalg.eigvalsh(hd)
    out = {}
    for ideal in (True, False):
        dav = D.Davidson(lambda x: H.apply_orc(s, x, s.ionic_grid), np.diag(hd), 8, 1e-13, 4, first_step=True,
                         max_iterations=200, preconditioner="qe", kinetic=s.kinetic_orc, ideal_reset=ideal)
        out[ideal] = dav.solve(D.random_orc_start(len(s.basis.g2), 8, 1))
        assert np.abs(out[ideal]["eigenvalues"] - w[:8]).max() < 1e-11
    r = out[True]
    assert r["ideal_resets"] > 0 and r["guard_fallbacks"] == 0
    assert r["periodic_refreshes"] == r["restarts"] // 4
    assert r["ideal_resets"] + r["periodic_refreshes"] == r["restarts"]
    assert out[False]["ideal_resets"] == 0 and out[False]["periodic_refreshes"] == 0


def test_ideal_reset_guard_trips_on_unnormalised_columns(si2_20ry):
    """Columns off unit norm by more than 1e-12 at a restart: the norm guard
    trips and every later restart is a measured refresh (workspace.rs:814-835)."""
    s = si2_20ry
    hd = H.dense_orc(s, s.ionic_grid)
    dav = D.Davidson(lambda x: H.apply_orc(s, x, s.ionic_grid), np.diag(hd), 8, 1e-13, 4, first_step=True,
                     max_iterations=200, preconditioner="qe", kinetic=s.kinetic_orc)
    original = dav._contract

    def skewed(m):  # scale the Ritz coefficients of the first restart by 1 + 1e-9
        if dav.contractions == 0:
            dav.C = dav.C * (1.0 + 1e-9)
        original(m)
    dav._contract = skewed
    r = dav.solve(D.random_orc_start(len(s.basis.g2), 8, 1))
    assert r["guard_fallbacks"] == 1 and r["ideal_resets"] == 0 and r["restarts"] > 1
    assert np.abs(r["eigenvalues"] - np.linalg.eigvalsh(hd)[:8]).max() < 1e-11


def test_kat1_scf_fixed_point(kat1):
    """KAT-1: the stored certified energy and eigenvalues of reference-1."""
    pp = P.synthetic_pseudo(4.0, 1.0, 801)
    s = P.build_gamma_system(fcc_rows(kat1["cell_a_bohr"]), [[0, 0, 0], [0.25] * 3], pp, kat1["ecut_ha"])
    r = scf(s, 8, threshold_ry=1e-13, final_tight=1e-13)
    assert abs(r["energy"]["total"] - kat1["certified_energy_ha"]) < 1e-9
    # The stored eigenvalues were converged only to ~1e-8 (degenerate
    # triplet split by 6e-9 in the file).
    assert np.abs(r["eigenvalues"] - np.array(kat1["eigenvalues_ha"])).max() < 2e-8
    for k in ("kinetic", "local", "nonlocal", "hartree", "xc"):
        assert abs(r["energy"][k] - kat1["energy_terms_ha"][k]) < 1e-7, k


def test_kat2_skew_cell_energy():
    """KAT-2: reference-results.json cli_runs[0] (spread of stored runs 4.5e-11)."""
    pp = P.synthetic_pseudo(2.0, 1.0, 161)
    s = P.build_gamma_system([[5.0, 0, 0], [0.2, 5.0, 0], [0, 0.1, 5.0]], [[0.13, 0.21, 0.32]], pp, 1.0)
    r = scf(s, 4, threshold_ry=1e-13, final_tight=1e-13)
    assert abs(r["energy"]["total"] - (-0.8646074276373095)) < 1e-9
