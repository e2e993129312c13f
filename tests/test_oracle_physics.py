"""Pins of the oracle's set-up layer: reference goldens (KAT-1), closed forms
and library special functions. CPU only."""
import math

import numpy as np
import pytest
from scipy.special import spherical_jn

from conftest import fcc_rows
from oracle import physics as P
from oracle import fields as F


@pytest.fixture(scope="module")
def si2(kat1):
    pp = P.synthetic_pseudo(4.0, 1.0, 801)
    return P.build_gamma_system(fcc_rows(kat1["cell_a_bohr"]), [[0, 0, 0], [0.25] * 3], pp, kat1["ecut_ha"])


def kat1_orbitals(kat1):
    v = np.array(kat1["orbitals_colmajor_re_im"])
    return (v[:, 0] + 1j * v[:, 1]).reshape(kat1["cols"], kat1["rows"]).T


def test_kat1_basis_and_grid(si2, kat1):
    # 137 rows of the stored orbitals; 15^3 density_shape of arrays.json.
    assert len(si2.basis.g2) == kat1["rows"]
    assert si2.dims == tuple(kat1["density_shape"])
    assert si2.basis.miller[0].tolist() == [0, 0, 0]
    assert np.all(np.diff(si2.basis.g2) >= 0)


def test_kat1_energy_terms_from_stored_orbitals(si2, kat1):
    """Each audited term of the stored converged state, recomputed from the
    stored orbitals, matches the reference audit (refscf/map.rs:466-540)."""
    c = kat1_orbitals(kat1)
    occ = np.array(kat1["occupations"])
    t = kat1["energy_terms_ha"]
    kinetic = np.sum(occ[None, :] * np.abs(c) ** 2 * 0.5 * si2.basis.g2[:, None])
    nonlocal_ = F.nonlocal_energies_complex(si2, c) @ occ
    rho = F.complex_density(si2, c, occ)
    dv = si2.cell.volume / rho.size
    f = F.fields(si2.cell, rho)
    assert abs(rho.sum() * dv - 8.0) < 1e-12
    assert abs(kinetic - t["kinetic"]) < 1e-12
    assert abs(nonlocal_ - t["nonlocal"]) < 1e-12
    assert abs(dv * np.sum(rho * si2.ionic_grid) - t["local"]) < 1e-12
    assert abs(f["hartree"] - t["hartree"]) < 1e-12
    assert abs(f["xc"] - t["xc"]) < 1e-12
    assert abs(si2.electrons * np.mean(si2.ionic_grid) - t["local_zero_mode"]) < 1e-12


def test_kat1_ewald_breakdown(si2, kat1):
    ref = kat1["ewald"]
    e = P.ewald(si2.cell, [4.0, 4.0], si2.species[0].cart)
    for k in ("energy", "real", "reciprocal", "self_energy", "background", "eta"):
        assert abs(e[k] - ref[k]) < 1e-12, k
    # The total is independent of the splitting parameter.
    e2 = P.ewald(si2.cell, [4.0, 4.0], si2.species[0].cart, eta=0.61)
    assert abs(e2["energy"] - ref["energy"]) < 1e-11


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_bessel_matches_scipy(l):
    x = np.concatenate([np.linspace(0, 4, 401), np.linspace(4, 60, 301)])
    assert np.allclose(P.bessel(l, x), spherical_jn(l, x), rtol=1e-13, atol=1e-15)


def test_real_harmonics_closed_forms():
    d = np.random.default_rng(3).standard_normal((50, 3))
    r = np.linalg.norm(d, axis=1)
    c1 = math.sqrt(3.0 / (4.0 * math.pi))
    assert np.allclose(P.real_ylm(0, 0, d), 1.0 / math.sqrt(4.0 * math.pi))
    assert np.allclose(P.real_ylm(1, -1, d), c1 * d[:, 1] / r)
    assert np.allclose(P.real_ylm(1, 0, d), c1 * d[:, 2] / r)
    assert np.allclose(P.real_ylm(1, 1, d), c1 * d[:, 0] / r)
    c2 = math.sqrt(15.0 / (4.0 * math.pi))
    assert np.allclose(P.real_ylm(2, -2, d), c2 * d[:, 0] * d[:, 1] / r ** 2)


@pytest.mark.parametrize("index,l,width", [(0, 0, 1.0), (1, 0, 2.0), (2, 1, 3.0)])
def test_projector_transform_gaussian_hankel(index, l, width):
    """4 pi int r^(l+2) e^{-r^2/w} j_l(qr) dr = 4 pi sqrt(pi) q^l e^{-q^2 w/4} / (2^(l+2) a^(l+3/2)), a=1/w."""
    pp = P.synthetic_pseudo(4.0, 1.0, 801)
    q = np.linspace(0.0, 6.0, 25)
    a = 1.0 / width
    exact = 4 * math.pi * math.sqrt(math.pi) * q ** l * np.exp(-q * q / (4 * a)) / (2 ** (l + 2) * a ** (l + 1.5))
    got = P.projector_transform(pp, index, q)
    assert np.allclose(got, exact, rtol=0, atol=2e-9 * np.abs(exact).max())


def test_local_transform_screened_coulomb_closed_form():
    """V = -Z erf(r/rc)/r: V(G) = -4 pi Z e^{-G^2 rc^2/4}/(Omega G^2), alpha_Z = pi Z rc^2/Omega."""
    for rc in (1.0, 1.4):
        pp = P.synthetic_pseudo(3.0, rc, 1601)
        vol = 123.0
        g = np.linspace(0.3, 8.0, 30)
        exact = -4 * math.pi * 3.0 * np.exp(-g * g * rc * rc / 4) / (vol * g * g)
        assert np.allclose(P.vloc_g(pp, vol, g), exact, rtol=0, atol=1e-9)
        assert abs(P.alpha_z(pp, vol) - math.pi * 3.0 * rc * rc / vol) < 1e-9


def test_orc_isometry(si2):
    """half.rs:179-202: unpack/pack round trip, inner products preserved."""
    n = len(si2.basis.g2)
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    a, b = P.orc_unpack(si2.half, x), P.orc_unpack(si2.half, y)
    assert np.allclose(P.orc_pack(si2.half, a), x, atol=1e-15)
    assert abs(np.vdot(a, b).real - x @ y) < 1e-13 and abs(np.vdot(a, b).imag) < 1e-13
    # Mirror is an involution; representatives are lexicographically positive.
    assert np.array_equal(si2.half.mirror[si2.half.mirror], np.arange(n))


def test_lda_pz_limits():
    """Threshold zero; the PZ81 branch values at rs = 1 (-0.0596 high-density
    side, -0.1423/2.3863 low-density side) and the Slater exchange closed form."""
    eps, v = F.lda_pz(np.array([1e-11]))
    assert eps[0] == 0.0 and v[0] == 0.0
    n1 = 3.0 / (4 * math.pi)  # rs = 1 exactly
    for n, ec in ((n1 * (1 + 1e-12), -0.0596), (n1 * (1 - 1e-12), -0.1423 / 2.3863)):
        e, _ = F.lda_pz(np.array([n]))
        ex = -0.75 * (3.0 / math.pi) ** (1 / 3) * n ** (1 / 3)
        assert abs(e[0] - ex - ec) < 1e-9
    # v_xc = d(n eps)/dn by central differences.
    n = np.array([0.003, 0.05, 0.7])
    h = 1e-6 * n
    ep, _ = F.lda_pz(n + h)
    em, _ = F.lda_pz(n - h)
    _, v = F.lda_pz(n)
    assert np.allclose(((n + h) * ep - (n - h) * em) / (2 * h), v, rtol=1e-7)
