"""Host logic of the k-point driver (baandi_amd/kpoint.py) against the oracle:
Monkhorst-Pack points and weights, Fermi-Dirac occupations, chemical
potential and entropy, including the capacity-1 form over realified levels;
and the Gamma SCF's check that smeared occupations have room for every
electron."""
import numpy as np

from baandi_amd import kpoint as K
from oracle import kpoint as OK


def test_monkhorst_pack_matches_oracle():
    for mesh in [(1, 1, 1), (2, 2, 2), (3, 3, 3), (4, 2, 3), (6, 6, 6)]:
        a, b = K.monkhorst_pack(mesh), OK.monkhorst_pack(mesh)
        assert [p for p, _ in a] == [p for p, _ in b]
        assert np.allclose([w for _, w in a], [w for _, w in b], rtol=0, atol=1e-15)
    assert len(K.monkhorst_pack((6, 6, 6))) == 108
    assert K.wrap_fractional((0.5, -0.5, 1.25)) == (-0.5, -0.5, 0.25)


def test_fermi_dirac_matches_oracle():
    rng = np.random.default_rng(4)
    e = np.sort(rng.uniform(-0.4, 0.6, (5, 9)), axis=1)
    w = rng.uniform(0.5, 1.5, 5)
    w /= w.sum()
    for theta, electrons in [(0.01, 7.0), (0.05, 3.3), (1e-3, 12.0)]:
        occ, mu, ts = K.fermi_dirac(e, w, electrons, theta)
        ro, rmu, rts = OK.fermi_dirac(e, w, electrons, theta)
        assert abs(mu - rmu) < 1e-12 and abs(ts - rts) < 1e-13 and np.abs(occ - ro).max() < 1e-12


def test_capacity_one_over_doubled_levels_equals_capacity_two():
    """Every realified level twice with capacity 1 gives the complex levels'
    capacity-2 occupations, mu and entropy."""
    rng = np.random.default_rng(8)
    e = np.sort(rng.uniform(-0.3, 0.5, (3, 6)), axis=1)
    w = np.array([0.25, 0.5, 0.25])
    occ2, mu2, ts2 = K.fermi_dirac(e, w, 5.0, 0.02)
    occ1, mu1, ts1 = K.fermi_dirac(np.repeat(e, 2, axis=1), w, 5.0, 0.02, capacity=1.0)
    assert abs(mu1 - mu2) < 1e-12 and abs(ts1 - ts2) < 1e-13
    assert np.abs(occ1[:, 0::2] + occ1[:, 1::2] - occ2).max() < 1e-12


def test_gamma_smearing_needs_more_bands_than_electron_pairs():
    """Fermi-Dirac occupations cannot hold every electron in electrons/2
    bands: the Gamma SCF rejects that before any device work."""
    import pytest
    from baandi_amd import lib as L
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.scf import run_scf, ScfOptions
    from conftest import fcc_rows
    s = GammaSystem(fcc_rows(10.26), [[0, 0, 0], [0.25] * 3], 4.0, 5.0)
    with pytest.raises(L.BaandiError, match="more bands"):
        run_scf(s, ScfOptions(bands=4, theta=0.01))
