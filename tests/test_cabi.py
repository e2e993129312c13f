"""CPU checks of the C ABI library: it loads, exports every symbol the header
declares, and its native set-up (basis order, grid, ORC maps, projector rows,
V_ion(G), Davidson diagonal, Ewald) matches the FP64 oracle. No GPU calls."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, fcc_rows
from oracle import physics as P
from oracle.scf import davidson_diagonal

HEADER = os.path.join(ROOT, "include", "baandi_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from baandi_amd import lib as L
    lib = L.load()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(L.EXPORTS)


CELLS = {
    "kat1": (fcc_rows(10.261212856717933), [[0, 0, 0], [0.25] * 3], 4.0, 5.0, 801),
    "si2_20ry": (fcc_rows(10.26), [[0, 0, 0], [0.25] * 3], 4.0, 10.0, 801),
    "kat2_skew": ([[5.0, 0, 0], [0.2, 5.0, 0], [0, 0.1, 5.0]], [[0.13, 0.21, 0.32]], 2.0, 1.0, 161),
}


@pytest.mark.parametrize("name", list(CELLS))
def test_native_setup_matches_oracle(name):
    from baandi_amd.hamiltonian import GammaSystem
    rows, atoms, z, ecut, points = CELLS[name]
    s = GammaSystem(rows, atoms, z, ecut, mesh_points=points)
    o = P.build_gamma_system(rows, atoms, P.synthetic_pseudo(z, 1.0, points), ecut)
    assert s.n == len(o.basis.g2) and s.dims == o.dims and s.nproj == o.orc_p.shape[1]
    assert np.array_equal(s.get("miller"), o.basis.miller)
    assert np.array_equal(s.get("density_miller"), o.density.miller)
    assert np.abs(s.get("g2") - o.basis.g2).max() == 0.0
    assert np.abs(s.get("kinetic_orc") - o.kinetic_orc).max() < 1e-14
    assert np.abs(s.get("projectors_orc") - o.orc_p).max() < 1e-13
    assert np.abs(s.get("diagonal_static") - davidson_diagonal(o)).max() < 1e-12
    vi = s.get("ionic_density_re") + 1j * s.get("ionic_density_im")
    assert np.abs(vi - P.ionic_coefficients(o.density, o.cell.volume, o.species)).max() < 1e-13
    flat = P.flat_index(P.map_miller(o.basis.miller, o.dims), o.dims)
    assert np.array_equal(s.get("plus_index"), flat[o.half.plus])
    assert np.array_equal(s.get("minus_index"), flat[o.half.minus])
    ew = P.ewald(o.cell, [z] * len(atoms), o.species[0].cart)["energy"]
    assert abs(s.get("ewald")[0] - ew) < 1e-11


def test_invalid_descriptors_are_errors():
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd import lib as L
    with pytest.raises(L.BaandiError) as e:
        GammaSystem(fcc_rows(10.26), [[0, 0, 0]], 4.0, -1.0)
    assert e.value.status == L.BD_ERR_INVALID_PARAMETER
    with pytest.raises(L.BaandiError):
        GammaSystem([[1, 0, 0], [2, 0, 0], [0, 0, 1]], [[0, 0, 0]], 4.0, 5.0)  # singular lattice


def test_operator_needs_a_gfx950_device():
    """Without a GPU the operator fails loudly (no CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from baandi_amd.hamiltonian import GammaSystem, DistributedRealHamiltonian
    from baandi_amd import lib as L
    s = GammaSystem(*CELLS["kat2_skew"][:4], mesh_points=161)
    with pytest.raises(L.BaandiError) as e:
        DistributedRealHamiltonian(s, capacity=4)
    assert e.value.status == L.BD_ERR_DEVICE


@pytest.mark.parametrize("mode", [0, 1])
def test_gemm_workspace_covers_smaller_shapes(mode):
    """One split-K buffer serves every A^T B up to its sizing shape: the size
    is monotone in M and N at a fixed K and never below M N. (Fewer output
    tiles take more K splits, so the split count alone is not monotone: the
    Davidson and operator buffers are sized once at the subspace capacity.)"""
    from baandi_amd import lib as L
    lib = L.load()
    for K in (100, 500, 23847, 36525, 200000):
        for mmax, nmax in ((640, 576), (576, 576), (320, 160), (64, 64)):
            top = lib.bd_gemm_workspace(mode, 1, mmax, nmax, K)
            for m in list(range(1, mmax + 1, 7)) + [mmax]:
                prev = 0
                for n in list(range(1, nmax + 1, 5)) + [nmax]:
                    w = lib.bd_gemm_workspace(mode, 1, m, n, K)
                    assert m * n <= w <= top, (mode, K, mmax, nmax, m, n, w, top)
                    if mode == 1:
                        # The BF16x3 TN kernel: 128 x 128 tiles, about 512
                        # workgroups, chunks of >= 256 rows (multiples of 32),
                        # FP32 partials two to a double.
                        tiles = -(-m // 128) * -(-n // 128)
                        splits = max(1, min(-(-512 // tiles), max(1, K // 256), 256))
                        chunk = -(-(-(-K // splits)) // 32) * 32
                        used = -(-K // chunk)
                        assert 2 * top >= used * m * n
                    assert w >= prev
                    prev = w
