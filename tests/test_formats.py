"""On-disk formats (baandi_amd.formats) on the CPU.

Fixtures under tests/golden/si30/ are the reference's own files:
tests/reference/si30_ours/ours/{density.bin,density.json,results.json} (a
Si2 run at 30 Ry by the reference CLI) and tests/fixtures_format/results.json
(the minimal record its tooling accepts)."""
import hashlib
import json
import math
import os
import struct

import numpy as np
import pytest

from conftest import fcc_rows

HERE = os.path.dirname(os.path.abspath(__file__))
SI30 = os.path.join(HERE, "golden", "si30")


def test_reference_density_files_decode():
    """The reference's density.bin: digest, size and ordering as its
    density.json states; the integral of the grid reproduces the stored
    charge 8.000000000000016 (C order, little-endian f64, electrons/Bohr^3)."""
    from baandi_amd import formats as F
    rho, meta = F.read_density(SI30)
    assert rho.shape == (25, 25, 25)
    integral = rho.sum() * meta["volume_bohr3"] / rho.size
    assert abs(integral - meta["normalization"]["integral_electrons"]) < 1e-13
    assert abs(integral - 8.0) < 1e-12
    # The Si2 cell is symmetric under cyclic permutations of its three
    # lattice vectors, and so is a decoded density (to the SCF's own
    # convergence, ~1e-9 relative); a dtype or byte-order slip breaks it.
    assert np.abs(rho - np.transpose(rho, (1, 2, 0))).max() < 1e-8 * rho.max()


def test_density_roundtrip_and_no_clobber(tmp_path):
    from baandi_amd import formats as F
    rng = np.random.default_rng(3)
    rho = np.abs(rng.standard_normal((6, 5, 4)))
    rows = fcc_rows(10.26)
    meta = F.write_density(str(tmp_path), rho, rows, 270.10716120683634, 8.0)
    back, meta2 = F.read_density(str(tmp_path))
    assert np.array_equal(back, rho) and meta2 == meta
    assert meta["lattice_columns_bohr"][0] == [r[0] for r in rows]
    with open(tmp_path / "density.json") as f:
        ref_keys = json.load(open(os.path.join(SI30, "density.json"))).keys()
        assert json.load(f).keys() == ref_keys
    with pytest.raises(FileExistsError):
        F.write_density(str(tmp_path), rho, rows, 1.0, 8.0)
    raw = bytearray(open(tmp_path / "density.bin", "rb").read())
    raw[3] ^= 1
    open(tmp_path / "density.bin", "wb").write(bytes(raw))
    with pytest.raises(F.FormatError):
        F.read_density(str(tmp_path))


def test_state_codec_roundtrip_and_errors():
    """u64 LE, f64 bits (signed zero and subnormals kept), bool tags,
    length-prefixed strings/vectors, options, column-major matrices."""
    from baandi_amd import formats as F
    w = F.StateWriter()
    w.u64(2 ** 63 + 5)
    w.f64(-0.0)
    w.f64(5e-324)
    w.boolean(True)
    w.string("Gamma ψ")
    w.f64s([1.5, -2.25])
    w.complexes([1 + 2j, -3.5j])
    w.option(None, w.u64)
    w.option(7, w.u64)
    m = np.arange(6.0).reshape(2, 3)  # two columns of three rows
    w.real_matrix(m)
    data = w.bytes()
    assert data[:8] == struct.pack("<Q", 2 ** 63 + 5)
    assert data[8:16] == struct.pack("<Q", 1 << 63)
    r = F.StateReader(data)
    assert r.u64() == 2 ** 63 + 5
    z = r.f64()
    assert z == 0.0 and math.copysign(1.0, z) < 0
    assert r.f64() == 5e-324
    assert r.boolean() is True
    assert r.string() == "Gamma ψ"
    assert r.f64s().tolist() == [1.5, -2.25]
    assert r.complexes().tolist() == [1 + 2j, -3.5j]
    assert r.option(r.u64) is None and r.option(r.u64) == 7
    mm = r.real_matrix()
    assert mm.shape == (2, 3) and np.array_equal(mm, m)
    r.finish()
    # rows, cols then values: rows = 3, cols = 2.
    off = len(data) - 8 * 8
    assert struct.unpack("<QQ", data[off:off + 16]) == (3, 2)
    with pytest.raises(F.FormatError):
        F.StateReader(data + b"\0").finish() if F.StateReader(data + b"\0").u64() else None
    rr = F.StateReader(data)
    rr.cursor = len(data) - 4
    with pytest.raises(F.FormatError):
        rr.u64()
    with pytest.raises(F.FormatError):
        F.StateReader(struct.pack("<Q", 2)).boolean()
    with pytest.raises(F.FormatError):  # claims more elements than bytes remain
        F.StateReader(struct.pack("<Q", 10) + b"\0" * 16).f64s()


def _identity():
    return {"input_sha256": "1" * 64, "upf_sha256": ["2" * 64], "binary_sha256": "3" * 64,
            "platform": "synthetic-test", "threads": 1, "ranks": 1}


def test_v2_container_rules():
    """The BAANDICP envelope and the v2 header rules: kinds, rank count,
    contiguous state ranges, result energy/hash/iteration, checksum."""
    from baandi_amd import formats as F
    state = b"\x01" * 40
    blob = F.encode_v2(_identity(), 3, [state])
    assert blob[:8] == b"BAANDICP"
    h, b = struct.unpack("<QQ", blob[8:24])
    assert len(blob) == 24 + h + b + 32 and blob[-32:] == hashlib.sha256(blob[:-32]).digest()
    head = json.loads(blob[24:24 + h])
    assert list(head) == ["schema_version", "kind", "identity", "iteration", "states", "result"]
    assert head["kind"] == "ExactContinuation" and head["states"] == [{"offset": 0, "count": 40}]
    d = F.decode_v2(blob)
    assert d["states"] == [state] and d["iteration"] == 3 and d["result"] is None
    bad = bytearray(blob)
    bad[30] ^= 1
    with pytest.raises(F.FormatError):
        F.decode_v2(bytes(bad))
    with pytest.raises(F.FormatError):
        F.decode_v2(blob[:-1])
    with pytest.raises(F.FormatError):  # two states for one rank
        F.encode_v2(_identity(), 3, [state, state])
    ident = dict(_identity(), binary_sha256="xyz")
    with pytest.raises(F.FormatError):
        F.encode_v2(ident, 3, [state])
    res = F.result_record(-4.5, np.ones(8), np.ones((2, 3), dtype=complex), True, 7, 0)
    blob2 = F.encode_v2(_identity(), 7, result=res)
    d2 = F.decode_v2(blob2)
    assert d2["kind"] == "ResultExport" and d2["states"] == [] and d2["result"] == res
    assert struct.unpack("<d", struct.pack("<Q", res["energy_ha_bits"]))[0] == -4.5
    assert res["density_sha256"] == hashlib.sha256(np.ones(8).tobytes()).hexdigest()
    with pytest.raises(F.FormatError):  # result iterations must equal the header's iteration
        F.encode_v2(_identity(), 6, result=res)
    # An envelope with a well-formed header naming bytes it does not cover.
    head = dict(json.loads(blob[24:24 + h]), states=[{"offset": 0, "count": 8}])
    with pytest.raises(F.FormatError):
        F.decode_v2(F.envelope(head, state))


def test_complex_orbitals_equal_oracle_unpack(kat1):
    """Delivered complex orbitals from ORC rows equal the oracle's Hermitian
    completion (half.rs unpack) bit for bit, keep the norm and satisfy
    c(-G) = conj c(G)."""
    from baandi_amd import formats as F
    from baandi_amd.hamiltonian import GammaSystem
    from oracle import physics as P
    rows, atoms = fcc_rows(kat1["cell_a_bohr"]), [[0, 0, 0], [0.25] * 3]
    s = GammaSystem(rows, atoms, 4.0, kat1["ecut_ha"])
    o = P.build_gamma_system(rows, atoms, P.synthetic_pseudo(4.0, 1.0, 801), kat1["ecut_ha"])
    x = np.random.default_rng(4).standard_normal((5, s.n))
    c = F.gamma_complex_orbitals(s, x)
    ref = np.ascontiguousarray(P.orc_unpack(o.half, x.T).T)
    assert np.array_equal(c.view(np.float64), ref.view(np.float64))
    assert np.allclose(np.linalg.norm(c, axis=1), np.linalg.norm(x, axis=1), rtol=1e-14)
    minus = s.get("minus_basis")
    plus = s.get("plus_basis")
    assert np.array_equal(c[:, minus[1:]], np.conj(c[:, plus[1:]]))


def test_si2_basis_matches_reference_run():
    """The reference's Si2 run at 30 Ry (its results.json: a = 10.2612 Bohr)
    has 749 plane waves, 5985 density G, a 25^3 grid and a 270.107... Bohr^3
    cell; the native set-up of the same cell gives the same counts."""
    from baandi_amd.hamiltonian import GammaSystem
    ref = json.load(open(os.path.join(SI30, "results.json")))
    rows = np.array(ref["cell"]["lattice_columns_bohr"]).T
    atoms = [a["fractional"] for a in ref["atoms"]]
    s = GammaSystem(rows, atoms, 4.0, ref["cutoff"]["ecutwfc_ha"])
    assert s.n == ref["basis"]["n_plane_waves"] == 749
    assert s.ndensity == ref["basis"]["n_density_g"] == 5985
    assert list(s.dims) == ref["basis"]["fft_grid"]
    assert abs(s.volume - ref["cell"]["volume_bohr3"]) < 1e-12 * s.volume


def test_results_record_fields(kat1):
    """Our results.json carries every field of the reference's format
    fixture with the same JSON types, and its top-level keys are the
    reference CLI's (its Si2 results.json)."""
    from baandi_amd import formats as F
    from baandi_amd.hamiltonian import GammaSystem
    s = GammaSystem(fcc_rows(kat1["cell_a_bohr"]), [[0, 0, 0], [0.25] * 3], 4.0, kat1["ecut_ha"])
    run = {"energy": {k: -1.0 for k in ("kinetic", "local", "nonlocal", "hartree", "xc", "ewald", "total",
                                         "band_sum")},
           "eigenvalues": np.linspace(-0.2, 0.3, 8), "occupations": np.array([2.0] * 4 + [0.0] * 4),
           "iterations": 2, "relative_charge_error": 1e-15, "max_eigen_residual_ha": 3e-9,
           "history": [{"iteration": 1, "dr2_ry": 1e-3, "ethr_ry": 1e-2, "davidson_iterations": 4,
                        "h_columns": 40, "mixed": False, "warm_start": False},
                       {"iteration": 2, "dr2_ry": 1e-9, "ethr_ry": 1e-5, "davidson_iterations": 3,
                        "h_columns": 20, "mixed": False, "warm_start": True}]}
    rec = F.results_record(s, run, {"ecut_ha": kat1["ecut_ha"], "density_residual_tolerance": 1e-8})
    rec = json.loads(json.dumps(rec))
    fixture = json.load(open(os.path.join(SI30, "format_fixture_results.json")))

    def check(want, got, path):
        for k, v in want.items():
            if k == "is_format_fixture":
                continue
            assert k in got, path + k
            if isinstance(v, dict):
                check(v, got[k], path + k + ".")
            else:
                assert type(got[k]) is type(v) or (isinstance(v, float) and isinstance(got[k], (int, float))), \
                    path + k
    check(fixture, rec, "")
    reference = json.load(open(os.path.join(SI30, "results.json")))
    extra = set(rec) - set(reference)
    assert extra == set(), extra
    assert rec["cell"]["lattice_columns_bohr"] == np.array(fcc_rows(kat1["cell_a_bohr"])).T.tolist()
    assert rec["eigenvalues"]["ev"][0] == rec["eigenvalues"]["ha"][0] * F.EV_PER_HA
    assert rec["energies"]["total"]["ry"] == 2.0 * rec["energies"]["total"]["ha"]
