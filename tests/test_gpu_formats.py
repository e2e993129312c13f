"""Checkpoint v2 continuation and result export of the Gamma SCF on MI355X.

A run stopped at a map boundary, written as an ExactContinuation file, read
back and resumed gives the bits of the uninterrupted run (energy, eigenvalues,
density, per-map record), as the reference requires of its v2 files
(crates/dft-cli/tests/checkpoint_v2.rs: file identity against an
uninterrupted trajectory). The warm-start entry crosses the file (its kept
fields are recomputed from X); the mixed-tier case stops while the mixed
operator holds the entry."""
import json
import struct

import numpy as np
import pytest

from conftest import fcc_rows

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tier,stop", [("fp64", 4), ("mixed", 2)])
def test_scf_checkpoint_resume_is_bitwise(kat1, tmp_path, tier, stop):
    from baandi_amd import formats as F, lib as L
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.scf import run_scf, ScfOptions
    s = GammaSystem(fcc_rows(kat1["cell_a_bohr"]), [[0, 0, 0], [0.25] * 3], 4.0, kat1["ecut_ha"])
    opts = ScfOptions(bands=8, threshold_ry=1e-12,
                      precision=L.BD_MIXED if tier == "mixed" else L.BD_FP64)
    full = run_scf(s, opts)
    part = run_scf(s, opts, stop_after=stop)
    st = part["state"]
    assert st["completed"] == stop and len(st["history"]) == stop and st["warm"] is not None
    if tier == "mixed":
        assert st["warm"]["op"] == "mixed"
    path = tmp_path / "scf.ckpt"
    size = F.write_checkpoint(str(path), s, opts, st)
    back = F.read_checkpoint(str(path), s, opts)
    resumed = run_scf(s, opts, resume=back)
    bits = lambda v: struct.pack("<d", v)
    print(f"[ckpt] {tier}: stop after {stop} of {full['iterations']} maps, file {size} bytes; "
          f"E full {full['energy']['total']!r} resumed {resumed['energy']['total']!r}", flush=True)
    assert bits(resumed["energy"]["total"]) == bits(full["energy"]["total"])
    assert np.array_equal(resumed["eigenvalues"], full["eigenvalues"])
    assert np.array_equal(resumed["density"], full["density"])
    assert resumed["history"] == full["history"]
    assert resumed["totals"]["warm_starts"] == full["totals"]["warm_starts"] - part["totals"]["warm_starts"]
    # Files: no clobbering, other options refused, result exports agree.
    with pytest.raises(FileExistsError):
        F.write_checkpoint(str(path), s, opts, st)
    other = ScfOptions(bands=8, threshold_ry=1e-11, precision=opts.precision)
    with pytest.raises(F.FormatError):
        F.read_checkpoint(str(path), s, other)
    r1 = F.export_result(str(tmp_path / "full.v2"), s, opts, full)
    r2 = F.export_result(str(tmp_path / "resumed.v2"), s, opts, resumed)
    assert r1 == r2 and r1["converged"] and r1["iterations"] == full["iterations"]
    assert F.decode_v2(open(tmp_path / "full.v2", "rb").read())["result"] == r1


def test_scf_results_and_density_files(kat1, tmp_path):
    """results.json and density.bin/json of a KAT-1 run: the format
    fixture's fields, the certified energy, and a density whose file
    integral is the electron count."""
    from baandi_amd import formats as F
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.scf import run_scf, ScfOptions
    rows = fcc_rows(kat1["cell_a_bohr"])
    s = GammaSystem(rows, [[0, 0, 0], [0.25] * 3], 4.0, kat1["ecut_ha"])
    opts = ScfOptions(bands=8, threshold_ry=1e-13)
    r = run_scf(s, opts)
    rec = F.results_record(s, r, {"ecut_ha": kat1["ecut_ha"], "density_residual_tolerance": opts.threshold_ry})
    (tmp_path / "results.json").write_text(json.dumps(rec, indent=2))
    back = json.loads((tmp_path / "results.json").read_text())
    assert abs(back["energies"]["total"]["ha"] - kat1["certified_energy_ha"]) < 1e-9
    assert back["convergence"]["last_iteration"]["max_eigen_residual_ha"] < 1e-5
    assert back["convergence"]["last_iteration"]["relative_charge_error"] < 1e-12
    meta = F.write_density(str(tmp_path), r["density"], rows, s.volume, s.electrons)
    rho, _ = F.read_density(str(tmp_path))
    assert np.array_equal(rho, r["density"])
    print(f"[files] KAT-1 results.json E={back['energies']['total']['ha']!r}, residual "
          f"{back['convergence']['last_iteration']['max_eigen_residual_ha']:.2e}, density integral "
          f"{meta['normalization']['integral_electrons']!r}", flush=True)
    assert abs(meta["normalization"]["integral_electrons"] - 8.0) < 1e-10
