"""GPU parity of the reference Davidson and the Gamma SCF against the FP64
oracle and the reference goldens, through the C ABI.

Bars: eigenvalues <= 1e-10 Ha against a dense FP64 eigensolve;
SCF total energy <= 1e-8 Ha/atom against the stored goldens (KAT-1 certified
-4.684561184277879 Ha; KAT-2 -0.8646074276373095 Ha)."""
import numpy as np
import pytest

from conftest import fcc_rows

pytestmark = pytest.mark.gpu


SI8 = ([[10.26, 0, 0], [0, 10.26, 0], [0, 0, 10.26]],
       [[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
        [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])


def si2(ecut=10.0, a=10.26):
    import torch
    from baandi_amd.hamiltonian import GammaSystem
    from oracle import physics as P
    rows, atoms = fcc_rows(a), [[0, 0, 0], [0.25] * 3]
    s = GammaSystem(rows, atoms, 4.0, ecut)
    o = P.build_gamma_system(rows, atoms, P.synthetic_pseudo(4.0, 1.0, 801), ecut)
    return torch, s, o


def si8(ecut=6.0):
    import torch
    from baandi_amd.hamiltonian import GammaSystem
    from oracle import physics as P
    s = GammaSystem(*SI8, 4.0, ecut)
    o = P.build_gamma_system(*SI8, P.synthetic_pseudo(4.0, 1.0, 801), ecut)
    return torch, s, o


@pytest.mark.parametrize("cell,bands,occupied,precond",
                         [("si2", 8, 4, 0), ("si2", 8, 4, 1), ("si8", 24, 16, 0), ("si8", 24, 16, 1)])
def test_davidson_matches_dense_eigensolve(cell, bands, occupied, precond):
    """Same H, same seeded start: GPU Davidson vs dense LAPACK. Occupied
    bands converge at tau = 1e-12 Ha and must match to <= 1e-10 Ha; empty
    bands use the reference's max(5 tau, 5e-6) rule (ref_davidson.rs:70-77)."""
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    from baandi_amd import ref_davidson
    from oracle.hamiltonian import dense_orc
    torch, s, o = si2() if cell == "si2" else si8()
    print(f"[davidson parity] cell={cell} n={s.n} bands={bands} precond={precond}", flush=True)
    pot = o.ionic_grid + 0.05 * np.cos(np.arange(o.ionic_grid.size)).reshape(o.dims)
    hd = dense_orc(o, pot)
    w = np.linalg.eigvalsh(hd)
    op = DistributedRealHamiltonian(s, capacity=bands)
    op.set_owned_potential(pot.reshape(-1))
    g = np.random.default_rng(1).standard_normal((bands, s.n))
    x0 = torch.from_numpy(g / np.linalg.norm(g, axis=1, keepdims=True)).cuda()
    diag = torch.from_numpy(s.get("diagonal_static") + pot.mean()).cuda()
    opts = ref_davidson.Options(bands=bands, threshold=1e-12, occupied=occupied, max_iterations=300,
                                return_hx=True, preconditioner=precond)
    sol = ref_davidson.solve_with_initial_hx(op, x0, diag, opts)
    print(f"[davidson parity] record={sol.record}", flush=True)
    assert sol.converged.all()
    assert np.abs(sol.eigenvalues[:occupied] - w[:occupied]).max() < 1e-10
    assert np.abs(sol.eigenvalues - w[:bands]).max() < 1e-3  # empty bands: Ritz change < 5e-6 only
    x = sol.eigenvectors.cpu().numpy()
    hx = sol.hx.cpu().numpy()
    assert np.abs(x @ x.T - np.eye(bands)).max() < 1e-10
    assert np.abs(hx - (hd @ x.T).T).max() < 1e-10


@pytest.mark.parametrize("cell,bands,occupied", [("si2", 8, 4), ("si8", 24, 16)])
@pytest.mark.parametrize("precond", ["qe", "tpa"])
def test_davidson_first_iterations_match_oracle(cell, bands, occupied, precond):
    """The Ritz values after 1, 2 and 3 iterations from the same start depend
    on every correction (QE denominator, or TPA weight with its kinetic
    scale), unlike converged values. GPU and oracle Davidson (the reference
    iteration, oracle/davidson.py) agree to 1e-10 Ha and take the same H
    columns; the threshold is below reach, so every band stays active."""
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    from baandi_amd import ref_davidson, lib as L
    from oracle.hamiltonian import apply_orc
    from oracle import davidson as OD
    torch, s, o = si2() if cell == "si2" else si8()
    pot = o.ionic_grid + 0.05 * np.cos(np.arange(o.ionic_grid.size)).reshape(o.dims)
    op = DistributedRealHamiltonian(s, capacity=bands)
    op.set_owned_potential(pot.reshape(-1))
    g = np.random.default_rng(3).standard_normal((bands, s.n))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    diag = s.get("diagonal_static") + pot.mean()
    code = L.BD_PRECOND_QE if precond == "qe" else L.BD_PRECOND_TPA
    for its in (1, 2, 3, 6):
        opts = ref_davidson.Options(bands=bands, threshold=1e-15, occupied=occupied, max_iterations=its,
                                    preconditioner=code)
        sol = ref_davidson.solve(op, torch.from_numpy(g).cuda(), torch.from_numpy(diag).cuda(), opts)
        ref = OD.Davidson(lambda x: apply_orc(o, x, pot), diag, bands, 1e-15, occupied, max_iterations=its,
                          preconditioner=precond, kinetic=s.get("kinetic_orc")).solve(g.T.copy())
        err = np.abs(sol.eigenvalues - ref["eigenvalues"]).max()
        rec = sol.record
        print(f"[davidson trajectory] {cell} {precond} iterations={its}: max |dE| {err:.2e} Ha, H columns "
              f"{rec['h_columns']}/{ref['h_columns']}, restarts {rec['restarts']}/{ref['restarts']} "
              f"(ideal {rec['ideal_resets']}/{ref['ideal_resets']}, periodic "
              f"{rec['periodic_refreshes']}/{ref['periodic_refreshes']})", flush=True)
        assert rec["iterations"] == ref["iterations"] == its
        assert rec["h_columns"] == ref["h_columns"]
        for key in ("restarts", "repairs", "ideal_resets", "periodic_refreshes", "guard_fallbacks"):
            assert rec[key] == ref[key], key
        assert err < 1e-10
    # A different preconditioner moves the iteration-3 values well above that bar.
    other = OD.Davidson(lambda x: apply_orc(o, x, pot), diag, bands, 1e-15, occupied, max_iterations=3,
                        preconditioner="tpa" if precond == "qe" else "qe",
                        kinetic=s.get("kinetic_orc")).solve(g.T.copy())
    assert np.abs(other["eigenvalues"] - ref["eigenvalues"]).max() > 1e-6


def test_davidson_rejects_dependent_start():
    """Equal start columns: InvalidParameter before any iteration
    (crates/dft-solvers/tests/ref_davidson.rs:345-356)."""
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    from baandi_amd import ref_davidson, lib as L
    torch, s, o = si2()
    op = DistributedRealHamiltonian(s, capacity=4)
    op.set_owned_potential(o.ionic_grid.reshape(-1))
    x0 = torch.ones((4, s.n), dtype=torch.float64).cuda() / np.sqrt(s.n)
    diag = torch.from_numpy(s.get("diagonal_static")).cuda()
    with pytest.raises(L.BaandiError) as e:
        ref_davidson.solve(op, x0, diag, ref_davidson.Options(bands=4, threshold=1e-8, occupied=2))
    assert e.value.status == L.BD_ERR_INVALID_PARAMETER
    assert "dependent" in str(e.value)


@pytest.mark.parametrize("precond", ["qe", "tpa"])
def test_davidson_rank_repair_when_capacity_is_short(precond):
    """A basis of 59 plane waves (Si2 at 2.5 Ha) and 40 bands: the capacity
    is the dimension, so the 40 corrections of the first iteration do not
    fit beside the 40 columns and the rank repair runs (workspace.rs:586-771).
    The first iteration equals the oracle's (same repair, H columns and Ritz
    values; the repaired basis spans the whole space, so later corrections are
    rounding noise whose 32 eps drops differ run to run), and the solve
    converges to the dense spectrum."""
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    from baandi_amd import ref_davidson, lib as L
    from oracle.hamiltonian import apply_orc, dense_orc
    from oracle import davidson as OD
    torch, s, o = si2(ecut=2.5)
    assert s.n == 59
    bands, occupied = 40, 4
    pot = o.ionic_grid + 0.05 * np.cos(np.arange(o.ionic_grid.size)).reshape(o.dims)
    op = DistributedRealHamiltonian(s, capacity=bands)
    op.set_owned_potential(pot.reshape(-1))
    g = np.random.default_rng(8).standard_normal((bands, s.n))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    diag = s.get("diagonal_static") + pot.mean()
    code = L.BD_PRECOND_QE if precond == "qe" else L.BD_PRECOND_TPA
    for its in (1,):
        sol = ref_davidson.solve(op, torch.from_numpy(g).cuda(), torch.from_numpy(diag).cuda(),
                                 ref_davidson.Options(bands=bands, threshold=1e-15, occupied=occupied,
                                                      max_iterations=its, preconditioner=code))
        ref = OD.Davidson(lambda x: apply_orc(o, x, pot), diag, bands, 1e-15, occupied, max_iterations=its,
                          preconditioner=precond, kinetic=s.get("kinetic_orc")).solve(g.T.copy())
        err = np.abs(sol.eigenvalues - ref["eigenvalues"]).max()
        print(f"[repair] {precond} iterations={its}: repairs {sol.record['repairs']}/{ref['repairs']}, "
              f"H columns {sol.record['h_columns']}/{ref['h_columns']}, max |dE| {err:.2e}", flush=True)
        assert sol.record["repairs"] == ref["repairs"] >= 1
        assert sol.record["h_columns"] == ref["h_columns"]
        assert err < 1e-10
    w = np.linalg.eigvalsh(dense_orc(o, pot))
    sol = ref_davidson.solve(op, torch.from_numpy(g).cuda(), torch.from_numpy(diag).cuda(),
                             ref_davidson.Options(bands=bands, threshold=1e-12, occupied=bands, max_iterations=100,
                                                  preconditioner=code, return_hx=True))
    print(f"[repair] {precond} converged: record={sol.record}", flush=True)
    assert sol.converged.all() and sol.record["repairs"] >= 1
    assert np.abs(sol.eigenvalues - w[:bands]).max() < 1e-10


def test_davidson_profile_history():
    """Options.profile: one row of Ritz changes and (previous, current)
    values per iteration (Solution::ritz_changes_ha / ritz_values_ha); the
    last row's current values are the returned eigenvalues and its changes
    fall under each converged band's threshold."""
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    from baandi_amd import ref_davidson
    torch, s, o = si8()
    bands, occupied = 24, 16
    op = DistributedRealHamiltonian(s, capacity=bands)
    op.set_owned_potential(o.ionic_grid.reshape(-1))
    g = np.random.default_rng(4).standard_normal((bands, s.n))
    x0 = torch.from_numpy(g / np.linalg.norm(g, axis=1, keepdims=True)).cuda()
    diag = torch.from_numpy(s.get("diagonal_static") + o.ionic_grid.mean()).cuda()
    sol = ref_davidson.solve(op, x0, diag, ref_davidson.Options(bands=bands, threshold=1e-9, occupied=occupied,
                                                                max_iterations=60, profile=True))
    its = sol.record["iterations"]
    print(f"[profile] iterations={its}, record={sol.record}", flush=True)
    assert len(sol.ritz_changes_ha) == len(sol.ritz_values_ha) == its > 1
    last = sol.ritz_values_ha[-1]
    assert np.array_equal(last[:, 1], sol.eigenvalues)
    assert np.array_equal(sol.ritz_changes_ha[-1], np.abs(last[:, 1] - last[:, 0]))
    assert np.all(sol.ritz_changes_ha[-1][sol.converged] < np.array(sol.band_thresholds_ha)[sol.converged])
    for a, b in zip(sol.ritz_values_ha[:-1], sol.ritz_values_ha[1:]):
        assert np.array_equal(a[:, 1], b[:, 0])
    assert sol.record["ideal_resets"] > 0
    assert sol.record["restarts"] >= sol.record["ideal_resets"] + sol.record["periodic_refreshes"]


def test_davidson_rejects_nonfinite_start():
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    from baandi_amd import ref_davidson, lib as L
    torch, s, o = si2()
    op = DistributedRealHamiltonian(s, capacity=4)
    pot = o.ionic_grid.copy()
    pot[0, 0, 0] = np.nan  # a NaN potential makes H X non-finite
    op.set_owned_potential(pot.reshape(-1))
    x0 = torch.eye(4, s.n, dtype=torch.float64).cuda()
    diag = torch.from_numpy(s.get("diagonal_static")).cuda()
    with pytest.raises(L.BaandiError) as e:
        ref_davidson.solve(op, x0, diag, ref_davidson.Options(bands=4, threshold=1e-8, occupied=2))
    assert e.value.status == L.BD_ERR_APPLICATION


def test_davidson_iteration_limit_is_reported():
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    from baandi_amd import ref_davidson
    torch, s, o = si2()
    op = DistributedRealHamiltonian(s, capacity=8)
    op.set_owned_potential(o.ionic_grid.reshape(-1))
    g = np.random.default_rng(2).standard_normal((8, s.n))
    x0 = torch.from_numpy(g).cuda()
    diag = torch.from_numpy(s.get("diagonal_static")).cuda()
    sol = ref_davidson.solve(op, x0, diag, ref_davidson.Options(bands=8, threshold=1e-14, occupied=4,
                                                                max_iterations=2))
    assert sol.record["iterations"] == 2 and not sol.converged.all()
    with pytest.raises(ref_davidson.NotConverged):
        ref_davidson.solve(op, x0, diag, ref_davidson.Options(bands=8, threshold=1e-14, occupied=4,
                                                              max_iterations=2), raise_not_converged=True)


def test_scf_kat1_golden(kat1):
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.scf import run_scf, ScfOptions
    s = GammaSystem(fcc_rows(kat1["cell_a_bohr"]), [[0, 0, 0], [0.25] * 3], 4.0, kat1["ecut_ha"])
    r = run_scf(s, ScfOptions(bands=8, threshold_ry=1e-13, final_threshold=1e-13))
    e = r["energy"]
    print(f"[kat1] E={e['total']!r} iterations={r['iterations']}", flush=True)
    assert abs(e["total"] - kat1["certified_energy_ha"]) / 2 < 1e-8  # per atom
    assert abs(e["total"] - kat1["certified_energy_ha"]) < 1e-9
    # The stored eigenvalues carry the reference's own certificate of 2.78e-8 Ha
    # (summary.json certificate_occupied_eigenvalue_change_hartree).
    assert np.abs(r["eigenvalues"] - np.array(kat1["eigenvalues_ha"])).max() < 5e-8
    # Single terms are first order in the density error of the stored run
    # (density residual 1.1e-8); the total above is variational. The formulas
    # of each term are pinned to 1e-12 by the stored-orbital audit test.
    for k in ("kinetic", "local", "nonlocal", "hartree", "xc", "ewald"):
        assert abs(e[k] - kat1["energy_terms_ha"][k]) < 5e-7, k


def test_scf_warm_start_matches_fresh(kat1):
    """KeptHx: the SCF whose maps start from H_old X + dV X (the kept pair
    fields of the previous band density) and the SCF with a fresh H·X of
    every start block reach the same fixed point; every map after the first
    starts warm."""
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.scf import run_scf, ScfOptions
    s = GammaSystem(fcc_rows(kat1["cell_a_bohr"]), [[0, 0, 0], [0.25] * 3], 4.0, kat1["ecut_ha"])
    warm = run_scf(s, ScfOptions(bands=8, threshold_ry=1e-13))
    fresh = run_scf(s, ScfOptions(bands=8, threshold_ry=1e-13, kept_hx=False))
    print(f"[kept] KAT-1 warm {warm['energy']['total']!r} ({warm['totals']['warm_starts']} warm maps, "
          f"{warm['totals']['h_columns']} H columns) fresh {fresh['energy']['total']!r} "
          f"({fresh['totals']['h_columns']} H columns)", flush=True)
    assert fresh["totals"]["warm_starts"] == 0
    assert warm["totals"]["warm_starts"] >= warm["iterations"] - 1
    assert all(h["warm_start"] for h in warm["history"][1:])
    assert abs(warm["energy"]["total"] - fresh["energy"]["total"]) < 1e-10
    assert np.abs(warm["eigenvalues"][:4] - fresh["eigenvalues"][:4]).max() < 1e-7


def test_scf_kat2_golden():
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.scf import run_scf, ScfOptions
    s = GammaSystem([[5.0, 0, 0], [0.2, 5.0, 0], [0, 0.1, 5.0]], [[0.13, 0.21, 0.32]], 2.0, 1.0, mesh_points=161)
    r = run_scf(s, ScfOptions(bands=4, threshold_ry=1e-13))
    assert abs(r["energy"]["total"] - (-0.8646074276373095)) < 1e-9


def test_scf_smeared_matches_oracle_k_at_gamma(kat1):
    """KAT-1 cell, 8 bands, Fermi-Dirac 0.01 Ha: the real Gamma SCF with
    smearing against the oracle's k SCF on the 1x1x1 mesh (dense solves of
    the complex H at k = 0). The cell is metallic at Gamma, so five bands
    carry fractional occupations."""
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.scf import run_scf, ScfOptions
    from oracle import physics as P
    from oracle import kpoint as OK
    rows, atoms = fcc_rows(kat1["cell_a_bohr"]), [[0, 0, 0], [0.25] * 3]
    s = GammaSystem(rows, atoms, 4.0, kat1["ecut_ha"])
    r = run_scf(s, ScfOptions(bands=8, theta=0.01, threshold_ry=1e-11))
    ref = OK.scf_k(rows, atoms, P.synthetic_pseudo(4.0, 1.0, 801), kat1["ecut_ha"], (1, 1, 1), 8, 0.01,
                   threshold_ry=1e-11)
    de = abs(r["energy"]["total"] - ref["energy"]["total"])
    deig = np.abs(r["eigenvalues"] - np.ravel(ref["eigenvalues"])).max()
    docc = np.abs(r["occupations"] - np.ravel(ref["occupations"])).max()
    print(f"[smeared gamma] E={r['energy']['total']:.12f} oracle={ref['energy']['total']:.12f} dE={de:.2e} "
          f"TS={r['energy']['entropy_term']:.3e} mu={r['mu']:.8f}/{ref['mu']:.8f} eig err {deig:.2e} "
          f"occ err {docc:.2e} maps={r['iterations']}", flush=True)
    assert abs(r["occupations"].sum() - 8.0) < 1e-10
    assert 0.0 < (r["occupations"] * (2.0 - r["occupations"])).sum()  # fractional bands present
    for key in ("kinetic", "nonlocal", "local", "hartree", "xc", "entropy_term"):
        assert abs(r["energy"][key] - ref["energy"][key]) < 1e-5, key
    assert de < 1e-8 * len(atoms)
    assert deig < 1e-6 and docc < 1e-4
    assert abs(r["mu"] - ref["mu"]) < 1e-6


def test_scf_si8_matches_oracle_scf():
    """Cubic Si8 at 6 Ha: GPU SCF vs oracle SCF from their own seeded starts."""
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.scf import run_scf, ScfOptions
    from oracle import physics as P
    from oracle.scf import scf as oracle_scf
    rows, atoms = SI8
    s = GammaSystem(rows, atoms, 4.0, 6.0)
    r = run_scf(s, ScfOptions(bands=20, threshold_ry=1e-13, final_threshold=1e-13))
    o = P.build_gamma_system(rows, atoms, P.synthetic_pseudo(4.0, 1.0, 801), 6.0)
    ro = oracle_scf(o, 20, threshold_ry=1e-13, final_tight=1e-13)
    print(f"[si8] gpu={r['energy']['total']!r} oracle={ro['energy']['total']!r}", flush=True)
    assert abs(r["energy"]["total"] - ro["energy"]["total"]) / 8 < 1e-8
    # Eigenvalues are first order in the residual density error of each run
    # (degenerate clusters split by ~1e-8 at dr2 < 1e-13 Ry); the 1e-10 Ha
    # eigenvalue bar is checked at a fixed H in test_davidson_matches_dense_eigensolve.
    assert np.abs(r["eigenvalues"][:16] - ro["eigenvalues"][:16]).max() < 1e-6
