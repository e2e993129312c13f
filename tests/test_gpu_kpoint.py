"""k-point path on MI355X against the oracle (oracle/kpoint.py): the basis,
projector table and diagonal at k, the complex H·X, the complex and the
realified Davidson against the dense spectrum of H_k, the complex Davidson's
iterations against the oracle's, k = 0 against Gamma, the band density and
nonlocal energies, and a Monkhorst-Pack k SCF with Fermi-Dirac smearing."""
import numpy as np
import pytest

from conftest import fcc_rows

pytestmark = pytest.mark.gpu

ROWS = fcc_rows(10.26)
ATOMS = [[0, 0, 0], [0.25] * 3]
K_GENERIC = (0.2, 0.1, -0.3)


def _pair(ecut, k):
    from oracle import physics as P
    from oracle import kpoint as K
    from baandi_amd.hamiltonian import GammaSystem
    system = GammaSystem(ROWS, ATOMS, 4.0, ecut)
    ks = K.build_k_system(ROWS, ATOMS, P.synthetic_pseudo(4.0, 1.0, 801), ecut, k)
    return system, ks


def _potential(ks, seed=0):
    rng = np.random.default_rng(seed)
    return ks.ionic_grid + 0.05 * rng.standard_normal(ks.dims)


def test_k_basis_tables_match_oracle():
    from baandi_amd.kpoint import KPointHamiltonian
    for k in [K_GENERIC, (0.5, -0.25, 0.75), (0.0, 0.0, 0.0)]:
        system, ks = _pair(5.0, k)
        op = KPointHamiltonian(system, k, capacity=4)
        print(f"[k basis] k={k} wrapped={op.k_frac} nk={op.nk} oracle={len(ks.kinetic)}", flush=True)
        assert op.k_frac == ks.k_frac
        assert np.array_equal(op.get("miller"), ks.basis.miller)
        assert np.array_equal(op.get("fft_index"), ks.fft_index)
        assert np.abs(op.get("q2") - ks.basis.q2).max() <= 1e-14 * ks.basis.q2.max()
        kin = op.get("kinetic")
        assert np.array_equal(kin[0::2], kin[1::2]) and np.abs(kin[0::2] - ks.kinetic).max() < 1e-13
        p = op.projectors_complex()
        assert np.abs(p - ks.projectors).max() <= 1e-13 * np.abs(ks.projectors).max()
        from oracle.kpoint import static_diagonal
        d = op.get("diagonal_static")
        assert np.array_equal(d[0::2], d[1::2]) and np.abs(d[0::2] - static_diagonal(ks)).max() < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_k_apply_matches_oracle(precision):
    import torch
    from baandi_amd import lib as L
    from baandi_amd.kpoint import KPointHamiltonian
    from oracle.kpoint import apply_k
    system, ks = _pair(5.0, K_GENERIC)
    prec = L.BD_FP64 if precision == "fp64" else L.BD_MIXED
    op = KPointHamiltonian(system, K_GENERIC, capacity=7, precision=prec)
    pot = _potential(ks)
    op.set_owned_potential(pot.reshape(-1))
    rng = np.random.default_rng(3)
    for m in (1, 4, 7):
        x = rng.standard_normal((m, op.nk)) + 1j * rng.standard_normal((m, op.nk))
        y = op.apply_complex(torch.from_numpy(x).cuda()).cpu().numpy()
        ref = apply_k(ks, x.T, pot).T
        err = np.abs(y - ref).max() / np.abs(ref).max()
        print(f"[k apply] {precision} m={m} nk={op.nk} max relative error {err:.2e}", flush=True)
        if precision == "fp64":
            assert err < 1e-13
        else:
            assert 1e-12 < err < 2e-5  # the Gamma mixed tier's FP32/BF16x3 budget


def _dense_solve(ecut, k, bands, precond, threshold=1e-12, seed=5, realified=False):
    import torch
    from baandi_amd import ref_davidson, lib as L
    from baandi_amd.kpoint import KPointHamiltonian, solve
    from oracle.kpoint import dense_k
    system, ks = _pair(ecut, k)
    op = KPointHamiltonian(system, k, capacity=2 * bands)
    pot = _potential(ks, seed)
    op.set_owned_potential(pot.reshape(-1))
    h = dense_k(ks, pot)
    w = np.linalg.eigvalsh(h)
    rng = np.random.default_rng(seed)
    x0 = rng.standard_normal((bands, op.nk)) + 1j * rng.standard_normal((bands, op.nk))
    x0 /= np.linalg.norm(x0, axis=1, keepdims=True)
    diag = torch.from_numpy(op.get("diagonal_static") + pot.mean()).cuda()
    sol = solve(op, torch.from_numpy(x0).cuda(), diag, ref_davidson.Options(
        bands=bands, threshold=threshold, occupied=bands, max_iterations=300, preconditioner=precond),
        realified=realified)
    return op, h, w, sol


@pytest.mark.parametrize("realified", [False, True])
@pytest.mark.parametrize("k", [K_GENERIC, (0.5, 0.5, 0.5)])
def test_k_davidson_matches_dense_spectrum(k, realified):
    from baandi_amd import lib as L
    op, h, w, sol = _dense_solve(6.0, k, 8, L.BD_PRECOND_TPA, realified=realified)
    err = np.abs(sol.eigenvalues - w[:8]).max()
    split = ""
    if realified:
        levels = sol.realified.eigenvalues
        split = f", pair split {np.abs(levels[0::2] - levels[1::2]).max():.1e}"
    print(f"[k davidson] k={k} realified={realified} nk={op.nk} eigenvalue error {err:.2e} Ha{split}, "
          f"record={sol.record}", flush=True)
    assert sol.converged.all()
    assert err < 1e-10
    x = sol.eigenvectors.cpu().numpy()
    assert np.abs(x.conj() @ x.T - np.eye(8)).max() < 1e-10
    r = h @ x.T - x.T * sol.eigenvalues[None, :]
    assert np.linalg.norm(r, axis=0).max() < 1e-5


def _k_oracle_pair(ecut, k, bands, precond, its, seed):
    """The GPU complex Davidson and the oracle's, `its` iterations from the
    same complex start at the same potential."""
    import torch
    from baandi_amd import ref_davidson, lib as L
    from baandi_amd.kpoint import KPointHamiltonian, solve
    from oracle import davidson as OD
    from oracle.kpoint import apply_k, static_diagonal
    system, ks = _pair(ecut, k)
    op = KPointHamiltonian(system, k, capacity=bands)
    pot = _potential(ks, seed)
    op.set_owned_potential(pot.reshape(-1))
    rng = np.random.default_rng(seed)
    x0 = rng.standard_normal((bands, op.nk)) + 1j * rng.standard_normal((bands, op.nk))
    x0 /= np.linalg.norm(x0, axis=1, keepdims=True)
    code = L.BD_PRECOND_QE if precond == "qe" else L.BD_PRECOND_TPA
    diag = torch.from_numpy(op.get("diagonal_static") + pot.mean()).cuda()
    sol = solve(op, torch.from_numpy(x0).cuda(), diag, ref_davidson.Options(
        bands=bands, threshold=1e-13, occupied=bands, max_iterations=its, preconditioner=code))
    ref = OD.Davidson(lambda x: apply_k(ks, x, pot), static_diagonal(ks) + pot.mean(), bands, 1e-13, bands,
                      max_iterations=its, preconditioner=precond, kinetic=ks.kinetic).solve(x0.T.copy())
    return op, ks, pot, sol, ref


@pytest.mark.parametrize("precond", ["qe", "tpa"])
def test_k_complex_davidson_follows_oracle(precond):
    """Iteration by iteration the GPU complex Davidson is the oracle's: the
    same H columns, restarts and ideal resets, the same Ritz values to
    rounding (capacity 2 bands, so it contracts from the second iteration)."""
    for its in (1, 2, 3, 5, 8):
        op, ks, pot, sol, ref = _k_oracle_pair(5.0, K_GENERIC, 8, precond, its, 7)
        err = np.abs(sol.eigenvalues - ref["eigenvalues"]).max()
        rec = sol.record
        print(f"[k complex vs oracle] {precond} its={its}: H columns {rec['h_columns']}/{ref['h_columns']}, "
              f"restarts {rec['restarts']}/{ref['restarts']}, ideal {rec['ideal_resets']}/{ref['ideal_resets']}, "
              f"max |dE| {err:.2e}", flush=True)
        assert rec["iterations"] == ref["iterations"] == its
        assert rec["h_columns"] == ref["h_columns"]
        assert rec["restarts"] == ref["restarts"] and rec["ideal_resets"] == ref["ideal_resets"]
        assert err < 1e-10


def test_k_complex_davidson_rank_repair():
    """More bands than half the plane waves (nk = 25, 15 bands): the capacity
    is the dimension, the first corrections do not fit and the rank repair
    runs, as in the oracle; the solve then reaches the dense spectrum."""
    import torch
    from baandi_amd import ref_davidson, lib as L
    from baandi_amd.kpoint import solve
    from oracle.kpoint import dense_k
    k = (0.25, -0.125, 0.375)
    op, ks, pot, sol, ref = _k_oracle_pair(1.5, k, 15, "qe", 1, 2)
    assert op.nk == 25
    print(f"[k complex repair] first iteration: repairs {sol.record['repairs']}/{ref['repairs']}, "
          f"H columns {sol.record['h_columns']}/{ref['h_columns']}", flush=True)
    assert sol.record["repairs"] == ref["repairs"] >= 1
    assert sol.record["h_columns"] == ref["h_columns"]
    assert np.abs(sol.eigenvalues - ref["eigenvalues"]).max() < 1e-10
    w = np.linalg.eigvalsh(dense_k(ks, pot))
    x0 = np.random.default_rng(2).standard_normal((15, 25)) + 0j
    x0 /= np.linalg.norm(x0, axis=1, keepdims=True)
    diag = torch.from_numpy(op.get("diagonal_static") + pot.mean()).cuda()
    s2 = solve(op, torch.from_numpy(x0).cuda(), diag, ref_davidson.Options(
        bands=15, threshold=1e-12, occupied=15, max_iterations=100, preconditioner=L.BD_PRECOND_QE))
    print(f"[k complex repair] converged: record={s2.record}", flush=True)
    assert s2.converged.all()
    assert np.abs(s2.eigenvalues - w[:15]).max() < 1e-10


def test_k_zero_matches_gamma():
    """At k = 0 the complex spectrum is the Gamma (ORC) one."""
    import torch
    from baandi_amd import ref_davidson, lib as L
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    op, h, w, sol = _dense_solve(6.0, (0.0, 0.0, 0.0), 8, L.BD_PRECOND_TPA)
    system = op.system
    g = DistributedRealHamiltonian(system, capacity=8)
    from oracle import physics as P
    o = P.build_gamma_system(ROWS, ATOMS, P.synthetic_pseudo(4.0, 1.0, 801), 6.0)
    pot = _potential(o, 5)
    g.set_owned_potential(pot.reshape(-1))
    x0 = np.random.default_rng(9).standard_normal((8, system.n))
    x0 /= np.linalg.norm(x0, axis=1, keepdims=True)
    diag = torch.from_numpy(system.get("diagonal_static") + pot.mean()).cuda()
    gs = ref_davidson.solve(g, torch.from_numpy(x0).cuda(), diag, ref_davidson.Options(
        bands=8, threshold=1e-12, occupied=8, max_iterations=300, preconditioner=L.BD_PRECOND_TPA))
    err = np.abs(gs.eigenvalues - sol.eigenvalues).max()
    print(f"[k=0 vs gamma] n={system.n} nk={op.nk} max difference {err:.2e} Ha", flush=True)
    assert op.nk == system.n and err < 1e-10


def test_k_density_and_nonlocal_match_oracle():
    import torch
    from oracle.kpoint import band_density, nonlocal_apply
    from baandi_amd.kpoint import KPointHamiltonian, as_real
    system, ks = _pair(5.0, K_GENERIC)
    op = KPointHamiltonian(system, K_GENERIC, capacity=6)
    op.set_owned_potential(_potential(ks).reshape(-1))
    rng = np.random.default_rng(11)
    x = rng.standard_normal((5, op.nk)) + 1j * rng.standard_normal((5, op.nk))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    occ = np.array([2.0, 1.7, 0.9, 0.3, 0.0])
    xr = as_real(torch.from_numpy(x).cuda()).contiguous()
    rho = torch.empty(int(np.prod(ks.dims)), dtype=torch.float64, device="cuda")
    op.band_density_sum(xr, 0.25 * occ, rho)
    ref = band_density(ks, x.T, occ, 0.25)
    err = np.abs(rho.cpu().numpy().reshape(ks.dims) - ref).max() / np.abs(ref).max()
    e_nl = op.nonlocal_band_energies(xr)
    ref_nl = np.real(np.einsum("gj,gj->j", x.T.conj(), nonlocal_apply(ks, x.T)))
    nerr = np.abs(e_nl - ref_nl).max()
    print(f"[k density] relative error {err:.2e}; nonlocal energy error {nerr:.2e}", flush=True)
    assert err < 1e-13 and nerr < 1e-13


@pytest.mark.parametrize("columns,lockstep", [("complex", True), ("complex", False), ("realified", False)])
def test_k_scf_matches_oracle(kat1, columns, lockstep):
    """KAT-1 cell at 5 Ha on a 2x2x2 Monkhorst-Pack mesh (4 points after time
    reversal), 8 bands, Fermi-Dirac 0.01 Ha: the Mermin free energy of the
    oracle's dense-solve k SCF, with the complex and the realified Davidson."""
    from oracle import physics as P
    from oracle import kpoint as OK
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.kpoint import KScfOptions, run_scf_k, monkhorst_pack
    rows = fcc_rows(kat1["cell_a_bohr"])
    kp = monkhorst_pack((2, 2, 2))
    assert [p for p, _ in kp] == [p for p, _ in OK.monkhorst_pack((2, 2, 2))]
    system = GammaSystem(rows, ATOMS, 4.0, kat1["ecut_ha"])
    r = run_scf_k(system, kp, KScfOptions(bands=8, theta=0.01, threshold_ry=1e-11, columns=columns,
                                          lockstep=lockstep),
                  log=lambda s: print(s, flush=True))
    ref = OK.scf_k(rows, ATOMS, P.synthetic_pseudo(4.0, 1.0, 801), kat1["ecut_ha"], (2, 2, 2), 8, 0.01,
                   threshold_ry=1e-11)
    de = abs(r["energy"]["total"] - ref["energy"]["total"])
    deig = np.abs(r["eigenvalues"] - ref["eigenvalues"]).max()
    print(f"[k scf] {columns} lockstep={lockstep} E={r['energy']['total']:.12f} oracle={ref['energy']['total']:.12f} dE={de:.2e} "
          f"TS={r['energy']['entropy_term']:.3e} mu={r['mu']:.8f}/{ref['mu']:.8f} eig err {deig:.2e} "
          f"maps={r['iterations']} wall={r['wall_s']:.2f}s eigen retries={r['totals']['eigen_retries']}",
          flush=True)
    # The two SCFs stop at different densities within dr2 < 1e-11 Ry: the
    # terms and eigenvalues differ to first order in that residual, the
    # stationary free energy only to second order.
    for key in ("kinetic", "nonlocal", "local", "hartree", "xc", "entropy_term"):
        assert abs(r["energy"][key] - ref["energy"][key]) < 1e-5, key
    assert de < 1e-8 * len(ATOMS)
    assert deig < 1e-6


def _al_ops(mesh=(2, 2, 2), betas=True, capacity=8, precision=None):
    from baandi_amd import workloads as W, lib as L
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.kpoint import KPointHamiltonian, monkhorst_pack
    wl = W.al4(mesh=mesh)
    system = GammaSystem(wl["rows"], wl["atoms"], wl["z"], wl["ecut_ha"],
                         betas=wl["betas"] if betas else (), dij_ry=wl["dij_ry"] if betas else ())
    kp = monkhorst_pack(mesh)
    prec = L.BD_FP64 if precision is None else precision
    ops = [KPointHamiltonian(system, k, capacity=capacity, precision=prec) for k, _ in kp]
    rng = np.random.default_rng(1)
    pot = 0.3 * rng.standard_normal(system.dims) - 0.4
    for op in ops:
        op.set_owned_potential(pot.reshape(-1))
    return system, ops


def _random_blocks(ops, m, seed):
    import torch
    rng = np.random.default_rng(seed)
    return [torch.from_numpy(rng.standard_normal((m, op.global_len()))).cuda() for op in ops]


@pytest.mark.parametrize("betas", [False, True])
def test_fused_matches_three_stage(betas):
    """The fused one-workgroup-per-column kernel against the three-stage
    stick/plane path on the Al cell (25^3 grid): bit for bit without
    projectors (same per-line FFT arithmetic); with projectors the nonlocal
    sums differ in order only."""
    import torch
    system, ops = _al_ops(betas=betas)
    for op in ops[:2]:
        shape = op.fused_shape()
        assert shape[0] > 0, "the Al k point fits one workgroup"
        x = _random_blocks([op], 7, 3)[0]
        y1 = torch.empty_like(x)
        op.apply_block(x, y1)
        op.set_fused_shape((0, 0, 0, 0))
        y0 = torch.empty_like(x)
        op.apply_block(x, y0)
        op.set_fused_shape(shape)
        torch.cuda.synchronize()
        d = (y1 - y0).abs().max().item() / y0.abs().max().item()
        print(f"[fused] betas={betas} nk={op.nk} shape={shape} max relative difference {d:.2e}", flush=True)
        if betas:
            assert d < 1e-14
        else:
            assert torch.equal(y1, y0)


def test_batch_apply_bitwise_equals_per_k():
    """All k points of a 2x2x2 mesh (uneven nk) in one batched launch against
    the per-k fused apply: bit for bit."""
    import torch
    from baandi_amd.kpoint import KPointBatch
    system, ops = _al_ops(mesh=(3, 3, 3))
    nks = sorted({op.nk for op in ops})
    assert len(nks) > 1, "uneven basis sizes"
    ms = [1 + (i % 8) for i in range(len(ops))]
    xs = [_random_blocks([op], m, 10 + i)[0] for i, (op, m) in enumerate(zip(ops, ms))]
    ys = [torch.empty_like(x) for x in xs]
    KPointBatch(ops).apply(xs, ys)
    torch.cuda.synchronize()
    same = True
    for op, x, y in zip(ops, xs, ys):
        r = torch.empty_like(x)
        op.apply_block(x, r)
        torch.cuda.synchronize()
        same = same and torch.equal(r, y)
    print(f"[batch] {len(ops)} k points, nk {nks[0]}..{nks[-1]}, columns {sum(ms)}: bitwise {same}", flush=True)
    assert same


def test_autotuned_shape_is_bitwise_neutral():
    import torch
    system, ops = _al_ops(capacity=16)
    op = ops[0]
    x = _random_blocks([op], 16, 5)[0]
    y0 = torch.empty_like(x)
    op.apply_block(x, y0)
    t, xl, zr, xp, us = op.autotune(16, reps=3)
    print(f"[autotune] nk={op.nk} picked threads={t} xlines={xl} zrows={zr} xplanes={xp}: {us:.1f} us per 16 "
          f"columns", flush=True)
    assert op.fused_shape() == (t, xl, zr, xp)
    y1 = torch.empty_like(x)
    op.apply_block(x, y1)
    torch.cuda.synchronize()
    assert torch.equal(y0, y1)


def test_fused_mixed_matches_fp64():
    import torch
    from baandi_amd import lib as L
    _, ops64 = _al_ops()
    _, ops32 = _al_ops(precision=L.BD_MIXED)
    x = _random_blocks([ops64[0]], 5, 8)[0]
    y64, y32 = torch.empty_like(x), torch.empty_like(x)
    ops64[0].apply_block(x, y64)
    ops32[0].apply_block(x, y32)
    torch.cuda.synchronize()
    d = (y64 - y32).abs().max().item() / y64.abs().max().item()
    print(f"[fused mixed] shape={ops32[0].fused_shape()} relative difference to FP64 {d:.2e}", flush=True)
    assert 1e-12 < d < 2e-5


def test_concurrent_solves_equal_sequential():
    """Davidson on several k points from host threads (one stream each) gives
    bit for bit the sequential solves."""
    import torch
    from baandi_amd import ref_davidson, lib as L
    from baandi_amd.kpoint import realified_start
    system, ops = _al_ops(mesh=(2, 2, 2), capacity=16)
    rng = np.random.default_rng(4)
    starts = []
    for op in ops:
        z = rng.standard_normal((8, op.nk)) + 1j * rng.standard_normal((8, op.nk))
        z /= np.linalg.norm(z, axis=1, keepdims=True)
        starts.append(realified_start(torch.from_numpy(z).cuda()))
    diags = [torch.from_numpy(op.get("diagonal_static") - 0.4).cuda() for op in ops]
    opts = ref_davidson.Options(bands=16, threshold=1e-10, occupied=16, max_iterations=200,
                                preconditioner=L.BD_PRECOND_TPA)
    seq = [ref_davidson.solve(op, s, d, opts) for op, s, d in zip(ops, starts, diags)]
    outs = [(torch.empty_like(s), None) for s in starts]
    par = ref_davidson.solve_many(ops, starts, diags, [opts] * len(ops), outs, threads=4)
    torch.cuda.synchronize()
    same = all(np.array_equal(a.eigenvalues, b.eigenvalues) and torch.equal(a.eigenvectors, b.eigenvectors)
               for a, b in zip(seq, par))
    print(f"[many] {len(ops)} k points, threads 4: bitwise equal to sequential {same}; iterations "
          f"{[s.record['iterations'] for s in par]}", flush=True)
    assert same and all(s.converged.all() for s in par)


@pytest.mark.parametrize("kind", ["gamma", "k"])
def test_graph_iteration_bitwise_equals_eager(kind):
    """The block iteration replayed from a captured HIP graph against the same
    iteration launched eagerly: bit for bit over several steps; and it
    converges to the dense spectrum."""
    import torch
    from baandi_amd import ref_davidson, lib as L
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    from baandi_amd.kpoint import KPointHamiltonian, realified_start
    from oracle.kpoint import dense_k
    system, ks = _pair(6.0, K_GENERIC if kind == "k" else (0.0, 0.0, 0.0))
    pot = _potential(ks, 2)
    rng = np.random.default_rng(6)
    if kind == "k":
        op = KPointHamiltonian(system, K_GENERIC, capacity=16)
        z = rng.standard_normal((8, op.nk)) + 1j * rng.standard_normal((8, op.nk))
        x0 = realified_start(torch.from_numpy(z / np.linalg.norm(z, axis=1, keepdims=True)).cuda())
        ref = np.repeat(np.linalg.eigvalsh(dense_k(ks, pot))[:8], 2)
    else:
        from oracle import physics as P
        from oracle.hamiltonian import dense_orc
        op = DistributedRealHamiltonian(system, capacity=16)
        o = P.build_gamma_system(ROWS, ATOMS, P.synthetic_pseudo(4.0, 1.0, 801), 6.0)
        pot = _potential(o, 2)
        x0 = torch.from_numpy(rng.standard_normal((16, system.n))).cuda()
        x0 = x0 / x0.norm(dim=1, keepdim=True)
        ref = np.linalg.eigvalsh(dense_orc(o, pot))[:16]
    op.set_owned_potential(pot.reshape(-1))
    diag = torch.from_numpy((op.get("diagonal_static") if kind == "k" else system.get("diagonal_static"))
                            + pot.mean()).cuda()
    runs = {}
    for graph in (False, True):
        it = ref_davidson.BlockIteration(op, 16, L.BD_PRECOND_TPA, graph=graph)
        it.set_diagonal(diag)
        it.start(x0)
        vals = [it.step() for _ in range(4)]
        x = torch.empty_like(x0)
        it.vectors(x)
        torch.cuda.synchronize()
        runs[graph] = (vals, x)
    same = all(np.array_equal(a, b) for a, b in zip(runs[False][0], runs[True][0])) and \
        torch.equal(runs[False][1], runs[True][1])
    # 16 columns, the lowest 12 converged (the rest a buffer against a
    # degenerate group cut by the block edge).
    vals, x, steps = ref_davidson.solve_block(op, x0, diag, 16, 1e-11, graph=True, converge_bands=12)
    err = np.abs(vals[:12] - ref[:12]).max()
    print(f"[graph] {kind}: replay bitwise equal to eager {same}; solve steps={steps}, eigenvalue error {err:.2e}",
          flush=True)
    assert same
    assert err < 1e-9


def test_plane_autotune_is_bitwise_neutral():
    """Gamma op: the tuned k_plane shape (threads, rows per z chunk) gives the
    same H·X bits as the default, FP64 and mixed."""
    import torch
    from baandi_amd import lib as L
    from baandi_amd.hamiltonian import GammaSystem, DistributedRealHamiltonian
    system = GammaSystem(fcc_rows(10.26), ATOMS, 4.0, 10.0)
    rng = np.random.default_rng(3)
    pot = rng.standard_normal(int(np.prod(system.dims))) * 0.1
    for prec in (L.BD_FP64, L.BD_MIXED):
        op = DistributedRealHamiltonian(system, capacity=12, precision=prec)
        op.set_owned_potential(pot)
        x = torch.from_numpy(rng.standard_normal((12, system.n))).cuda()
        y0, y1 = torch.empty_like(x), torch.empty_like(x)
        op.apply_block(x, y0)
        shape = op.autotune(12, reps=2)
        op.apply_block(x, y1)
        torch.cuda.synchronize()
        print(f"[plane autotune] precision={prec} threads={int(shape[0])} chunk={int(shape[1])} "
              f"{shape[4]:.1f} us per apply", flush=True)
        assert torch.equal(y0, y1)


def test_kblocks_equal_per_k_block_iterations():
    """All k points of the Al 2x2x2 mesh stepped together (per-k work on the
    ops' streams, one fused H·W launch over every column of every k point,
    eager and replayed from the captured graph) give the bits of the per-k
    block iterations: Ritz values of the start and three steps, and the
    final Ritz blocks."""
    import torch
    from baandi_amd import ref_davidson, lib as L
    from baandi_amd.kpoint import KBlocks, realified_start
    system, ops = _al_ops(capacity=16)
    nb = 16
    rng = np.random.default_rng(3)
    starts, diags = [], []
    for op in ops:
        z = rng.standard_normal((8, op.nk)) + 1j * rng.standard_normal((8, op.nk))
        starts.append(realified_start(torch.from_numpy(z / np.linalg.norm(z, axis=1, keepdims=True)).cuda()))
        diags.append(torch.from_numpy(op.get("diagonal_static") - 0.4).cuda())
    ref_vals, ref_x = [], []
    for op, x0, d in zip(ops, starts, diags):
        it = ref_davidson.BlockIteration(op, nb, L.BD_PRECOND_TPA, graph=False)
        it.set_diagonal(d)
        v0 = it.start(x0)
        vals = [v0] + [it.step() for _ in range(3)]
        x = torch.empty_like(x0)
        it.vectors(x)
        torch.cuda.synchronize()
        ref_vals.append(vals)
        ref_x.append(x)
    for graph in (False, True):
        kb = KBlocks(ops, nb, L.BD_PRECOND_TPA)
        kb.set_diagonals(diags)
        vals = [kb.start(starts)] + [kb.step(1, use_graph=graph) for _ in range(3)]
        xs = [torch.empty_like(x0) for x0 in starts]
        kb.vectors(xs)
        torch.cuda.synchronize()
        same = all(np.array_equal(vals[s][k], ref_vals[k][s]) for k in range(len(ops)) for s in range(4)) and \
            all(torch.equal(xs[k], ref_x[k]) for k in range(len(ops)))
        print(f"[kblocks] {len(ops)} k points, graph={graph}: bitwise equal to the per-k iterations {same}",
              flush=True)
        assert same


def test_k_scf_block_solver_matches_davidson_and_oracle(kat1):
    """The k SCF with the block iterations of all k points stepped together
    (one fused H·W launch per step, the step replayed from a HIP graph)
    reaches the fixed point of the Davidson k SCF (Al 2x2x2) and of the
    oracle's dense k SCF (KAT-1 2x2x2)."""
    from oracle import physics as P
    from oracle import kpoint as OK
    from baandi_amd import workloads as W
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.kpoint import KScfOptions, run_scf_k, monkhorst_pack
    wl = W.al4(mesh=(2, 2, 2))
    system = W.build(wl)
    kp = monkhorst_pack((2, 2, 2))
    base = dict(bands=wl["bands"], theta=wl["theta"], threshold_ry=1e-10, preconditioning="thomas_fermi")
    rd = run_scf_k(system, kp, KScfOptions(**base))
    rb = run_scf_k(system, kp, KScfOptions(solver="block", columns="realified", **base))
    de = abs(rd["energy"]["total"] - rb["energy"]["total"])
    deig = np.abs(rd["eigenvalues"] - rb["eigenvalues"]).max()
    print(f"[k scf block] Al 2x2x2: davidson {rd['energy']['total']:.12f} ({rd['wall_s']:.2f}s) block "
          f"{rb['energy']['total']:.12f} ({rb['wall_s']:.2f}s) dE={de:.2e} eig {deig:.2e}", flush=True)
    assert de < 1e-8 * 4
    assert deig < 1e-6
    rows = fcc_rows(kat1["cell_a_bohr"])
    s2 = GammaSystem(rows, ATOMS, 4.0, kat1["ecut_ha"])
    r = run_scf_k(s2, kp, KScfOptions(bands=8, theta=0.01, threshold_ry=1e-11, solver="block",
                                      columns="realified"))
    ref = OK.scf_k(rows, ATOMS, P.synthetic_pseudo(4.0, 1.0, 801), kat1["ecut_ha"], (2, 2, 2), 8, 0.01,
                   threshold_ry=1e-11)
    de = abs(r["energy"]["total"] - ref["energy"]["total"])
    print(f"[k scf block] KAT-1 2x2x2: {r['energy']['total']:.12f} oracle {ref['energy']['total']:.12f} "
          f"dE={de:.2e}", flush=True)
    assert de < 1e-8 * len(ATOMS)
    assert np.abs(r["eigenvalues"] - ref["eigenvalues"]).max() < 1e-6


def _al_starts(ops, bands, seed):
    import torch
    from baandi_amd.kpoint import as_real
    rng = np.random.default_rng(seed)
    starts, diags = [], []
    for op in ops:
        z = rng.standard_normal((bands, op.nk)) + 1j * rng.standard_normal((bands, op.nk))
        z /= np.linalg.norm(z, axis=1, keepdims=True)
        starts.append(as_real(torch.from_numpy(z).cuda()).contiguous())
    return starts


@pytest.mark.parametrize("precond", ["qe", "tpa"])
def test_lockstep_k_davidson_follows_per_k_solves(precond):
    """The Davidsons of all Al 2x2x2 k points in lock step (one launch per
    stage for every k point, the subspace step in one workgroup per k point)
    take each k point along the path of its own complex Davidson: the same
    iterations, H columns, restarts and ideal resets, the same Ritz values to
    rounding, after 1, 2, 3, 5 and 8 iterations and at convergence."""
    import torch
    from baandi_amd import ref_davidson, lib as L
    from baandi_amd.kpoint import KDavidson, solve, as_complex
    bands = 8
    system, ops = _al_ops(capacity=bands)
    code = L.BD_PRECOND_QE if precond == "qe" else L.BD_PRECOND_TPA
    starts = _al_starts(ops, bands, 17)
    diags = [torch.from_numpy(op.get("diagonal_static") - 0.4).cuda() for op in ops]
    kd = KDavidson(ops, bands)
    for its in (1, 2, 3, 5, 8, 200):
        thr = 1e-13 if its < 200 else 1e-11
        o = ref_davidson.Options(bands=bands, threshold=thr, occupied=bands, max_iterations=its,
                                 preconditioner=code)
        outs = [(torch.empty_like(x), None) for x in starts]
        sols = kd.solve(starts, diags, o, outs)
        for k, op in enumerate(ops):
            ref = solve(op, as_complex(starts[k]), diags[k], o)
            err = np.abs(sols[k].eigenvalues - ref.eigenvalues).max()
            a, b = sols[k].record, ref.record
            print(f"[lockstep] {precond} its={its} k={k}: iterations {a['iterations']}/{b['iterations']}, "
                  f"H columns {a['h_columns']}/{b['h_columns']}, restarts {a['restarts']}/{b['restarts']}, "
                  f"ideal {a['ideal_resets']}/{b['ideal_resets']}, repairs {a['repairs']}/{b['repairs']}, "
                  f"max |dE| {err:.2e}", flush=True)
            assert a["iterations"] == b["iterations"] and a["h_columns"] == b["h_columns"]
            assert a["restarts"] == b["restarts"] and a["ideal_resets"] == b["ideal_resets"]
            assert a["repairs"] == b["repairs"]
            assert err < 1e-10
            if its == 200:
                assert sols[k].converged.all()
                x = as_complex(outs[k][0])
                gram = (x.conj() @ x.T).cpu().numpy()
                assert np.abs(gram - np.eye(bands)).max() < 1e-10


def test_lockstep_k_davidson_matches_dense_spectrum():
    """Lock step on the Si2 cell at two k points (one generic, one on the zone
    boundary) converges to the dense spectra of H_k. Near convergence
    borders turn dependent and the rank repair runs inside the lock step
    (record.repairs), as in the single-k solver on the same input."""
    import torch
    from baandi_amd import ref_davidson, lib as L
    from baandi_amd.kpoint import KPointHamiltonian, KDavidson, as_real, as_complex, solve
    from oracle.kpoint import dense_k
    from oracle import physics as P
    from oracle import kpoint as OK
    ks_list, ops = [], []
    system, ks0 = _pair(6.0, K_GENERIC)
    pot = _potential(ks0, 5)
    for k in (K_GENERIC, (0.5, 0.5, 0.5)):
        ks = OK.build_k_system(ROWS, ATOMS, P.synthetic_pseudo(4.0, 1.0, 801), 6.0, k)
        op = KPointHamiltonian(system, k, capacity=8)
        op.set_owned_potential(pot.reshape(-1))
        ks_list.append(ks)
        ops.append(op)
    starts = _al_starts(ops, 8, 3)
    diags = [torch.from_numpy(op.get("diagonal_static") + pot.mean()).cuda() for op in ops]
    kd = KDavidson(ops, 8)
    outs = [(torch.empty_like(x), torch.empty_like(x)) for x in starts]
    sols = kd.solve(starts, diags, ref_davidson.Options(bands=8, threshold=1e-12, occupied=8, max_iterations=300,
                                                        preconditioner=L.BD_PRECOND_TPA, return_hx=True), outs)
    for k, (ks, op) in enumerate(zip(ks_list, ops)):
        h = dense_k(ks, pot)
        w = np.linalg.eigvalsh(h)
        x = as_complex(outs[k][0]).cpu().numpy()
        hx = as_complex(outs[k][1]).cpu().numpy()
        err = np.abs(sols[k].eigenvalues - w[:8]).max()
        single = solve(op, as_complex(starts[k]), diags[k], ref_davidson.Options(
            bands=8, threshold=1e-12, occupied=8, max_iterations=300, preconditioner=L.BD_PRECOND_TPA))
        print(f"[lockstep dense] k={k} nk={op.nk} error {err:.2e} record={sols[k].record} "
              f"single-k iterations {single.record['iterations']} repairs {single.record['repairs']}", flush=True)
        assert sols[k].converged.all() and err < 1e-10
        assert np.abs(x.conj() @ x.T - np.eye(8)).max() < 1e-10
        assert np.abs(hx - (h @ x.T).T).max() < 1e-9
    assert any(s.record["repairs"] >= 1 for s in sols)


def test_k_batch_density_and_potential_equal_per_k_calls():
    """bd_k_band_density_dev (every k point's density side by side, summed in
    k order) gives the bits of the per-k densities added in k order, and
    bd_k_set_potential_dev the H·X of per-op potentials."""
    import torch
    from baandi_amd.kpoint import band_density_all, set_potential_all, as_real
    system, ops = _al_ops(capacity=6)
    rng = np.random.default_rng(4)
    xs, ws = [], []
    for op in ops:
        z = rng.standard_normal((5, op.nk)) + 1j * rng.standard_normal((5, op.nk))
        xs.append(as_real(torch.from_numpy(z / np.linalg.norm(z, axis=1, keepdims=True)).cuda()).contiguous())
        ws.append(rng.uniform(0.0, 0.25, 5))
    n = int(np.prod(system.dims))
    ref = torch.zeros(n, dtype=torch.float64, device="cuda")
    part = torch.empty_like(ref)
    for op, x, w in zip(ops, xs, ws):
        op.band_density_sum(x, w, part)
        ref += part
    got = torch.empty_like(ref)
    band_density_all(ops, xs, ws, got)
    print(f"[k batch density] bitwise equal {torch.equal(got, ref)}, max diff {float((got - ref).abs().max()):.1e}",
          flush=True)
    assert torch.equal(got, ref)
    v = torch.from_numpy(0.2 * rng.standard_normal(n) - 0.3).cuda()
    y0 = []
    for op, x in zip(ops, xs):
        op.set_owned_potential(v)
        y = torch.empty_like(x)
        op.apply_block(x, y)
        y0.append(y)
    set_potential_all(ops, 2.0 * v)
    set_potential_all(ops, v)
    for op, x, y in zip(ops, xs, y0):
        y1 = torch.empty_like(x)
        op.apply_block(x, y1)
        assert torch.equal(y, y1)
