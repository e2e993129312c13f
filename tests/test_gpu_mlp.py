"""Learned preconditioner on MI355X: the MFMA forward pass and the training
kernels against the oracle (oracle/mlp.py, bf16 emulation), training
progress, and Davidson with the network (same fixed point, QE trajectory at
K = 1, QE fallback, gain over QE on a held-out cell)."""
import numpy as np
import pytest

from conftest import fcc_rows

pytestmark = pytest.mark.gpu


def _theta(seed):
    from oracle import mlp as M
    rng = np.random.default_rng(seed)
    th = M.init_params(seed)
    p = M.unpack(th)
    p["b1"] = rng.uniform(-0.3, 0.3, M.HIDDEN)
    p["b2"] = rng.uniform(-0.3, 0.3, M.HIDDEN)
    p["w3"] = rng.uniform(-0.8, 0.8, M.HIDDEN)
    p["b3"] = np.array([0.3])
    return M.pack(p).astype(np.float32).astype(np.float64)


def _synthetic_pairs(n, bands, seed):
    rng = np.random.default_rng(seed)
    g = np.sort(rng.uniform(0.0, 12.0, n)).astype(np.float32)
    R = rng.standard_normal((bands, n)).astype(np.float32)
    Tg = (-R / (g[None, :] + rng.uniform(0.3, 1.5, (bands, 1))) + 0.05 * rng.standard_normal((bands, n)))
    scal = np.stack([rng.uniform(0.4, 2.5, bands), rng.uniform(-0.4, 0.6, bands), rng.uniform(-0.5, 0.1, bands)],
                    axis=1)
    return g, R, Tg.astype(np.float32), scal.astype(np.float32)


def _pairs_dev(g, R, Tg, scal):
    import torch
    from baandi_amd.learned import Pairs
    return Pairs(torch.from_numpy(g).cuda(), torch.from_numpy(R).cuda(), torch.from_numpy(Tg).cuda(),
                 torch.from_numpy(scal).cuda())


def test_forward_matches_oracle():
    import torch
    from oracle import mlp as M
    from baandi_amd.learned import Network
    th = _theta(1)
    net = Network(th)
    g = np.linspace(0.0, 14.0, 1000).astype(np.float32)
    for T, ritz, vbar in [(1.3, 0.2, -0.1), (0.1, -0.4, -0.6), (3.1, 0.9, 0.05)]:
        K = net.weights(torch.from_numpy(g).cuda(), T, ritz, vbar).cpu().numpy()
        ref = M.forward(th, M.features(g.astype(np.float64), np.float32(T), np.float32(ritz), np.float32(vbar)),
                        bf16=True)["K"]
        err = np.abs(K / ref - 1.0).max()
        print(f"[mlp fwd] T={T}: max relative K error {err:.2e}")
        assert err < 2e-3


def test_loss_and_gradient_match_oracle():
    from oracle import mlp as M
    from baandi_amd.learned import Network, loss_and_grad
    th = _theta(2)
    g, R, Tg, scal = _synthetic_pairs(700, 6, 3)
    net = Network(th)
    batch = [0, 2, 3, 5]
    loss, grad = loss_and_grad(net, _pairs_dev(g, R, Tg, scal), batch)
    grad = grad.cpu().numpy().astype(np.float64)
    ref_batch = [(g.astype(np.float64), float(scal[b, 0]), float(scal[b, 1]), float(scal[b, 2]),
                  R[b].astype(np.float64), Tg[b].astype(np.float64)) for b in batch]
    ref_loss, ref_grad = M.loss_and_grad(th, ref_batch, bf16=True)
    gerr = np.linalg.norm(grad - ref_grad) / np.linalg.norm(ref_grad)
    print(f"[mlp grad] loss {loss:.8f} vs oracle {ref_loss:.8f}; gradient relative error {gerr:.2e}")
    assert abs(loss - ref_loss) < 1e-5 * max(abs(ref_loss), 1e-3)
    assert gerr < 5e-3
    # Layer by layer, so that a wrong block cannot hide behind the others.
    for name, shape in M.SIZES:
        a, b = M.unpack(grad)[name], M.unpack(ref_grad)[name]
        if np.linalg.norm(b) > 1e-8:
            assert np.linalg.norm(a - b) / np.linalg.norm(b) < 2e-2, name


def test_training_reduces_loss_below_tpa():
    from baandi_amd.learned import Network, train, tpa_loss, loss_and_grad
    g, R, Tg, scal = _synthetic_pairs(2000, 48, 4)
    pairs = _pairs_dev(g, R, Tg, scal)
    net = Network.initial(0)
    everything = list(range(pairs.bands))
    l0, _ = loss_and_grad(net, pairs, everything)
    losses = train(net, pairs, steps=300, batch=12, lr=3e-3, seed=1)
    l1, _ = loss_and_grad(net, pairs, everything)
    lt = tpa_loss(pairs)
    print(f"[mlp train] loss {l0:.4f} -> {l1:.4f} (TPA {lt:.4f}); last batch losses {losses[-3:]}")
    assert np.all(np.isfinite(losses))
    assert l1 < 0.5 * l0 and l1 < lt


def _si2():
    import torch
    from baandi_amd.hamiltonian import GammaSystem, DistributedRealHamiltonian
    from oracle import physics as P
    from oracle.hamiltonian import dense_orc
    rows, atoms = fcc_rows(10.26), [[0, 0, 0], [0.25] * 3]
    s = GammaSystem(rows, atoms, 4.0, 10.0)
    o = P.build_gamma_system(rows, atoms, P.synthetic_pseudo(4.0, 1.0, 801), 10.0)
    pot = o.ionic_grid
    op = DistributedRealHamiltonian(s, capacity=8)
    op.set_owned_potential(pot.reshape(-1))
    w = np.linalg.eigvalsh(dense_orc(o, pot))
    gen = np.random.default_rng(1).standard_normal((8, s.n))
    x0 = torch.from_numpy(gen / np.linalg.norm(gen, axis=1, keepdims=True)).cuda()
    diag = torch.from_numpy(s.get("diagonal_static") + pot.mean()).cuda()
    return op, x0, diag, w, float(pot.mean())


def test_davidson_with_network_keeps_fixed_point():
    from baandi_amd import ref_davidson, lib as L
    from baandi_amd.learned import Network
    op, x0, diag, w, vbar = _si2()
    net = Network(_theta(5))
    sol = ref_davidson.solve(op, x0, diag, ref_davidson.Options(
        bands=8, threshold=1e-12, occupied=4, max_iterations=200, preconditioner=L.BD_PRECOND_MLP, mlp=net,
        mlp_vbar=vbar))
    err = np.abs(sol.eigenvalues[:4] - w[:4]).max()
    print(f"[mlp davidson] occupied error {err:.2e} Ha, record={sol.record}")
    assert sol.converged[:4].all() and err < 1e-10


def _unit_theta():
    """Parameters whose network output is exactly 0 (K = 1): the QE correction."""
    from oracle import mlp as M
    p = M.unpack(_theta(3))
    p["w3"] = np.zeros(M.HIDDEN)
    p["b3"] = np.zeros(1)
    return M.pack(p)


@pytest.mark.parametrize("iterations", [1, 2, 3])
def test_unit_network_follows_the_qe_trajectory(iterations):
    """K = 1 makes the network's correction -r / d_QE, the QE correction up to
    its sign and the order of one sum, so the Ritz values after 1-3 iterations
    match the QE solve's to rounding; a wrong diagonal, Ritz value or
    denominator in the network's correction moves them by far more."""
    from baandi_amd import ref_davidson, lib as L
    from baandi_amd.learned import Network
    op, x0, diag, w, vbar = _si2()
    net = Network(_unit_theta())
    common = dict(bands=8, threshold=1e-14, occupied=8, max_iterations=iterations, first_step=True)
    mlp = ref_davidson.solve(op, x0, diag, ref_davidson.Options(preconditioner=L.BD_PRECOND_MLP, mlp=net,
                                                                mlp_vbar=vbar, **common))
    qe = ref_davidson.solve(op, x0, diag, ref_davidson.Options(preconditioner=L.BD_PRECOND_QE, **common))
    tpa = ref_davidson.solve(op, x0, diag, ref_davidson.Options(preconditioner=L.BD_PRECOND_TPA, **common))
    d = np.abs(mlp.eigenvalues - qe.eigenvalues).max()
    d_tpa = np.abs(tpa.eigenvalues - qe.eigenvalues).max()
    print(f"[mlp unit] {iterations} iterations: |mlp - qe| {d:.2e} Ha, |tpa - qe| {d_tpa:.2e} Ha")
    assert mlp.record["mlp_fallback"] == -1
    assert d < 1e-11
    assert d_tpa > 1e3 * max(d, 1e-14)


def test_davidson_falls_back_to_qe_on_nonfinite_network():
    from baandi_amd import ref_davidson, lib as L
    from baandi_amd.learned import Network
    op, x0, diag, w, vbar = _si2()
    th = _theta(6)
    th[-1] = np.nan  # b3: every K is NaN
    net = Network(th)
    sol = ref_davidson.solve(op, x0, diag, ref_davidson.Options(
        bands=8, threshold=1e-12, occupied=4, max_iterations=200, preconditioner=L.BD_PRECOND_MLP, mlp=net,
        mlp_vbar=vbar))
    qe = ref_davidson.solve(op, x0, diag, ref_davidson.Options(
        bands=8, threshold=1e-12, occupied=4, max_iterations=200, preconditioner=L.BD_PRECOND_QE))
    print(f"[mlp fallback] record={sol.record}")
    assert sol.record["mlp_fallback"] == 0
    # From the first step on this is the QE solve, bit for bit.
    assert np.array_equal(sol.eigenvalues, qe.eigenvalues)
    assert sol.record["h_columns"] == qe.record["h_columns"]
    assert np.abs(sol.eigenvalues[:4] - w[:4]).max() < 1e-10


def test_trained_network_beats_qe_on_held_out_cell():
    """The committed network (trained on displaced Si64 cells, seeds 101-132)
    on a held-out displaced Si64 cell (seed 2024) with a partly screened
    potential: fewer H columns than the QE correction it scales, from the
    same start, and the same converged eigenvalues (the fixed point does not
    move)."""
    import os
    import torch
    from baandi_amd import ref_davidson, workloads, lib as L
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    from baandi_amd.learned import Network
    from baandi_amd.scf import run_scf, ScfOptions
    from baandi_amd.train_preconditioner import displaced, DEFAULT_OUT
    assert os.path.exists(DEFAULT_OUT)
    net = Network.load(DEFAULT_OUT)
    wl = displaced(workloads.si64(), 0.05, 2024)
    system = workloads.build(wl)
    potential = run_scf(system, ScfOptions(bands=wl["bands"], threshold_ry=1e-3, max_iterations=12,
                                           allow_unconverged=True))["potential"]
    op = DistributedRealHamiltonian(system, capacity=wl["bands"])
    op.set_owned_potential(torch.from_numpy(potential.reshape(-1)).cuda())
    vbar = float(potential.mean())
    diag = torch.from_numpy(system.get("diagonal_static") + vbar).cuda()
    g = torch.Generator(device="cpu").manual_seed(7)
    x0 = torch.randn(wl["bands"], system.n, generator=g, dtype=torch.float64)
    x0 = (x0 / x0.norm(dim=1, keepdim=True)).cuda()
    occ = int(round(system.electrons)) // 2
    out = {}
    for name, pre in [("tpa", L.BD_PRECOND_TPA), ("mlp", L.BD_PRECOND_MLP), ("qe", L.BD_PRECOND_QE)]:
        sol = ref_davidson.solve(op, x0, diag, ref_davidson.Options(
            bands=wl["bands"], threshold=1e-10, occupied=occ, max_iterations=600, preconditioner=pre,
            mlp=net if pre == L.BD_PRECOND_MLP else None, mlp_vbar=vbar))
        out[name] = sol
        print(f"[mlp gain] {name}: iterations {sol.record['iterations']}, H columns {sol.record['h_columns']}, "
              f"fallback {sol.record['mlp_fallback']}, converged {bool(sol.converged.all())}")
    qe, mlp = out["qe"], out["mlp"]
    assert qe.converged[:occ].all() and mlp.converged[:occ].all()
    assert np.abs(qe.eigenvalues[:occ] - mlp.eigenvalues[:occ]).max() < 1e-9
    gain = 1.0 - mlp.record["h_columns"] / qe.record["h_columns"]
    print(f"[mlp gain] H columns saved against QE: {100 * gain:.1f}%, "
          f"iterations {mlp.record['iterations']} against {qe.record['iterations']}")
    assert gain > 0.0
