"""Parity at the measured workloads' own sizes, through the C ABI.

* Si64 (2x2x2 diamond, a = 20.52 Bohr, 30 Ry, n = 23,847, 72^3 grid, 320
  projectors): H·X against the oracle for every plane-kernel shape the
  autotuner may pick (256 threads x 16/64 rows per z chunk, 512 threads x
  16/72 rows; radix 9 x 8 passes), and after autotuning; a 160-band Davidson at a fixed potential
  checked a posteriori with the oracle's H (Rayleigh quotients, residuals,
  orthonormality).
* Si128 (4x2x2 cells displaced, 25 Ry, 135 x 72 x 72 grid, 640 projectors):
  H·X against the oracle.
* Al 4-atom fcc cell, 6x6x6 Monkhorst-Pack mesh (108 k points, 25 Ry, 25^3):
  H_k·X against the oracle at the k points of smallest and largest basis and
  two more; one batched launch over all 108 k points bit for bit against the
  per-k applies.

Bars: 1e-12 relative for H·X (FP64 FFT rounding; the reference's dense-vs-FFT
test uses 1e-11), 1e-10 Ha for eigenvalues.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gamma(wl, capacity, seed=7):
    import torch
    from baandi_amd import workloads as W
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    from oracle import physics as P
    system = W.build(wl)
    orc = P.build_gamma_system(wl["rows"], wl["atoms"], P.synthetic_pseudo(4.0, 1.0, 801), wl["ecut_ha"])
    assert tuple(system.dims) == tuple(orc.dims)
    assert np.array_equal(system.get("miller"), orc.basis.miller)
    op = DistributedRealHamiltonian(system, capacity=capacity)
    rng = np.random.default_rng(seed)
    pot = orc.ionic_grid + 0.05 * rng.standard_normal(orc.dims)
    op.set_owned_potential(torch.from_numpy(pot.reshape(-1)).cuda())
    return torch, system, orc, op, pot


def _apply(torch, op, x):
    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    yd = torch.full_like(xd, float("nan"))
    op.apply_block(xd, yd)
    torch.cuda.synchronize()
    return yd


@pytest.fixture(scope="module")
def si64():
    from baandi_amd import workloads as W
    return _gamma(W.si64(), capacity=160)


def test_si64_hx_every_plane_shape(si64):
    """Si64 H·X on 32 columns for each plane-kernel shape, against the oracle;
    all shapes give the same bits."""
    from oracle.hamiltonian import apply_orc
    torch, system, orc, op, pot = si64
    plan = op.plan_info()
    print(f"[si64 hx] n={system.n} dims={system.dims} plan={plan} default shape={op.plane_shape()}", flush=True)
    assert (system.n, plan["sticks"], plan["z_planes"]) == (23847, 997, 35)
    x = np.random.default_rng(11).standard_normal((32, system.n))
    ref = apply_orc(orc, x.T, pot).T
    default = op.plane_shape()
    first = None
    for threads, chunk in ((256, 16), (256, 64), (512, 16), (512, 72)):
        op.set_plane_shape(threads, chunk)
        assert op.plane_shape() == (threads, chunk)
        yd = _apply(torch, op, x)
        err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
        print(f"[si64 hx] threads={threads} chunk={chunk}: max relative error {err:.2e}", flush=True)
        assert err < 1e-12
        if first is None:
            first = yd
        assert torch.equal(first, yd)
    op.set_plane_shape(*default)
    tuned = op.autotune(160, reps=2)
    yd = _apply(torch, op, x)
    print(f"[si64 hx] autotuned shape threads={int(tuned[0])} chunk={int(tuned[1])}", flush=True)
    assert torch.equal(first, yd)


def test_si64_plane_shape_rejects_invalid(si64):
    from baandi_amd import lib as L
    torch, system, orc, op, pot = si64
    # 72 rows of radix-9 butterflies exceed what 256 threads hold in one pass.
    for bad in ((128, 16), (256, 0), (512, 73), (256, 72)):
        with pytest.raises(L.BaandiError):
            op.set_plane_shape(*bad)


def test_si64_davidson_fixed_potential(si64):
    """160 bands (128 occupied) of the Si64 H at a fixed potential from a
    seeded random start. The returned pairs are checked with the oracle's H:
    the Rayleigh quotients x^T H x equal the returned eigenvalues to 1e-10 Ha,
    X is orthonormal to 1e-10, and each residual |Hx - theta x| bounds the
    distance to an eigenvalue of H."""
    from baandi_amd import ref_davidson
    from oracle.hamiltonian import apply_orc
    torch, system, orc, op, pot = si64
    op.set_plane_shape(512, 72)
    bands, occ = 160, 128
    g = np.random.default_rng(1).standard_normal((bands, system.n))
    x0 = torch.from_numpy(g / np.linalg.norm(g, axis=1, keepdims=True)).cuda()
    diag = torch.from_numpy(system.get("diagonal_static") + pot.mean()).cuda()
    opts = ref_davidson.Options(bands=bands, threshold=1e-10, occupied=occ, max_iterations=200, return_hx=True)
    sol = ref_davidson.solve_with_initial_hx(op, x0, diag, opts)
    print(f"[si64 davidson] record={sol.record}, unconverged bands {np.flatnonzero(~sol.converged).tolist()}",
          flush=True)
    # Every occupied band meets 1e-10 Ha. From a random start the highest
    # empty band (in a near-degenerate cluster at the window edge) may still
    # move by more than its 5e-6 Ha rule at the iteration limit; the
    # a posteriori checks below cover every band.
    assert sol.converged[:occ].all()
    x = sol.eigenvectors.cpu().numpy()
    hx_gpu = sol.hx.cpu().numpy()
    hx = apply_orc(orc, x.T, pot).T
    assert np.abs(hx_gpu - hx).max() < 1e-12 * np.abs(hx).max()
    rq = np.einsum("ij,ij->i", x, hx)
    res = np.linalg.norm(hx - sol.eigenvalues[:, None] * x, axis=1)
    orth = np.abs(x @ x.T - np.eye(bands)).max()
    print(f"[si64 davidson] |rq - theta| {np.abs(rq - sol.eigenvalues).max():.2e}, orthonormality {orth:.2e}, "
          f"residual occupied {res[:occ].max():.2e} empty {res[occ:].max():.2e}", flush=True)
    assert np.abs(rq - sol.eigenvalues).max() < 1e-10
    assert orth < 1e-10
    assert np.all(np.diff(sol.eigenvalues) >= -1e-12)
    assert res[:occ].max() < 1e-4
    # Projected back onto the subspace, H is diagonal with the returned values.
    hsub = x @ hx.T
    assert np.abs(hsub - np.diag(sol.eigenvalues)).max() < 1e-9


def test_si128_hx_matches_oracle():
    from baandi_amd import workloads as W
    from oracle.hamiltonian import apply_orc
    torch, system, orc, op, pot = _gamma(W.si128(), capacity=8)
    plan = op.plan_info()
    print(f"[si128 hx] n={system.n} dims={system.dims} nproj={system.nproj} plan={plan}", flush=True)
    assert system.n == 36525 and tuple(system.dims) == (135, 72, 72) and system.nproj == 640
    x = np.random.default_rng(13).standard_normal((8, system.n))
    ref = apply_orc(orc, x.T, pot).T
    yd = _apply(torch, op, x)
    err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
    print(f"[si128 hx] max relative error {err:.2e}", flush=True)
    assert err < 1e-12


@pytest.fixture(scope="module")
def al108():
    from baandi_amd import workloads as W
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.kpoint import KPointHamiltonian, monkhorst_pack
    wl = W.al4()
    system = GammaSystem(wl["rows"], wl["atoms"], wl["z"], wl["ecut_ha"], betas=wl["betas"], dij_ry=wl["dij_ry"])
    kp = monkhorst_pack(wl["mesh"])
    assert len(kp) == 108
    ops = [KPointHamiltonian(system, k, capacity=32) for k, _ in kp]
    rng = np.random.default_rng(2)
    pot = 0.3 * rng.standard_normal(system.dims) - 0.4
    for op in ops:
        op.set_owned_potential(pot.reshape(-1))
    return wl, system, kp, ops, pot


def test_al_hk_matches_oracle(al108):
    """H_k·X (16 complex columns) at four of the 108 k points against the oracle."""
    import torch
    from oracle import physics as P, kpoint as OK
    wl, system, kp, ops, pot = al108
    nks = np.array([op.nk for op in ops])
    print(f"[al hk] 108 k points, nk {nks.min()}..{nks.max()}", flush=True)
    pseudo = P.synthetic_pseudo(wl["z"], 1.0, 801, betas=wl["betas"], dij_ry=wl["dij_ry"])
    picks = sorted({int(nks.argmin()), int(nks.argmax()), 1, 57})
    rng = np.random.default_rng(4)
    for i in picks:
        op = ops[i]
        ks = OK.build_k_system(wl["rows"], wl["atoms"], pseudo, wl["ecut_ha"], kp[i][0])
        assert np.array_equal(op.get("miller"), ks.basis.miller)
        z = rng.standard_normal((16, op.nk)) + 1j * rng.standard_normal((16, op.nk))
        hz = op.apply_complex(torch.from_numpy(z).cuda()).cpu().numpy()
        ref = OK.apply_k(ks, z.T, pot).T
        err = np.abs(hz - ref).max() / np.abs(ref).max()
        print(f"[al hk] k #{i} {op.k_frac} nk={op.nk} shape={op.fused_shape()}: max relative error {err:.2e}",
              flush=True)
        assert err < 1e-12


def test_al_batch_of_108_bitwise_equals_per_k(al108):
    """All 108 k points, 32 realified columns each (the 16 complex bands),
    in one batched launch against the per-k applies: bit for bit."""
    import torch
    from baandi_amd.kpoint import KPointBatch
    wl, system, kp, ops, pot = al108
    rng = np.random.default_rng(5)
    xs = [torch.from_numpy(rng.standard_normal((32, op.global_len()))).cuda() for op in ops]
    ys = [torch.empty_like(x) for x in xs]
    KPointBatch(ops).apply(xs, ys)
    torch.cuda.synchronize()
    same = 0
    for op, x, y in zip(ops, xs, ys):
        r = torch.empty_like(x)
        op.apply_block(x, r)
        torch.cuda.synchronize()
        same += int(torch.equal(r, y))
    print(f"[al batch] {same} of {len(ops)} k points bitwise equal", flush=True)
    assert same == len(ops)
