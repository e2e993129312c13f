"""Mixed tier on MI355X: the MFMA operand layout, BF16x3 split GEMMs against
FP64, and the FP32-FFT / BF16x3 H·X and Davidson against the FP64 oracle.
Each test states the error it measured."""
import numpy as np
import pytest

from conftest import fcc_rows

pytestmark = pytest.mark.gpu


def bf16_bits(x):
    """Round-to-nearest-even float32 -> bf16 bit patterns (test helper)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def test_mfma_operand_layout():
    """A 16x16x16 bf16 MFMA on small integers (exact in bf16 and in FP32)
    equals the host product: pins the lane layout the BF16x3 kernels assume."""
    import torch
    from baandi_amd import lib as L
    rng = np.random.default_rng(0)
    a = rng.integers(-4, 5, (16, 16)).astype(np.float32)
    b = rng.integers(-4, 5, (16, 16)).astype(np.float32)
    ad = torch.from_numpy(bf16_bits(a).astype(np.int16)).cuda()
    bd = torch.from_numpy(bf16_bits(b).astype(np.int16)).cuda()
    d = torch.zeros(256, dtype=torch.float32, device="cuda")
    L.check(L.load().bd_mfma_probe_dev(L.ptr(ad), L.ptr(bd), L.ptr(d)))
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy().reshape(16, 16), a @ b)


def test_mfma_f64_operand_layout():
    """One v_mfma_f64_16x16x4f64 on random doubles against the host product
    (4-term sums: equal to a few ulps): pins the lane layout of the FP64 GEMMs."""
    import torch
    from baandi_amd import lib as L
    rng = np.random.default_rng(2)
    a = rng.standard_normal((16, 4))
    b = rng.standard_normal((4, 16))
    ad, bd = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()  # held until the kernel has run
    d = torch.zeros(256, dtype=torch.float64, device="cuda")
    L.check(L.load().bd_mfma_f64_probe_dev(L.ptr(ad), L.ptr(bd), L.ptr(d)))
    torch.cuda.synchronize()
    assert np.abs(d.cpu().numpy().reshape(16, 16) - a @ b).max() < 1e-14


# Products that leave most CUs idle with 64-wide tiles run on 32 x 32 tiles
# (NN 300 x 47, 320 x 320, 33 x 65, 64 x 16 x 16; TN with one K chunk: 65 x
# 64 x 17, 320 x 160 x 320, 33 x 7 x 100); 1 x 1 x 1 (K below a slice), the
# split-K TN (70 x 33 x 5000) and the tall products keep the 64-wide kernels.
@pytest.mark.parametrize("trans_a,M,N,K", [(False, 300, 47, 130), (True, 70, 33, 5000), (False, 1, 1, 1),
                                           (True, 65, 64, 17), (False, 320, 320, 320), (False, 33, 65, 17),
                                           (False, 64, 16, 16), (False, 4000, 160, 300), (True, 320, 160, 320),
                                           (True, 33, 7, 100)])
def test_fp64_gemm_matches_numpy(trans_a, M, N, K):
    """The FP64 MFMA GEMMs (NN, and TN with split-K) against NumPy: 1e-14 of
    sum |a||b| (edge tiles, beta = 0.5 included)."""
    import torch
    from baandi_amd import lib as L
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((K, M) if trans_a else (M, K))
    B = rng.standard_normal((K, N))
    C0 = rng.standard_normal((M, N))
    ref = 1.5 * ((A.T if trans_a else A) @ B) + 0.5 * C0
    bound = 1e-14 * (np.abs(A.T if trans_a else A) @ np.abs(B)).max() * 1.5 + 1e-14
    Ad = torch.from_numpy(np.asfortranarray(A).ravel(order="F").copy()).cuda()
    Bd = torch.from_numpy(np.asfortranarray(B).ravel(order="F").copy()).cuda()
    Cd = torch.from_numpy(np.asfortranarray(C0).ravel(order="F").copy()).cuda()
    lib = L.load()
    work = torch.empty(max(lib.bd_gemm_workspace(0, int(trans_a), M, N, K), 1), dtype=torch.float64, device="cuda")
    L.check(lib.bd_gemm_dev(0, int(trans_a), M, N, K, 1.5, L.ptr(Ad), A.shape[0], L.ptr(Bd), K, 0.5, L.ptr(Cd), M,
                            L.ptr(work)))
    torch.cuda.synchronize()
    out = Cd.cpu().numpy().reshape(N, M).T
    err = np.abs(out - ref).max()
    print(f"[fp64 gemm] trans_a={trans_a} {M}x{N}x{K}: max err {err:.2e} (bound {bound:.2e})")
    assert err < bound


@pytest.mark.parametrize("trans_a,M,N,K", [(False, 23847, 160, 320), (True, 320, 160, 23847),
                                           (True, 37, 5, 1000), (False, 100, 7, 33)])
def test_bf16x3_gemm_error_bound(trans_a, M, N, K):
    import torch
    from baandi_amd import lib as L
    rng = np.random.default_rng(1)
    A = rng.standard_normal((K, M) if trans_a else (M, K))
    B = rng.standard_normal((K, N))
    ref = (A.T if trans_a else A) @ B
    lib = L.load()
    out = {}
    for mode in (0, 1):
        Ad = torch.from_numpy(np.asfortranarray(A).ravel(order="F")).cuda()
        Bd = torch.from_numpy(np.asfortranarray(B).ravel(order="F")).cuda()
        Cd = torch.zeros(M * N, dtype=torch.float64, device="cuda")
        wsz = lib.bd_gemm_workspace(mode, int(trans_a), M, N, K)
        W = torch.empty(max(wsz, 1), dtype=torch.float64, device="cuda")
        lda = A.shape[0]
        L.check(lib.bd_gemm_dev(mode, int(trans_a), M, N, K, 1.0, L.ptr(Ad), lda, L.ptr(Bd), K, 0.0, L.ptr(Cd), M,
                                L.ptr(W)))
        out[mode] = Cd.cpu().numpy().reshape(N, M).T
    # FP64 tiles: rounding only.
    assert np.abs(out[0] - ref).max() < 1e-12 * np.abs(ref).max()
    # BF16x3: per-product relative error ~2^-16, FP32 accumulation; the bound
    # scales with the sum of |a_ik b_kj| (random signs cancel in the sum).
    scale = (np.abs(A.T if trans_a else A) @ np.abs(B)).max()
    err = np.abs(out[1] - ref).max()
    print(f"[bf16x3] trans_a={trans_a} {M}x{N}x{K}: max err {err:.3e}, relative to sum|ab| {err / scale:.3e}")
    assert err < 2.0 ** -15 * scale


@pytest.mark.parametrize("tier", ["mixed", "fp32_hx"])
def test_mixed_hx_against_fp64_oracle(tier):
    """H·X of the reduced tiers against the FP64 oracle: MIXED (FP32 FFT and
    BF16x3 projector products) and FP32_HX (FP32 FFT, FP64 products)."""
    import torch
    from baandi_amd.hamiltonian import GammaSystem, DistributedRealHamiltonian
    from baandi_amd import lib as L
    from oracle import physics as P
    from oracle.hamiltonian import apply_orc
    rows, atoms = fcc_rows(10.26), [[0, 0, 0], [0.25] * 3]
    s = GammaSystem(rows, atoms, 4.0, 10.0)
    o = P.build_gamma_system(rows, atoms, P.synthetic_pseudo(4.0, 1.0, 801), 10.0)
    pot = o.ionic_grid
    op = DistributedRealHamiltonian(s, capacity=9, precision=L.BD_MIXED if tier == "mixed" else L.BD_FP32_HX)
    op.set_owned_potential(pot.reshape(-1))
    x = np.random.default_rng(3).standard_normal((9, s.n))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    op.apply_block(xd, yd)
    torch.cuda.synchronize()
    ref = apply_orc(o, x.T, pot).T
    err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
    print(f"[{tier} H·X] max relative error vs FP64 oracle: {err:.3e}")
    # really reduced precision, and within the budget: FP32 FFT ~1e-7, BF16x3 ~2^-16
    assert 1e-12 < err < (2e-5 if tier == "mixed" else 2e-6)


@pytest.mark.parametrize("tier", ["mixed", "fp32_hx"])
@pytest.mark.parametrize("precond", [0, 1])
def test_mixed_davidson_eigenvalues_vs_fp64(precond, tier):
    """Reduced-tier Davidson (FP32 FFT H·X; BF16x3 Grams/rotations in MIXED,
    FP64 in FP32_HX; FP64 factor and eigensolve) against the dense FP64
    eigenvalues at a loose threshold."""
    import torch
    from baandi_amd.hamiltonian import GammaSystem, DistributedRealHamiltonian
    from baandi_amd import ref_davidson, lib as L
    from oracle import physics as P
    from oracle.hamiltonian import dense_orc
    rows, atoms = fcc_rows(10.26), [[0, 0, 0], [0.25] * 3]
    s = GammaSystem(rows, atoms, 4.0, 10.0)
    o = P.build_gamma_system(rows, atoms, P.synthetic_pseudo(4.0, 1.0, 801), 10.0)
    pot = o.ionic_grid
    w = np.linalg.eigvalsh(dense_orc(o, pot))
    op = DistributedRealHamiltonian(s, capacity=8, precision=L.BD_MIXED if tier == "mixed" else L.BD_FP32_HX)
    op.set_owned_potential(pot.reshape(-1))
    g = np.random.default_rng(1).standard_normal((8, s.n))
    x0 = torch.from_numpy(g / np.linalg.norm(g, axis=1, keepdims=True)).cuda()
    diag = torch.from_numpy(s.get("diagonal_static") + pot.mean()).cuda()
    sol = ref_davidson.solve(op, x0, diag, ref_davidson.Options(bands=8, threshold=1e-6, occupied=4,
                                                                max_iterations=100, preconditioner=precond))
    err = np.abs(sol.eigenvalues[:4] - w[:4]).max()
    print(f"[{tier} davidson] occupied eigenvalue error vs FP64 dense: {err:.3e} Ha, record={sol.record}")
    assert sol.converged[:4].all()
    # BF16x3 Grams carry ~2^-16 relative noise: the mixed tier's eigenvalue
    # floor is ~2e-5 Ha, which is why the SCF finishes in FP64. FP32 H·X
    # alone leaves ~1e-7 relative in H·X: a floor near 1e-6 Ha.
    assert err < (5e-5 if tier == "mixed" else 2e-6)


@pytest.mark.parametrize("mode", [0, 1])
def test_gemm_workspace_sized_at_capacity_serves_narrow_products(mode):
    """A^T B with a workspace sized for (640, 576) and products of 64..576
    columns, as the operator's projector product P^T X runs them: the partial
    sums stay inside the sized region (a guard band after it keeps its fill)
    and each product matches the FP64 host product."""
    import torch
    from baandi_amd import lib as L
    lib = L.load()
    K, M, cap = 36525, 640, 576
    rng = np.random.default_rng(7)
    A = rng.standard_normal((K, M))
    B = rng.standard_normal((K, cap))
    Ad = torch.from_numpy(np.asfortranarray(A).ravel(order="F")).cuda()
    Bd = torch.from_numpy(np.asfortranarray(B).ravel(order="F")).cuda()
    size = lib.bd_gemm_workspace(mode, 1, M, cap, K)
    guard = 1 << 20
    W = torch.full((size + guard,), 7.25, dtype=torch.float64, device="cuda")
    for n in (64, 96, 128, 288, 576):
        print(f"[gemm workspace] mode={mode} {M}x{n}x{K} sized {size}", flush=True)
        Cd = torch.zeros(M * n, dtype=torch.float64, device="cuda")
        L.check(lib.bd_gemm_dev(mode, 1, M, n, K, 1.0, L.ptr(Ad), K, L.ptr(Bd), K, 0.0, L.ptr(Cd), M, L.ptr(W)))
        torch.cuda.synchronize()
        assert bool((W[size:] == 7.25).all()), n
        ref = A.T @ B[:, :n]
        out = Cd.cpu().numpy().reshape(n, M).T
        err = np.abs(out - ref).max()
        if mode == 0:
            assert err < 1e-12 * np.abs(ref).max()
        else:  # the bf16x3 bound of test_bf16x3_gemm_error_bound
            assert err < 2.0 ** -15 * (np.abs(A.T) @ np.abs(B[:, :n])).max()
