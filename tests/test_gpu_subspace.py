"""The FP64 subspace step on MI355X, through the C ABI: the GPU eigensolver
(tridiagonalisation, bisection, inverse iteration, cluster Gram-Schmidt,
back-transformation) against LAPACK (numpy.linalg.eigh), and the bordered
inverse Cholesky against the oracle's factor."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _eigen_dev(A, want):
    import torch
    from baandi_amd import lib as L
    lib = L.load()
    m = A.shape[0]
    Ad = torch.from_numpy(np.asfortranarray(A).ravel(order="F").copy()).cuda()
    vals = torch.zeros(want, dtype=torch.float64, device="cuda")
    Y = torch.zeros(m * want, dtype=torch.float64, device="cuda")
    work = torch.zeros(lib.bd_sym_eigen_workspace(m, want), dtype=torch.float64, device="cuda")
    L.check(lib.bd_sym_eigen_dev(m, want, L.ptr(Ad), m, L.ptr(vals), L.ptr(Y), m, L.ptr(work)))
    return vals.cpu().numpy(), Y.cpu().numpy().reshape(want, m).T


def _check_pairs(A, want, label):
    w, Y = _eigen_dev(A, want)
    ref = np.linalg.eigvalsh(A)[:want]
    scale = max(np.abs(A).sum(axis=0).max(), 1e-300)
    verr = np.abs(w - ref).max() / scale
    res = np.abs(A @ Y - Y * w[None, :]).max() / scale
    orth = np.abs(Y.T @ Y - np.eye(want)).max()
    print(f"[eigen] {label} m={A.shape[0]} want={want}: value {verr:.2e}, residual {res:.2e}, orthogonality {orth:.2e}")
    # Backward-stable reduction and bisection: O(m eps ||A||) values.
    assert verr < 1e-13 * max(A.shape[0], 10)
    assert res < 1e-12 * max(A.shape[0], 10)
    assert orth < 1e-12 * max(A.shape[0], 10)


# The blocked tridiagonalisation holds m <= 523 in LDS (523 and 524 straddle
# that limit); the LDS inverse iteration holds 8 vectors per workgroup up to
# m = 476 and 4 up to 910. Above those the one-step and global-memory kernels
# run. The back-transform keeps the vector in registers, 5 rows a lane up to
# m = 320 and 10 up to 640 (320/321 and 640/641 straddle those), and stages it
# in LDS above, past 64 KB of LDS from m = 683.
@pytest.mark.parametrize("m,want", [(1, 1), (2, 2), (3, 3), (5, 2), (64, 64), (160, 160), (320, 160), (321, 100),
                                    (523, 60), (524, 100), (576, 288), (640, 64), (641, 64), (659, 40), (660, 30),
                                    (700, 60), (910, 40), (911, 40), (974, 30), (1025, 30)])
def test_eigen_random_symmetric(m, want):
    rng = np.random.default_rng(m * 7 + want)
    B = rng.standard_normal((m, m))
    _check_pairs(0.5 * (B + B.T), want, "random")


@pytest.mark.parametrize("half", [4, 16, 32, 80])
def test_eigen_realified_hermitian(half):
    """The realified form [[Re A, -Im A], [Im A, Re A]] of a Hermitian A (the
    k-point Davidson's reduced matrix): every eigenvalue exactly twice."""
    rng = np.random.default_rng(half)
    B = rng.standard_normal((half, half)) + 1j * rng.standard_normal((half, half))
    A = 0.5 * (B + B.conj().T)
    R = np.block([[A.real, -A.imag], [A.imag, A.real]])
    _check_pairs(R, half, "realified")
    _check_pairs(R, 2 * half, "realified")


def test_eigen_realified_above_lds_inverse_iteration():
    """m = 912 (above the LDS inverse iteration's 910): the realified
    Hermitian matrix, every eigenvalue exactly twice, through the
    global-memory inverse iteration."""
    rng = np.random.default_rng(912)
    B = rng.standard_normal((456, 456)) + 1j * rng.standard_normal((456, 456))
    A = 0.5 * (B + B.conj().T)
    _check_pairs(np.block([[A.real, -A.imag], [A.imag, A.real]]), 40, "realified")


@pytest.mark.parametrize("m", [12, 96, 240])
def test_eigen_degenerate_and_clustered(m):
    """Exact multiplicities (the triplets of a cubic crystal), a 1e-13 cluster
    and a diagonal block structure (zero off-diagonals after reduction)."""
    rng = np.random.default_rng(m)
    vals = np.sort(rng.uniform(-1.0, 3.0, m))
    vals[3:6] = vals[3]            # triple
    vals[8:10] = vals[8]           # pair
    vals[-4:-1] = vals[-4] + np.array([0.0, 1e-13, 2e-13])
    vals = np.sort(vals)
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    A = (Q * vals[None, :]) @ Q.T
    _check_pairs(0.5 * (A + A.T), m // 2 + 1, "degenerate")
    D = np.diag(vals.copy())
    _check_pairs(D, m // 2, "diagonal")


def test_eigen_davidson_like_spectrum():
    """T H T^T of a Davidson subspace: a low occupied block, a wide tail."""
    rng = np.random.default_rng(5)
    m = 320
    vals = np.concatenate([np.sort(rng.uniform(-0.3, 0.2, 160)), np.sort(rng.uniform(0.2, 40.0, m - 160))])
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    A = (Q * vals[None, :]) @ Q.T
    _check_pairs(0.5 * (A + A.T), 160, "davidson-like")


def _factor_dev(S, T, cap, frm, to):
    import ctypes as C
    import torch
    from baandi_amd import lib as L
    lib = L.load()
    Sd = torch.from_numpy(np.asfortranarray(S).ravel(order="F").copy()).cuda()
    Td = torch.from_numpy(np.asfortranarray(T).ravel(order="F").copy()).cuda()
    acc = C.c_int32(0)
    ratio = C.c_double(0.0)
    L.check(lib.bd_inverse_cholesky_dev(L.ptr(Sd), L.ptr(Td), cap, frm, to, C.byref(acc), C.byref(ratio)))
    return bool(acc.value), ratio.value, Td.cpu().numpy().reshape(cap, cap).T


def test_inverse_cholesky_border_against_oracle():
    from oracle.davidson import Davidson
    rng = np.random.default_rng(2)
    cap, n = 200, 900
    V = rng.standard_normal((n, cap)) * rng.uniform(0.5, 2.0, cap)[None, :]
    S = V.T @ V
    T = np.zeros((cap, cap))
    for frm, to in [(0, 90), (90, 150), (150, 200)]:
        ok, ratio, T = _factor_dev(S, T, cap, frm, to)
        ok_ref, Tref = Davidson._factor(None, S, frm, to, None)
        assert ok and ok_ref
        err = np.abs(T[:to, :to] - Tref).max() / np.abs(Tref).max()
        ident = np.abs(T[:to, :to] @ S[:to, :to] @ T[:to, :to].T - np.eye(to)).max()
        print(f"[factor] {frm}->{to}: T vs oracle {err:.2e}, |T S T^T - I| {ident:.2e}, min ratio {ratio:.3e}")
        assert err < 1e-11 and ident < 1e-11
        assert np.all(T[:to, :to][np.triu_indices(to, 1)] == 0.0)


def test_inverse_cholesky_rejects_dependent_border():
    from oracle.davidson import Davidson
    rng = np.random.default_rng(3)
    cap, n = 64, 300
    V = rng.standard_normal((n, cap))
    V[:, 40] = V[:, 3] + 1e-9 * V[:, 41]   # numerically dependent on column 3
    S = V.T @ V
    T = np.zeros((cap, cap))
    ok, _, T = _factor_dev(S, T, cap, 0, 32)
    assert ok
    ok, ratio, _ = _factor_dev(S, T, cap, 32, 64)
    ok_ref, _ = Davidson._factor(None, S, 32, 64, None)
    print(f"[factor] dependent border: accepted={ok}, min ratio {ratio:.3e}, oracle accepted={ok_ref}")
    assert not ok and not ok_ref


def test_inverse_cholesky_wide_borders_against_oracle():
    """Borders wider than one factor block (288 new columns as three blocks
    of 96, 300 as four of 75) against the oracle's factor."""
    from oracle.davidson import Davidson
    rng = np.random.default_rng(4)
    cap, n = 600, 2000
    V = rng.standard_normal((n, cap)) * rng.uniform(0.5, 2.0, cap)[None, :]
    S = V.T @ V
    T = np.zeros((cap, cap))
    for frm, to in [(0, 288), (288, 588), (588, 600)]:
        print(f"[factor wide] {frm}->{to}", flush=True)
        ok, ratio, T = _factor_dev(S, T, cap, frm, to)
        ok_ref, Tref = Davidson._factor(None, S, frm, to, None)
        assert ok and ok_ref
        err = np.abs(T[:to, :to] - Tref).max() / np.abs(Tref).max()
        ident = np.abs(T[:to, :to] @ S[:to, :to] @ T[:to, :to].T - np.eye(to)).max()
        print(f"[factor wide] {frm}->{to}: T vs oracle {err:.2e}, |T S T^T - I| {ident:.2e}, min ratio {ratio:.3e}")
        assert err < 1e-11 and ident < 1e-11
        assert np.all(T[:to, :to][np.triu_indices(to, 1)] == 0.0)
        assert 0.0 < ratio <= 1.0


@pytest.mark.parametrize("dependent", [60, 250])
def test_inverse_cholesky_wide_border_rejects_dependence(dependent):
    """A 300-column border (four blocks of 75) with one column numerically
    dependent on an earlier one, in the first or the last block: rejected,
    as the oracle rejects it, with the failing pivot ratio reported."""
    from oracle.davidson import Davidson
    rng = np.random.default_rng(5)
    cap, n = 320, 1200
    V = rng.standard_normal((n, cap))
    V[:, dependent] = V[:, 10] + 1e-9 * V[:, dependent + 1]
    S = V.T @ V
    T = np.zeros((cap, cap))
    ok, ratio, _ = _factor_dev(S, T, cap, 0, 300)
    ok_ref, _ = Davidson._factor(None, S, 0, 300, None)
    print(f"[factor wide] dependent column {dependent}: accepted={ok}, min ratio {ratio:.3e}, oracle {ok_ref}")
    assert not ok and not ok_ref
    assert ratio < np.finfo(float).eps / 1e-12


def _coop(setting=None):
    """Read, or set and return the previous, {workers, from, to, spins} of the
    cooperative tridiagonalisation."""
    from baandi_amd import lib as L
    v = np.zeros(4, dtype=np.int32)
    L.check(L.load().bd_sym_eigen_coop(L.ptr(v), 0))
    old = v.copy()
    if setting is not None:
        v[:] = setting
        L.check(L.load().bd_sym_eigen_coop(L.ptr(v), 1))
    return old


# m and the worker count: one worker (the whole matrix in one LDS), worker
# counts that do not divide m, the Si64 subspaces, and the largest sizes the
# workers' LDS holds (Si128's 576 and 700 over 24 and 32 workers).
@pytest.mark.parametrize("m,want,workers", [(3, 3, 1), (5, 2, 2), (40, 40, 3), (141, 70, 1), (160, 160, 8),
                                            (240, 160, 5), (320, 160, 8), (320, 160, 16), (321, 100, 7),
                                            (576, 288, 24), (700, 60, 32)])
def test_eigen_cooperative_tridiagonalisation(m, want, workers):
    """The tridiagonalisation split by columns over cooperating workgroups
    against LAPACK: random, degenerate (exact triples) and realified
    (exact pairs) matrices."""
    old = _coop([workers, 3, 1025, 1 << 21])
    try:
        rng = np.random.default_rng(m + workers)
        B = rng.standard_normal((m, m))
        _check_pairs(0.5 * (B + B.T), want, f"coop x{workers} random")
        vals = np.sort(rng.uniform(-1.0, 3.0, m))
        if m >= 9:
            vals[3:6] = vals[3]
        Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
        _check_pairs((Q * vals) @ Q.T, want, f"coop x{workers} degenerate")
        if m % 2 == 0 and m <= 320:
            h = m // 2
            C = rng.standard_normal((h, h)) + 1j * rng.standard_normal((h, h))
            H = 0.5 * (C + C.conj().T)
            _check_pairs(np.block([[H.real, -H.imag], [H.imag, H.real]]), min(want, m), f"coop x{workers} realified")
    finally:
        _coop(old)


@pytest.mark.parametrize("m,want", [(64, 64), (320, 160), (523, 60)])
def test_eigen_panel_kernel_without_cooperation(m, want):
    """With the cooperative path off, the one-workgroup panel kernel."""
    old = _coop([0, 3, 1025, 1 << 21])
    try:
        rng = np.random.default_rng(m * 3 + 1)
        B = rng.standard_normal((m, m))
        _check_pairs(0.5 * (B + B.T), want, "panel")
    finally:
        _coop(old)


def test_eigen_cooperative_abandon_is_reported():
    """A worker that may not wait at all (spins < 0, forced from the host)
    abandons the solve: bd_sym_eigen_dev reports the failure instead of
    returning a tridiagonal form, and the next solve with the default
    setting is correct again."""
    from baandi_amd import lib as L
    rng = np.random.default_rng(17)
    B = rng.standard_normal((200, 200))
    A = 0.5 * (B + B.T)
    old = _coop([8, 3, 1025, -1])
    try:
        with pytest.raises(L.BaandiError):
            _eigen_dev(A, 20)
    finally:
        _coop(old)
    _check_pairs(A, 20, "coop after abandon")


def _herm_dev(A, want):
    import torch
    from baandi_amd import lib as L
    lib = L.load()
    m = A.shape[0]
    Af = np.asfortranarray(A).ravel(order="F")
    Ad = torch.from_numpy(np.stack([Af.real, Af.imag], axis=1).ravel().copy()).cuda()
    vals = torch.zeros(want, dtype=torch.float64, device="cuda")
    Y = torch.zeros(2 * m * want, dtype=torch.float64, device="cuda")
    work = torch.zeros(lib.bd_sym_eigen_workspace(m, want), dtype=torch.float64, device="cuda")
    L.check(lib.bd_herm_eigen_dev(m, want, L.ptr(Ad), m, L.ptr(vals), L.ptr(Y), m, L.ptr(work)))
    y = Y.cpu().numpy().reshape(want, m, 2)
    return vals.cpu().numpy(), (y[..., 0] + 1j * y[..., 1]).T


def _check_herm(A, want, label):
    w, Y = _herm_dev(A, want)
    ref = np.linalg.eigvalsh(A)[:want]
    m = A.shape[0]
    scale = max(np.abs(A).sum(axis=0).max(), 1e-300)
    verr = np.abs(w - ref).max() / scale
    res = np.abs(A @ Y - Y * w[None, :]).max() / scale
    orth = np.abs(Y.conj().T @ Y - np.eye(want)).max()
    print(f"[herm] {label} m={m} want={want}: value {verr:.2e}, residual {res:.2e}, orthogonality {orth:.2e}")
    assert verr < 1e-13 * max(m, 10)
    assert res < 1e-12 * max(m, 10)
    assert orth < 1e-12 * max(m, 10)


@pytest.mark.parametrize("m,want", [(1, 1), (2, 2), (3, 2), (16, 16), (32, 16), (64, 40), (160, 80), (300, 120)])
def test_herm_eigen_random(m, want):
    """The complex Hermitian eigensolver (complex Householder reduction to a
    real tridiagonal, the real bisection and inverse iteration, the complex
    back-transform) against LAPACK's zheevd through numpy."""
    rng = np.random.default_rng(1000 + m)
    B = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    _check_herm(0.5 * (B + B.conj().T), want, "random")


@pytest.mark.parametrize("m", [12, 48, 96])
def test_herm_eigen_degenerate_and_real(m):
    """Exact multiplicities and a 1e-13 cluster under a random unitary, a
    real symmetric input (zero imaginary parts: every reflector real) and a
    diagonal one (every tau zero)."""
    rng = np.random.default_rng(m + 5)
    vals = np.sort(rng.uniform(-1.0, 3.0, m))
    vals[3:6] = vals[3]
    vals[8:10] = vals[8]
    vals[-4:-1] = vals[-4] + np.array([0.0, 1e-13, 2e-13])
    vals = np.sort(vals)
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
    A = (Q * vals[None, :]) @ Q.conj().T
    _check_herm(0.5 * (A + A.conj().T), m // 2 + 1, "degenerate")
    B = rng.standard_normal((m, m))
    _check_herm(0.5 * (B + B.T).astype(complex), m // 2, "real")
    _check_herm(np.diag(vals).astype(complex), m // 2, "diagonal")
