"""The oracle of the learned preconditioner: backward pass against central
finite differences, loss invariants, bf16 rounding and the fallback rule."""
import math

import numpy as np
import pytest

from oracle import mlp as M


def _batch(seed, bands=3, n=40):
    rng = np.random.default_rng(seed)
    g = np.sort(rng.uniform(0.0, 12.0, n))
    out = []
    for _ in range(bands):
        r = rng.standard_normal(n)
        t = -r / (g + 1.0) + 0.1 * rng.standard_normal(n)
        out.append((g, rng.uniform(0.5, 3.0), rng.uniform(-0.3, 0.8), rng.uniform(-0.5, 0.1), r, t))
    return out


def test_backward_matches_finite_differences():
    theta = M.init_params(3)
    theta[M.COUNT - 1] = 0.2  # b3
    batch = _batch(1)
    L, grad = M.loss_and_grad(theta, batch)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(M.COUNT, 150, replace=False), [M.COUNT - 1, 0, 300]])
    h = 1e-6
    worst = 0.0
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (M.loss_and_grad(tp, batch)[0] - M.loss_and_grad(tm, batch)[0]) / (2 * h)
        # central differences carry ~1e-11 of rounding noise at h = 1e-6
        worst = max(worst, abs(fd - grad[i]) / (1e-5 * max(abs(fd), abs(grad[i])) + 1e-9))
    assert 0.0 <= L <= 1.0
    assert worst < 1.0, worst
    # b3 scales every K of a band alike, which the loss ignores.
    assert abs(grad[M.COUNT - 1]) < 1e-12


def test_band_loss_invariants():
    rng = np.random.default_rng(2)
    r = rng.standard_normal(50)
    K = np.exp(rng.standard_normal(50))
    t = rng.standard_normal(50)
    L1, _ = M.band_loss(K, r, t)
    L2, _ = M.band_loss(7.5 * K, r, t)
    assert abs(L1 - L2) < 1e-14  # a correction's scale is irrelevant to Davidson
    L0, d0 = M.band_loss(K, r, -3.0 * K * r)
    assert abs(L0) < 1e-14 and np.abs(d0).max() < 1e-12  # exact direction: minimum
    # A correction orthogonal to the target: loss 1.
    a = K * r
    t_perp = t - (a @ t) / (a @ a) * a
    assert abs(M.band_loss(K, r, t_perp)[0] - 1.0) < 1e-12


def test_bf16_rounding_ties_to_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.0e38, 1.0 + 2.0 ** -7])
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, -2.5, np.float32(3.0e38).astype(np.float64), 1.0 + 2.0 ** -7])
    got = M.bf16_round(x)
    assert np.array_equal(got[:4], want[:4]) and got[5] == want[5]
    assert abs(got[4] - 3.0e38) / 3.0e38 < 2.0 ** -8


def test_bf16_forward_close_to_exact():
    theta = M.init_params(5)
    g = np.linspace(0.0, 14.0, 200)
    u = M.features(g, 1.3, 0.2, -0.1)
    K = M.forward(theta, u)["K"]
    Kq = M.forward(theta, u, bf16=True)["K"]
    assert np.abs(Kq / K - 1.0).max() < 0.05  # bf16 weights: a few 2^-8 per layer


def test_fallback_rule():
    assert M.fallback_after([0.5, 0.7, 0.9, 0.4]) is None
    assert M.fallback_after([0.5, 1.2, 0.8, 1.1, 1.3, 0.2]) == 4
    assert M.fallback_after([1.0, 1.0]) == 1
    assert M.fallback_after([0.3, float("nan"), 2.0]) == 2
    assert M.fallback_after([1.5, 0.9, 1.5], strikes=1) == 0


def test_adam_moves_against_gradient():
    theta = M.init_params(0)
    batch = _batch(4)
    L0, g = M.loss_and_grad(theta, batch)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    th = theta
    for step in range(1, 30):
        L, g = M.loss_and_grad(th, batch)
        th, m, v = M.adam_step(th, g, m, v, step, lr=3e-3)
    assert M.loss_and_grad(th, batch)[0] < L0


def test_correction_scales_the_qe_correction():
    """A zero output layer (K = 1) gives -r / d_QE(diag, theta) exactly; a
    nonzero b3 alone scales every row by e^b3."""
    from oracle import davidson as D
    rng = np.random.default_rng(11)
    n = 300
    g = np.sort(rng.uniform(0.0, 15.0, n))
    diag = g + rng.uniform(-1.0, 2.0, n)
    r, x = rng.standard_normal(n), rng.standard_normal(n)
    ritz, vbar = -0.2, -0.4
    p = M.unpack(M.init_params(2))
    p["w3"] = np.zeros(M.HIDDEN)
    p["b3"] = np.zeros(1)
    e = M.correction(M.pack(p), r, x, g, ritz, vbar, diag, bf16=False)
    xq = 2.0 * (diag - ritz)
    d_ry = (1.0 + xq + np.sqrt(1.0 + (xq - 1.0) ** 2)) / 2.0  # ref_davidson.rs:246-255, Ry form
    np.testing.assert_allclose(e, -r / (d_ry / 2.0), rtol=1e-13)
    p["b3"] = np.array([0.7])
    e2 = M.correction(M.pack(p), r, x, g, ritz, vbar, diag, bf16=False)
    np.testing.assert_allclose(e2, math.exp(0.7) * e, rtol=1e-13)
