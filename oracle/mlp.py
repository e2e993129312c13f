"""CPU restatement of the learned Davidson preconditioner.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker of the HIP kernels; never shipped
or called by the product path.

The reference has no learned preconditioner: it applies the Teter-Payne-Allan
diagonal (crates/dft-solvers/src/lib.rs:176-280) as e = -r K(T_G/<T>), or the
QE correction e = r / d_QE(h, theta) (ref_davidson/stream.rs:16-22). The
learned form scales the QE correction by a row factor, e = -K r / d_QE(h,
theta) (the sign is immaterial to Davidson), with K from a 4 -> 64 -> 64 -> 1
tanh network of per-plane-wave features:

    u = [log(1 + g/T), log(T), (theta - vbar)/T, vbar]
    h1 = tanh(W1 u + b1), h2 = tanh(W2 h1 + b2), o = w3 . h2 + b3, K = exp(o)

g = |G|^2/2 of the row (Ha), T = the band's kinetic scale <x|T|x>/<x|x>
floored at 0.1 Ha (lib.rs:178-193), theta the Ritz value, vbar the mean local
potential. With bf16=True the weights, and h1 before the 64 x 64 product, are
rounded to bfloat16 (round to nearest even) as the MFMA kernels do; sums stay
in high precision. Gradients treat that rounding as the identity.

Training loss per band (scale invariant, because Davidson normalises every
correction): L = 1 - <a, t>^2 / (|a|^2 |t|^2), a = K r, t = the ideal
correction (the converged eigenvector's component outside the current Ritz
block), r the QE-scaled residual r / d_QE. The batch loss is the mean over
bands. K = 1 (a zero output layer) is the QE correction.

Parity: the backward pass is pinned by central finite differences of the
float64 forward (tests/test_oracle_mlp.py); the GPU kernels are checked
against both modes.
"""
from __future__ import annotations

import math

import numpy as np

HIDDEN = 64
FEATURES = 4
KINETIC_FLOOR = 0.1  # kinetic_scale's floor (crates/dft-solvers/src/lib.rs:178-193)
# Packed parameter order (float32): W1 (64x4 row-major), b1, W2 (64x64 row-major), b2, w3, b3.
SIZES = [("W1", (HIDDEN, FEATURES)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
         ("w3", (HIDDEN,)), ("b3", (1,))]
COUNT = sum(int(np.prod(s)) for _, s in SIZES)


def bf16_round(x):
    """float -> nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in SIZES])


def init_params(seed=0):
    """Glorot-uniform weights, zero biases; w3 small so K starts near 1."""
    rng = np.random.default_rng(seed)
    p = {}
    lim1 = math.sqrt(6.0 / (FEATURES + HIDDEN))
    lim2 = math.sqrt(6.0 / (2 * HIDDEN))
    p["W1"] = rng.uniform(-lim1, lim1, (HIDDEN, FEATURES))
    p["b1"] = np.zeros(HIDDEN)
    p["W2"] = rng.uniform(-lim2, lim2, (HIDDEN, HIDDEN))
    p["b2"] = np.zeros(HIDDEN)
    p["w3"] = rng.uniform(-0.1, 0.1, HIDDEN)
    p["b3"] = np.zeros(1)
    return pack(p).astype(np.float32).astype(np.float64)


def kinetic_scale(x, g):
    """<x|T|x>/<x|x> floored at 0.1 Ha (lib.rs:178-193)."""
    w = x * x
    return max(float(np.dot(w, g) / np.sum(w)), KINETIC_FLOOR)


def features(g, T, theta, vbar):
    g = np.asarray(g, dtype=np.float64)
    u = np.empty((g.size, FEATURES))
    u[:, 0] = np.log1p(g / T)
    u[:, 1] = math.log(T)
    u[:, 2] = (theta - vbar) / T
    u[:, 3] = vbar
    return u


def forward(theta, u, bf16=False):
    p = unpack(theta)
    rnd = bf16_round if bf16 else (lambda a: a)
    W1, W2, w3 = rnd(p["W1"]), rnd(p["W2"]), rnd(p["w3"])
    z1 = u @ W1.T + p["b1"]
    h1 = np.tanh(z1)
    h1q = rnd(h1)
    z2 = h1q @ W2.T + p["b2"]
    h2 = np.tanh(z2)
    o = h2 @ w3 + p["b3"][0]
    K = np.exp(o)
    return dict(u=u, h1=h1, h1q=h1q, h2=h2, o=o, K=K, W1=W1, W2=W2, w3=w3)


def backward(theta, cache, dL_do, bf16=False):
    """Parameter gradient (packed) for an upstream dL/do per element."""
    rnd = bf16_round if bf16 else (lambda a: a)
    d2 = (dL_do[:, None] * cache["w3"][None, :]) * (1.0 - cache["h2"] ** 2)
    d2q = rnd(d2)
    g = {}
    g["w3"] = dL_do @ cache["h2"]
    g["b3"] = np.array([dL_do.sum()])
    g["W2"] = d2q.T @ cache["h1q"]
    g["b2"] = d2.sum(axis=0)
    d1 = (d2q @ cache["W2"]) * (1.0 - cache["h1"] ** 2)
    g["W1"] = d1.T @ cache["u"]
    g["b1"] = d1.sum(axis=0)
    return pack(g)


def band_loss(K, r, t):
    """L = 1 - <a,t>^2/(|a|^2 |t|^2), a = K r; returns L and dL/dK."""
    a = K * r
    c, A, B = float(a @ t), float(a @ a), float(t @ t)
    L = 1.0 - c * c / (A * B)
    dL_da = -(2.0 * c / (A * B)) * t + (2.0 * c * c / (A * A * B)) * a
    return L, dL_da * r


def loss_and_grad(theta, batch, bf16=False):
    """batch: list of (g, T, theta_ritz, vbar, r, t). Mean band loss and its gradient."""
    total, grad = 0.0, np.zeros(COUNT)
    for g, T, th, vb, r, t in batch:
        cache = forward(theta, features(g, T, th, vb), bf16)
        L, dL_dK = band_loss(cache["K"], r, t)
        total += L
        grad += backward(theta, cache, dL_dK * cache["K"], bf16)
    n = len(batch)
    return total / n, grad / n


def adam_step(theta, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) on float32 master weights."""
    m = b1 * m + (1.0 - b1) * grad
    v = b2 * v + (1.0 - b2) * grad * grad
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def correction(theta, r, x, g, ritz, vbar, diag, bf16=True):
    """e = -K r / d_QE(diag, ritz) for one band (the inference the Davidson
    step applies: the network's row factor on the QE correction)."""
    from .davidson import sqrt_denominator
    T = kinetic_scale(x, g)
    return -forward(theta, features(g, T, ritz, vbar), bf16)["K"] * r / sqrt_denominator(diag, ritz)


def fallback_after(ratios, limit=1.0, strikes=2):
    """The fallback rule (back to the QE correction): the first iteration index at which the
    residual-reduction ratio has been >= limit for `strikes` consecutive
    iterations (None if never). A non-finite ratio counts as a strike."""
    run = 0
    for i, q in enumerate(ratios):
        run = run + 1 if (not np.isfinite(q) or q >= limit) else 0
        if run >= strikes:
            return i
    return None
