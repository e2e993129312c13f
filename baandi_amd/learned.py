"""Learned Davidson preconditioner: the network, its training pairs and its
training, all on the GPU through the bd_mlp_* kernels of libbaandi_amd.so.

The network scales the QE correction row by row: e = -K r / d_QE(h, theta)
(ref_davidson/stream.rs:16-22 has the QE denominator and
crates/dft-solvers/src/lib.rs:176-280 the preconditioners it sits beside;
there is no learned counterpart in the reference). K = exp(MLP(u)) with
u = [log(1 + g/T), log T, (theta - vbar)/T, vbar] per row: g = |G|^2/2 of the
row, T the column's kinetic scale, theta its Ritz value, vbar the mean local
potential. K = 1 is QE itself, and Davidson falls back to QE when the
network's residual reduction regresses (bd_davidson_solve_dev, BD_PRECOND_MLP).

Training pairs come from the FP64 GPU Davidson itself (QE corrections): from
a seeded start, the Ritz block after k iterations gives residuals
r = HX - theta X, stored QE-scaled as r / d_QE; the converged eigenvectors
give each column's ideal correction t (its closest exact eigenvector, within a
degenerate cluster, minus the column, with the current Ritz block projected
out). The loss per column, 1 - <K r, t>^2 / (|K r|^2 |t|^2), is blind to the
correction's scale, as Davidson is.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import lib as L

HIDDEN = 64
FEATURES = 4


class Network:
    """FP32 master parameters plus the device blob (bf16 weights) that the
    Davidson kernels read."""

    def __init__(self, theta):
        import torch
        lib = L.load()
        self.count = lib.bd_mlp_param_count()
        theta = np.asarray(theta, dtype=np.float32).ravel()
        if theta.size != self.count:
            raise ValueError(f"expected {self.count} parameters, got {theta.size}")
        self.theta = torch.from_numpy(theta.copy()).cuda()
        self.blob = torch.zeros(lib.bd_mlp_blob_floats(), dtype=torch.float32, device="cuda")
        self.m = torch.zeros_like(self.theta)
        self.v = torch.zeros_like(self.theta)
        self.steps = 0
        self.refresh()

    @classmethod
    def initial(cls, seed=0, output_scale=0.1):
        """Glorot-uniform weights, zero biases, output layer U(-s, s) with
        s = output_scale (0: K = 1 everywhere, the QE correction)."""
        rng = np.random.default_rng(seed)
        l1 = math.sqrt(6.0 / (FEATURES + HIDDEN))
        l2 = math.sqrt(6.0 / (2 * HIDDEN))
        parts = [rng.uniform(-l1, l1, HIDDEN * FEATURES), np.zeros(HIDDEN),
                 rng.uniform(-l2, l2, HIDDEN * HIDDEN), np.zeros(HIDDEN),
                 rng.uniform(-output_scale, output_scale, HIDDEN), np.zeros(1)]
        return cls(np.concatenate(parts))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as f:
            return cls(f["theta"])

    def save(self, path, **meta):
        np.savez(path, theta=self.parameters(), **{k: np.asarray(v) for k, v in meta.items()})

    def parameters(self):
        return self.theta.cpu().numpy().copy()

    def refresh(self):
        L.check(L.load().bd_mlp_quantize_dev(L.ptr(self.theta), L.ptr(self.blob)))

    def weights(self, g, T, ritz, vbar):
        """K for the rows g (CUDA float32) of one column."""
        import torch
        K = torch.empty_like(g)
        L.check(L.load().bd_mlp_forward_dev(g.numel(), L.ptr(g), float(T), float(ritz), float(vbar),
                                            L.ptr(self.blob), L.ptr(K)))
        return K


@dataclass
class Pairs:
    """Training columns on the GPU (float32): R, Tg (bands x n), scal (bands x 3:
    T, theta, vbar) and the shared row energies g (n)."""
    g: object
    R: object
    Tg: object
    scal: object

    @property
    def bands(self):
        return int(self.R.shape[0])

    @property
    def n(self):
        return int(self.R.shape[1])

    def extend(self, other):
        import torch
        if other.n != self.n:
            raise ValueError("pairs of different bases")
        return Pairs(self.g, torch.cat([self.R, other.R]), torch.cat([self.Tg, other.Tg]),
                     torch.cat([self.scal, other.scal]))


def kinetic_scale(x, g):
    """<x|T|x>/<x|x> floored at 0.1 Ha (lib.rs:178-193), per row of x."""
    w = x * x
    return ((w * g[None, :]).sum(dim=1) / w.sum(dim=1)).clamp_min(0.1)


def qe_denominator(h, e):
    """The QE smooth denominator d_QE(h, e) in Ha (ref_davidson/stream.rs:16-22),
    elementwise on tensors."""
    import torch
    x = 2.0 * (h - e)
    z = (x - 1.0).abs()
    return 0.25 * (2.0 * torch.clamp_min(x, 1.0) + 1.0 / (torch.sqrt(1.0 + z * z) + z))


def generate_pairs(operator, diag, vbar, bands, occupied, seed,
                   iterations=(1, 2, 3, 4, 5, 6, 8, 10, 12, 15, 19, 24, 30, 37),
                   extra=8, cluster_ha=1e-6, min_overlap=0.3):
    """Residual -> ideal-correction pairs from the FP64 Davidson on `operator`
    (potential installed), started from a seeded N(0,1) block."""
    import torch
    from .ref_davidson import Options, solve
    n = operator.global_len()
    g64 = torch.from_numpy(operator.system.get("kinetic_orc")).cuda()
    gen = torch.Generator(device="cpu").manual_seed(int(seed))
    nref = min(bands + extra, n)
    x0 = torch.randn(nref, n, generator=gen, dtype=torch.float64)
    x0 = (x0 / x0.norm(dim=1, keepdim=True)).cuda()
    ref = solve(operator, x0, diag, Options(bands=nref, threshold=1e-11, occupied=nref, first_step=True,
                                            max_iterations=400, preconditioner=L.BD_PRECOND_QE))
    w = torch.from_numpy(ref.eigenvalues).cuda()
    Psi = ref.eigenvectors
    start = x0[:bands].contiguous()
    Rs, Ts, S = [], [], []
    for k in iterations:
        sol = solve(operator, start, diag, Options(bands=bands, threshold=1e-14, occupied=bands, first_step=True,
                                                   max_iterations=int(k), return_hx=True,
                                                   preconditioner=L.BD_PRECOND_QE))
        X, HX = sol.eigenvectors, sol.hx
        th = torch.from_numpy(sol.eigenvalues).cuda()
        R = (HX - th[:, None] * X) / qe_denominator(diag[None, :], th[:, None])
        ov = X @ Psi.T                                    # (bands, nref)
        best = ov.abs().argmax(dim=1)
        keep = []
        for j in range(bands):
            jm = int(best[j])
            cl = (w - w[jm]).abs() < cluster_ha
            if jm >= nref - 2 or float((ov[j, cl] ** 2).sum()) < min_overlap:
                continue
            psi = (ov[j, cl][:, None] * Psi[cl]).sum(dim=0)
            psi = psi / psi.norm()
            t = psi - X[j]
            t = t - (X @ t) @ X
            if float(t.norm()) > 1e-10 and float(R[j].norm()) > 1e-12:
                keep.append((j, t))
        if not keep:
            continue
        idx = torch.tensor([j for j, _ in keep], device="cuda")
        T = kinetic_scale(X[idx], g64)
        Rs.append(R[idx].float())
        Ts.append(torch.stack([t for _, t in keep]).float())
        S.append(torch.stack([T, th[idx], torch.full_like(T, float(vbar))], dim=1).float())
    return Pairs(g64.float(), torch.cat(Rs).contiguous(), torch.cat(Ts).contiguous(), torch.cat(S).contiguous())


def loss_and_grad(net, pairs, batch_idx):
    """Mean batch loss and the FP32 gradient (CUDA tensor) of the network."""
    import torch
    lib = L.load()
    b = torch.as_tensor(np.asarray(batch_idx, dtype=np.int32)).cuda()
    work = torch.empty(lib.bd_mlp_train_workspace(pairs.n, len(batch_idx)), dtype=torch.float64, device="cuda")
    grad = torch.empty_like(net.theta)
    loss = C.c_double(0.0)
    L.check(lib.bd_mlp_loss_grad_dev(pairs.n, pairs.bands, L.ptr(pairs.g), L.ptr(pairs.R), L.ptr(pairs.Tg),
                                     L.ptr(pairs.scal), L.ptr(b), len(batch_idx), L.ptr(net.blob), L.ptr(grad),
                                     C.byref(loss), L.ptr(work)))
    return loss.value, grad


def train(net, pairs, steps, batch=16, lr=2e-3, seed=0):
    """`steps` Adam steps on the GPU; returns the per-step batch losses."""
    import torch
    lib = L.load()
    batch = min(batch, pairs.bands)
    work = torch.empty(lib.bd_mlp_train_workspace(pairs.n, batch), dtype=torch.float64, device="cuda")
    bidx = torch.empty(batch, dtype=torch.int32, device="cuda")
    grad = torch.empty_like(net.theta)
    losses = np.zeros(steps)
    # Adam's bias correction restarts with each call (fresh moments).
    net.m.zero_()
    net.v.zero_()
    L.check(lib.bd_mlp_train_dev(pairs.n, pairs.bands, L.ptr(pairs.g), L.ptr(pairs.R), L.ptr(pairs.Tg),
                                 L.ptr(pairs.scal), L.ptr(net.theta), L.ptr(net.m), L.ptr(net.v), int(steps),
                                 int(batch), float(lr), int(seed), losses.ctypes.data_as(C.c_void_p), L.ptr(bidx),
                                 L.ptr(net.blob), L.ptr(grad), L.ptr(work)))
    net.steps += steps
    return losses


def qe_loss(pairs):
    """The same loss at K = 1: for pairs whose R holds the QE-scaled residual,
    the loss of the QE correction (the baseline the network must beat)."""
    import torch
    a = pairs.R.double()
    t = pairs.Tg.double()
    c = (a * t).sum(dim=1)
    return float((1.0 - c * c / ((a * a).sum(dim=1) * (t * t).sum(dim=1))).mean())


def tpa_loss(pairs):
    """The same loss for the TPA weight applied to R as stored."""
    import torch
    g = pairs.g[None, :].double()
    T = pairs.scal[:, 0:1].double()
    x = g / T
    y = T / g.clamp_min(1e-300)
    p1 = 27.0 + x * (18.0 + x * (12.0 + x * 8.0))
    small = p1 / (p1 + 16.0 * x ** 4)
    p2 = y * (8.0 + y * (12.0 + y * (18.0 + y * 27.0)))
    big = p2 / (p2 + 16.0)
    K = torch.where(x <= 1.0, small, big)
    a = K * pairs.R.double()
    t = pairs.Tg.double()
    c = (a * t).sum(dim=1)
    return float((1.0 - c * c / ((a * a).sum(dim=1) * (t * t).sum(dim=1))).mean())
