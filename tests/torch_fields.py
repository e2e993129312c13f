"""Torch FP64 restatement of the SCF field step, the reference the field
kernels (bd_field_*) are checked against: density-sphere coefficients and
torch.fft grid transforms, LDA-PZ per point, the half-sphere Hartree metric
and the difference-history Anderson mixer
(crates/dft-engine/src/refscf/map.rs:41-198, metric.rs:1-5,
broyden.rs:106-113 and :317-445, crates/dft-physics/src/xc.rs:48-73).
Test infrastructure: it is pinned by the KAT-1 SCF golden through the field
kernels it checks and by the NumPy LDA in oracle/fields.py."""
from __future__ import annotations

import math

import numpy as np


def lda_pz_torch(n):
    """Slater exchange + PZ81 correlation of a CUDA tensor: (eps_xc, v_xc),
    zero at or below 1e-10 (xc.rs:48-73)."""
    import torch
    live = n > 1e-10
    d = torch.where(live, n, torch.ones_like(n))
    cd = torch.pow(d, 1.0 / 3.0)
    rs = (3.0 / (4.0 * math.pi)) ** (1.0 / 3.0) / cd
    ex = -0.75 * (3.0 / math.pi) ** (1.0 / 3.0) * cd
    low = rs >= 1.0
    sr = torch.sqrt(rs)
    den = 1.0 + 1.0529 * sr + 0.3334 * rs
    lg = torch.log(rs)
    ec = torch.where(low, -0.1423 / den, 0.0311 * lg - 0.048 + 0.0020 * rs * lg - 0.0116 * rs)
    dec = torch.where(low, 0.1423 * (1.0529 / (2.0 * sr) + 0.3334) / (den * den),
                      0.0311 / rs + 0.0020 * (lg + 1.0) - 0.0116)
    zero = torch.zeros_like(n)
    eps = torch.where(live, ex + ec, zero)
    vxc = torch.where(live, 4.0 * ex / 3.0 + ec - rs * dec / 3.0, zero)
    return eps, vxc


class DensityField:
    """Density-sphere coefficients (4 Ecut) and their grid transforms."""

    def __init__(self, system):
        import torch
        dev = torch.device("cuda")
        self.dims = tuple(system.dims)
        self.size = int(np.prod(self.dims))
        self.flat = torch.from_numpy(system.get("density_index").astype(np.int64)).to(dev)
        g2 = system.get("density_g2")
        miller = system.get("density_miller")
        zero = np.all(miller == 0, axis=1)
        # Positive representatives carry the half-sphere Hartree weight
        # 8 pi Omega / G^2 in Ry.
        first = np.where(miller[:, 0] != 0, miller[:, 0], np.where(miller[:, 1] != 0, miller[:, 1], miller[:, 2]))
        positive = first > 0
        metric = np.zeros(len(g2))
        metric[positive] = 8.0 * math.pi * system.volume / g2[positive]
        coulomb = np.zeros(len(g2))
        coulomb[~zero] = 4.0 * math.pi / g2[~zero]
        self.g2 = g2
        self.zero_np = zero
        self.zero = torch.from_numpy(zero).to(dev)
        self.metric = torch.from_numpy(metric).to(dev)
        self.coulomb = torch.from_numpy(coulomb).to(dev)

    def to_grid(self, coeffs):
        import torch
        g = torch.zeros(self.size, dtype=torch.complex128, device=coeffs.device)
        g[self.flat] = coeffs
        return torch.fft.ifftn(g.view(self.dims)) * self.size

    def from_grid(self, field):
        import torch
        return (torch.fft.fftn(field.view(self.dims).to(torch.complex128)) / self.size).reshape(-1)[self.flat]

    def dot(self, a, b):
        return float((self.metric * (a.real * b.real + a.imag * b.imag)).sum())

    def gram(self, rows, f):
        w = self.metric
        re, im = rows.real, rows.imag
        wre, wim = re * w, im * w
        a = (wre[:, None, :] * re[None, :, :]).sum(dim=-1) + (wim[:, None, :] * im[None, :, :]).sum(dim=-1)
        c = (wre * f.real).sum(dim=-1) + (wim * f.imag).sum(dim=-1)
        return a.cpu().numpy(), c.cpu().numpy()


class Mixer:
    """Difference-history Anderson update with the Hartree metric."""

    def __init__(self, field: DensityField, beta, history, preconditioning="plain", volume=None, electrons=None):
        import torch
        self.field, self.beta, self.history = field, beta, history
        self.pairs = []
        self.prev = None
        self.screening = None
        if preconditioning == "thomas_fermi":
            rs = (3.0 * volume / (4.0 * math.pi * electrons)) ** (1.0 / 3.0)
            a = (12.0 / math.pi) ** (2.0 / 3.0) / rs
            g2 = torch.from_numpy(field.g2).to(field.metric.device)
            self.screening = g2 / (g2 + a)
        elif preconditioning != "plain":
            raise ValueError(f"unknown preconditioning {preconditioning!r}")

    def update(self, x, f):
        import torch
        if self.prev is not None:
            px, pf = self.prev
            self.pairs.append((px - x, pf - f))
            if len(self.pairs) > self.history:
                self.pairs.pop(0)
        self.prev = (x.clone(), f.clone())
        xx, ff = x.clone(), f.clone()
        while self.pairs:
            p = len(self.pairs)
            a, c = self.field.gram(torch.stack([df for _, df in self.pairs]), f)
            a = 0.5 * (a + a.T)
            lam, q = np.linalg.eigh(a)
            if lam[-1] <= 0.0 or lam[0] <= math.sqrt(p * np.finfo(float).eps) * lam[-1]:
                self.pairs.pop(0)
                continue
            gamma = torch.from_numpy(q @ ((q.T @ c) / lam)).to(x.device)
            dX = torch.stack([dx for dx, _ in self.pairs])
            dF = torch.stack([df for _, df in self.pairs])
            xx = xx - (gamma[:, None] * dX).sum(dim=0)
            ff = ff - (gamma[:, None] * dF).sum(dim=0)
            break
        if self.screening is not None:
            ff = ff * self.screening
        return xx + self.beta * ff
