"""Host mirror of the SCF field step (bd_field_*): the density-sphere
coefficients and their grid transforms, V = V_ion + V_H + V_xc of an input
density, the residual and its Hartree-metric norm dr2, the mixing products
and update, and the Hartree / XC / local energy sums, all in
libbaandi_amd.so (crates/dft-engine/src/refscf/map.rs:41-198, metric.rs,
broyden.rs:317-445). Coefficients are complex128 CUDA tensors on the density
sphere (4 Ecut), fields real float64 CUDA tensors on the FFT grid.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import lib as L

# bd_field_gram_dev and bd_field_mix_dev take at most this many history rows.
MAX_HISTORY = 9


def gaussian_start(system):
    """Start density on the density sphere: superposed Gaussian atomic charges
    Z (alpha/pi)^1.5 e^{-alpha r^2}, alpha = 1/rc^2, i.e. coefficients
    Z/Omega e^{-G^2 rc^2/4} sum_a e^{-i G.tau_a}, with G = 0 holding
    electrons/Omega (complex128 numpy array, density-sphere order)."""
    atoms = system.get("atoms_cart")
    recip = 2.0 * math.pi * np.linalg.inv(np.array(system.lattice_rows)).T
    miller = system.get("density_miller")
    g2 = system.get("density_g2")
    sfac = np.exp(-1j * ((miller.astype(np.float64) @ recip) @ atoms.T)).sum(axis=1)
    x = system.z_valence / system.volume * np.exp(-g2 * system.rc ** 2 / 4.0) * sfac
    x[np.all(miller == 0, axis=1)] = system.electrons / system.volume
    return x


class DensityField:
    def __init__(self, system):
        import torch
        lib = L.load()
        self.system = system
        h = C.c_void_p()
        L.check(lib.bd_field_create(system.handle, C.byref(h)))
        self.handle = h
        self.dims = tuple(system.dims)
        self.size = int(system.dims[0]) * int(system.dims[1]) * int(system.dims[2])
        self.nd = int(system.ndensity)
        self._torch_stream = torch.cuda.ExternalStream(lib.bd_field_stream(h))

    def enter(self):
        """Order the field's stream after torch's current stream."""
        import torch
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream())
        self._torch_stream.wait_event(ev)

    def leave(self):
        import torch
        ev = torch.cuda.Event()
        ev.record(self._torch_stream)
        torch.cuda.current_stream().wait_event(ev)

    def _complex(self, t):
        import torch
        assert t.dtype == torch.complex128 and t.is_contiguous() and t.numel() == self.nd
        return L.ptr(torch.view_as_real(t))

    def to_grid(self, coeffs):
        """Unnormalised inverse transform of sphere coefficients: complex grid."""
        import torch
        out = torch.empty(self.dims, dtype=torch.complex128, device="cuda")
        self.enter()
        L.check(L.load().bd_field_to_grid_dev(self.handle, self._complex(coeffs.contiguous()),
                                              L.ptr(torch.view_as_real(out))))
        self.leave()
        return out

    def from_grid(self, field):
        """Forward transform / N of a real grid field, gathered on the sphere."""
        import torch
        out = torch.empty(self.nd, dtype=torch.complex128, device="cuda")
        f = field.reshape(-1).contiguous()
        self.enter()
        L.check(L.load().bd_field_from_grid_dev(self.handle, L.ptr(f), self._complex(out)))
        self.leave()
        return out

    def potential(self, x, vion, out):
        """out = V_ion + V_H[x] + v_xc(|rho_in|) on the grid; returns mean(V)."""
        mean = C.c_double(0.0)
        self.enter()
        L.check(L.load().bd_field_potential_dev(self.handle, self._complex(x), L.ptr(vion.reshape(-1)),
                                                L.ptr(out.reshape(-1)), C.byref(mean)))
        self.leave()
        return mean.value

    def residual(self, rho_out, x, fres):
        """fres = coefficients of rho_out (with the G = 0 charge) - x; returns dr2 (Ry)."""
        dr2 = C.c_double(0.0)
        self.enter()
        L.check(L.load().bd_field_residual_dev(self.handle, L.ptr(rho_out.reshape(-1).contiguous()), self._complex(x),
                                               self._complex(fres), C.byref(dr2)))
        self.leave()
        return dr2.value

    def gram(self, rows, f):
        """Metric products of stacked coefficient rows (p, nd) with each other and with f."""
        import numpy as np
        import torch
        p = rows.shape[0]
        a = np.zeros((p, p))
        c = np.zeros(p)
        self.enter()
        L.check(L.load().bd_field_gram_dev(self.handle, p, L.ptr(torch.view_as_real(rows.contiguous())),
                                           self._complex(f), L.ptr(a), L.ptr(c)))
        self.leave()
        return a, c

    def mix(self, dX, dF, gamma, x, f, beta, screening, out):
        """out = (x - sum gamma dX) + beta (f - sum gamma dF) [* Thomas-Fermi screening]."""
        import numpy as np
        import torch
        p = 0 if dX is None else dX.shape[0]
        g = np.ascontiguousarray(gamma if p else np.zeros(1), dtype=np.float64)
        self.enter()
        L.check(L.load().bd_field_mix_dev(
            self.handle, p, L.ptr(torch.view_as_real(dX)) if p else None,
            L.ptr(torch.view_as_real(dF)) if p else None, L.ptr(g), self._complex(x), self._complex(f), float(beta),
            1 if screening else 0, self._complex(out)))
        self.leave()
        return out

    def energies(self, rho, vion):
        """(Hartree, XC, local) energies of a grid density (Ha)."""
        import numpy as np
        out = np.zeros(3)
        self.enter()
        L.check(L.load().bd_field_energies_dev(self.handle, L.ptr(rho.reshape(-1).contiguous()),
                                               L.ptr(vion.reshape(-1).contiguous()), L.ptr(out)))
        self.leave()
        return float(out[0]), float(out[1]), float(out[2])

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                L.load().bd_field_destroy(h)
            except (TypeError, AttributeError):
                pass
            self.handle = None


class Mixer:
    """Difference-history Anderson update in the half-sphere Hartree metric
    (crates/dft-engine/src/refscf/broyden.rs:317-445) on the field kernels.

    Keeps the last ``history`` pairs (x_prev - x, f_prev - f); the mixing
    coefficients solve the p x p metric Gram system of the residual
    differences by an eigendecomposition on the host (a rank-deficient history
    drops its oldest pair and tries again). ``preconditioning``
    "thomas_fermi" scales the history-reduced residual by |G|^2/(|G|^2 + a),
    a = (12/pi)^(2/3)/r_s (broyden.rs:106-113, applied to the residual term
    only, :439); "plain" (profile.rs:17-26, the default) leaves it unscaled."""

    def __init__(self, field: DensityField, beta, history, preconditioning="plain"):
        if preconditioning not in ("plain", "thomas_fermi"):
            raise ValueError(f"unknown preconditioning {preconditioning!r}")
        if not 0 <= int(history) <= MAX_HISTORY:
            raise ValueError(f"mixing history must be 0..{MAX_HISTORY}")
        self.field, self.beta, self.history = field, float(beta), int(history)
        self.screening = preconditioning == "thomas_fermi"
        self.pairs = []  # (dx, dF), newest last
        self.prev = None

    def update(self, x, f):
        """Next input density coefficients from input x and residual f."""
        import torch
        if self.prev is not None:
            px, pf = self.prev
            self.pairs.append((px - x, pf - f))
            if len(self.pairs) > self.history:
                self.pairs.pop(0)
        self.prev = (x.clone(), f.clone())
        gamma = dX = dF = None
        while self.pairs:
            p = len(self.pairs)
            dF = torch.stack([df for _, df in self.pairs])
            a, c = self.field.gram(dF, f)
            a = 0.5 * (a + a.T)
            lam, q = np.linalg.eigh(a)
            if lam[-1] <= 0.0 or lam[0] <= math.sqrt(p * np.finfo(float).eps) * lam[-1]:
                self.pairs.pop(0)  # rank-deficient history: drop the oldest pair
                dF = None
                continue
            gamma = q @ ((q.T @ c) / lam)
            dX = torch.stack([dx for dx, _ in self.pairs])
            break
        out = torch.empty_like(x)
        return self.field.mix(dX, dF, gamma, x, f, self.beta, self.screening, out)
