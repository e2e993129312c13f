"""Oracle: Gamma ORC Hamiltonian apply H·X and the dense brute-force H.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Parity: ``apply_orc`` is checked against the dense assembly V(G-G') + T +
V_NL (``dense_orc``) on small cells to 1e-12 and against the free-electron
spectrum under constant V (tests/test_oracle_hamiltonian.py); the pair
packing/unpacking is checked against independent single-column transforms.
"""
from __future__ import annotations

import math

import numpy as np

from .physics import orc_unpack, orc_pack, map_miller

SQRT2 = math.sqrt(2.0)


def pack_pair(system, xa, xb):
    """Two ORC columns -> one complex full-sphere vector psi_a + i psi_b
    (packed_value, distributed_real_gamma.rs:651-682). ``xb`` may be None."""
    half = system.half
    n = len(xa)
    out = np.empty(n, dtype=np.complex128)
    a = (xa[1::2] + 1j * xa[2::2]) / SQRT2
    b = (xb[1::2] + 1j * xb[2::2]) / SQRT2 if xb is not None else np.zeros_like(a)
    out[0] = xa[0] + 1j * (xb[0] if xb is not None else 0.0)
    out[half.plus[1:]] = a + 1j * b  # a + (-b.im, b.re)
    out[half.minus[1:]] = np.conj(a) + 1j * np.conj(b)  # conj(a) + (b.im, b.re)
    return out


def unpack_pair(system, h):
    """Even/odd split of h(+G), h(-G) back to two ORC columns
    (mirror_back_and_unpack, distributed_real_gamma.rs:1316-1339)."""
    half = system.half
    n = len(h)
    ya = np.empty(n)
    yb = np.empty(n)
    hp = h[half.plus[1:]]
    hm = h[half.minus[1:]]
    even = (hp + np.conj(hm)) * 0.5
    odd = (hp - np.conj(hm)) * 0.5
    ya[0] = h[0].real
    yb[0] = h[0].imag
    ya[1::2] = even.real * SQRT2
    ya[2::2] = even.imag * SQRT2
    yb[1::2] = odd.imag * SQRT2
    yb[2::2] = -odd.real * SQRT2
    return ya, yb


def local_apply_full(system, c, potential):
    """Scatter, unscaled G->r, times V(r), r->G with 1/N, gather
    (DistributedPairFft::apply_local_batch_from, distributed_fft.rs:1240-1288)."""
    dims = system.dims
    ntot = dims[0] * dims[1] * dims[2]
    g = np.zeros(ntot, dtype=np.complex128)
    g[system.fft_index] = c
    psi = np.fft.ifftn(g.reshape(dims)) * ntot
    h = np.fft.fftn(psi * potential) / ntot
    return h.reshape(-1)[system.fft_index]


def apply_nonlocal_orc(system, x, y):
    """Y += P D P^T X with per-atom D blocks (distributed_real_gamma.rs:1166-1251)."""
    if system.orc_p.shape[1] == 0:
        return y
    q = system.orc_p.T @ x
    w = np.empty_like(q)
    for start, nc, d in system.coupling:
        w[start:start + nc] = d @ q[start:start + nc]
    y += system.orc_p @ w
    return y


def apply_orc(system, x, potential):
    """H·X for an ORC block (apply_block, distributed_real_gamma.rs:1062-1264):
    pair packing, local FFT apply, unpack, + T_G x, + P D P^T x. Overwrites."""
    x = np.asarray(x, dtype=np.float64)
    n, m = x.shape
    y = np.empty_like(x)
    for t in range((m + 1) // 2):
        ca, cb = 2 * t, 2 * t + 1
        xb = x[:, cb] if cb < m else None
        c = pack_pair(system, x[:, ca], xb)
        h = local_apply_full(system, c, potential)
        ya, yb = unpack_pair(system, h)
        y[:, ca] = ya
        if cb < m:
            y[:, cb] = yb
    y += system.kinetic_orc[:, None] * x
    return apply_nonlocal_orc(system, x, y)


def dense_complex(system, potential):
    """Dense complex H over the full sphere: V(G-G') from the grid transform
    of V, T on the diagonal, beta D beta^H (the test oracle of
    crates/dft-physics/tests/all/hamiltonian.rs:288-417)."""
    dims = system.dims
    vg = np.fft.fftn(potential) / potential.size
    m = system.basis.miller
    diff = m[:, None, :] - m[None, :, :]
    idx = np.stack([np.mod(diff[..., i], dims[i]) for i in range(3)], -1)
    h = vg[idx[..., 0], idx[..., 1], idx[..., 2]].astype(np.complex128)
    h[np.diag_indices_from(h)] += 0.5 * system.basis.g2
    p = system.projectors
    for start, nc, d in system.coupling:
        pp = p[:, start:start + nc]
        h += pp @ d @ pp.conj().T
    return h


def dense_orc(system, potential):
    """Dense real ORC matrix U^H H U with U the ORC->complex isometry."""
    n = len(system.basis.g2)
    eye = np.eye(n)
    u = orc_unpack(system.half, eye)  # (n complex rows, n ORC columns)
    h = dense_complex(system, potential)
    m = u.conj().T @ h @ u
    return m.real
