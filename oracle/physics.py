"""Oracle: cell, plane-wave basis, FFT grid, ORC half sphere, synthetic
pseudopotential, radial transforms, local potential and projector tables.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Parity: pinned by KAT-1 (kinetic, nonlocal, local, Hartree, XC energies of the
stored converged Si2 orbitals), by the Gaussian Hankel closed forms of the
synthetic projectors, and by Ewald/grid/basis-count checks in
``tests/test_oracle_physics.py``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

TAU = 2.0 * math.pi
CUTOFF_TOLERANCE = 1e-12  # crates/dft-basis/src/lib.rs:12


# --------------------------------------------------------------------------
# Cell (crates/dft-geometry/src/lib.rs:43-96)
# --------------------------------------------------------------------------
def _inverse3(m):
    """Cofactor 3x3 inverse in the operation order of nalgebra's do_inverse3
    (the dependency behind Cell::from_matrix, crates/dft-geometry/src/lib.rs:66).
    Exact for 0/1 matrices, which keeps fcc/cubic reciprocal vectors exact."""
    m11 = m[1][1] * m[2][2] - m[2][1] * m[1][2]
    m12 = m[1][0] * m[2][2] - m[2][0] * m[1][2]
    m13 = m[1][0] * m[2][1] - m[2][0] * m[1][1]
    det = m[0][0] * m11 - m[0][1] * m12 + m[0][2] * m13
    out = [[0.0] * 3 for _ in range(3)]
    out[0][0] = m11 / det
    out[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) / det
    out[0][2] = (m[0][1] * m[1][2] - m[1][1] * m[0][2]) / det
    out[1][0] = -m12 / det
    out[1][1] = (m[0][0] * m[2][2] - m[2][0] * m[0][2]) / det
    out[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) / det
    out[2][0] = m13 / det
    out[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) / det
    out[2][2] = (m[0][0] * m[1][1] - m[1][0] * m[0][1]) / det
    return out, det


@dataclass
class Cell:
    """Lattice vectors are the COLUMNS of ``lattice`` (Bohr); user rows are
    vectors (Cell::from_user_rows_bohr, crates/dft-geometry/src/lib.rs:46)."""
    lattice: list  # [cartesian][vector]
    inverse: list
    volume: float

    @staticmethod
    def from_rows(rows):
        lat = [[float(rows[j][i]) for j in range(3)] for i in range(3)]
        scale = max(abs(x) for r in lat for x in r)
        scaled = [[x / scale for x in r] for r in lat]
        inv, det = _inverse3(scaled)
        inv = [[x / scale for x in r] for r in inv]
        volume = ((abs(det) * scale) * scale) * scale
        return Cell(lat, inv, volume)

    def reciprocal(self):
        """B[i][j] = 2*pi*inverse[j][i] (lib.rs:91-93): columns are b_j."""
        return [[TAU * self.inverse[j][i] for j in range(3)] for i in range(3)]

    def frac_to_cart(self, f):
        a = self.lattice
        return [sum(a[i][j] * f[j] for j in range(3)) for i in range(3)]

    def vector_lengths(self):
        a = self.lattice
        return [math.hypot(math.hypot(a[0][i], a[1][i]), a[2][i]) for i in range(3)]


# --------------------------------------------------------------------------
# Plane-wave basis and grid (crates/dft-basis/src/lib.rs:54-269, 359-502)
# --------------------------------------------------------------------------
@dataclass
class Basis:
    miller: np.ndarray  # (n, 3) int
    cart: np.ndarray  # (n, 3) float, G (not k+G)
    g2: np.ndarray  # (n,)
    k_frac: tuple = (0.0, 0.0, 0.0)
    kcart: np.ndarray = field(default_factory=lambda: np.zeros(3))
    q2: np.ndarray = None  # |k+G|^2


def _bounds(cell: Cell, ecut):
    """bounding_box (lib.rs:54-73) then the extra +1 of at_k (lib.rs:238)."""
    radius = math.sqrt(2.0 * ecut * (1.0 + CUTOFF_TOLERANCE))
    return [int(math.ceil(L * radius / TAU) + 1.0) + 1 for L in cell.vector_lengths()]


def plane_wave_basis(cell: Cell, ecut, k_frac=(0.0, 0.0, 0.0)) -> Basis:
    """PlaneWaveBasis::at_k (lib.rs:234-269): all Miller indices with
    |k+G|^2/2 <= Ecut (1+1e-12), sorted by |G|^2 (total order) then Miller."""
    b = cell.reciprocal()
    bounds = _bounds(cell, ecut)
    kc = [sum(b[i][j] * k_frac[j] for j in range(3)) for i in range(3)]
    r = [np.arange(-n, n + 1, dtype=np.int64) for n in bounds]
    m = np.stack(np.meshgrid(*r, indexing="ij"), axis=-1).reshape(-1, 3)
    mf = m.astype(np.float64)
    # cartesian[i] = ((0 + b[i][0] m0) + b[i][1] m1) + b[i][2] m2 (include, lib.rs:107)
    cart = np.empty_like(mf)
    for i in range(3):
        cart[:, i] = ((b[i][0] * mf[:, 0]) + b[i][1] * mf[:, 1]) + b[i][2] * mf[:, 2]
    g2 = (cart[:, 0] * cart[:, 0] + cart[:, 1] * cart[:, 1]) + cart[:, 2] * cart[:, 2]
    qc = cart + np.asarray(kc)[None, :]
    q2 = (qc[:, 0] ** 2 + qc[:, 1] ** 2) + qc[:, 2] ** 2
    keep = 0.5 * q2 <= ecut * (1.0 + CUTOFF_TOLERANCE)
    m, cart, g2, q2 = m[keep], cart[keep], g2[keep], q2[keep]
    order = np.lexsort((m[:, 2], m[:, 1], m[:, 0], g2))
    return Basis(m[order], cart[order], g2[order], tuple(k_frac), np.asarray(kc), q2[order])


def _smooth(n):
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def grid_dims(cell: Cell, ecut):
    """FftGridSpec::grid_dims (lib.rs:414-455): required = max(2|m|+1) over the
    density sphere at 4*Ecut, each raised to the next 2,3,5-smooth size."""
    dens = plane_wave_basis(cell, 4.0 * ecut)
    req = [int(2 * np.abs(dens.miller[:, i]).max() + 1) for i in range(3)]
    dims = []
    for d in req:
        while not _smooth(d):
            d += 1
        dims.append(d)
    return tuple(dims), dens


def map_miller(miller, dims):
    """FftGridSpec::map_miller (lib.rs:491-502): canonical [-n/2,(n-1)/2]."""
    out = np.empty_like(miller)
    for i in range(3):
        n = dims[i]
        mi = miller[:, i]
        if np.any(mi < -(n // 2)) or np.any(mi > (n - 1) // 2):
            raise ValueError("Miller index outside the canonical frequencies")
        out[:, i] = np.where(mi >= 0, mi, n + mi)
    return out


def flat_index(idx, dims):
    return (idx[:, 0] * dims[1] + idx[:, 1]) * dims[2] + idx[:, 2]


@dataclass
class HalfSphere:
    """crates/dft-basis/src/half.rs:31-153."""
    plus: np.ndarray
    minus: np.ndarray
    mirror: np.ndarray


def half_sphere(basis: Basis) -> HalfSphere:
    """Representatives: zero, then lexicographically positive Miller triples in
    full-list order (half.rs:41-79)."""
    n = len(basis.miller)
    lookup = {tuple(m): i for i, m in enumerate(basis.miller.tolist())}
    mirror = np.array([lookup[(-a, -b, -c)] for a, b, c in basis.miller.tolist()], dtype=np.int64)
    plus, minus = [0], [0]
    for i, m in enumerate(basis.miller.tolist()):
        if tuple(m) > (0, 0, 0):
            plus.append(i)
            minus.append(int(mirror[i]))
    assert 2 * len(plus) - 1 == n
    return HalfSphere(np.array(plus), np.array(minus), mirror)


def orc_pack(half: HalfSphere, full):
    """HalfSphere::pack (half.rs:113-138): real[0]=Re c0, real[2i-1]=sqrt2 Re,
    real[2i]=sqrt2 Im of the representative. ``full`` is (n,) or (n, m)."""
    full = np.asarray(full)
    out = np.empty(full.shape, dtype=np.float64)
    out[0] = full[0].real
    z = full[half.plus[1:]]
    out[1::2] = math.sqrt(2.0) * z.real
    out[2::2] = math.sqrt(2.0) * z.imag
    return out


def orc_unpack(half: HalfSphere, real):
    """HalfSphere::unpack (half.rs:140-153): Hermitian completion; each
    component divided by sqrt 2 on its own, as Complex64 / f64 does."""
    real = np.asarray(real, dtype=np.float64)
    full = np.empty(real.shape, dtype=np.complex128)
    full[0] = real[0]
    z = np.empty(real[1::2].shape, dtype=np.complex128)
    z.real = real[1::2] / math.sqrt(2.0)
    z.imag = real[2::2] / math.sqrt(2.0)
    full[half.plus[1:]] = z
    full[half.minus[1:]] = np.conj(z)
    return full


# --------------------------------------------------------------------------
# Radial quadrature (crates/dft-upf/src/radial.rs)
# --------------------------------------------------------------------------
def integrate(values):
    """Composite Simpson in the mesh index with a final 3/8 panel for even
    counts, Kahan-compensated (radial.rs:56-81). ``values`` already include rab."""
    v = np.asarray(values, dtype=np.float64)
    n = len(v)
    if n < 2:
        return 0.0
    if n == 2:
        return (v[0] + v[1]) * 0.5
    last = n - 4 if n % 2 == 0 else n - 1
    terms = (v[0:last:2] + 4.0 * v[1:last + 1:2] + v[2:last + 2:2]) / 3.0
    if n % 2 == 0:
        tail = 3.0 * (v[last] + 3.0 * v[last + 1] + 3.0 * v[last + 2] + v[last + 3]) / 8.0
        terms = np.append(terms, tail)
    s = 0.0
    c = 0.0
    for x in terms.tolist():
        y = x - c
        t = s + y
        c = (t - s) - y
        s = t
    return s


def integrate_rows(values):
    """Vectorised ``integrate`` over the last axis (same panels; the Kahan
    loop runs over panels for all rows at once)."""
    v = np.asarray(values, dtype=np.float64)
    n = v.shape[-1]
    last = n - 4 if n % 2 == 0 else n - 1
    terms = (v[..., 0:last:2] + 4.0 * v[..., 1:last + 1:2] + v[..., 2:last + 2:2]) / 3.0
    if n % 2 == 0:
        tail = 3.0 * (v[..., last] + 3.0 * v[..., last + 1] + 3.0 * v[..., last + 2]
                      + v[..., last + 3]) / 8.0
        terms = np.concatenate([terms, tail[..., None]], axis=-1)
    s = np.zeros(v.shape[:-1])
    c = np.zeros(v.shape[:-1])
    for k in range(terms.shape[-1]):
        y = terms[..., k] - c
        t = s + y
        c = (t - s) - y
        s = t
    return s


def bessel(l, x):
    """Spherical Bessel j_l, l<=3 (radial.rs:90-117): power series below 4,
    upward recurrence from sin/cos above. Vectorised over x."""
    x = np.asarray(x, dtype=np.float64)
    out = np.empty_like(x)
    small = x < 4.0
    xs = x[small]
    term = np.ones_like(xs)
    for k in range(1, l + 1):
        term = term * (xs / (2 * k + 1))
    s = term.copy()
    active = np.ones(xs.shape, dtype=bool)
    for k in range(1, 80):
        term = np.where(active, term * (-xs * xs / (2 * k * (2 * l + 2 * k + 1))), term)
        s = np.where(active, s + term, s)
        active &= ~(np.abs(term) <= np.finfo(np.float64).eps * np.abs(s))
        if not active.any():
            break
    out[small] = s
    xl = x[~small]
    sin, cos = np.sin(xl), np.cos(xl)
    prev = sin / xl
    if l == 0:
        out[~small] = prev
        return out
    cur = (prev - cos) / xl
    for k in range(1, l):
        nxt = (2 * k + 1) / xl * cur - prev
        prev, cur = cur, nxt
    out[~small] = cur
    return out


def real_ylm(l, m, direction):
    """Real normalised harmonics with Condon-Shortley and the extra (-1)^m
    (radial.rs:121-164). ``direction`` is (k, 3); zero rows are invalid."""
    d = np.asarray(direction, dtype=np.float64)
    norm = np.hypot(np.hypot(d[:, 0], d[:, 1]), d[:, 2])
    z = np.clip(d[:, 2] / norm, -1.0, 1.0)
    phi = np.arctan2(d[:, 1], d[:, 0])
    a = abs(m)
    p = np.ones_like(z)
    for k in range(1, a + 1):
        p = p * (-(2 * k - 1) * np.sqrt((1.0 - z) * (1.0 + z)))
    if l > a:
        prev = p
        p = p * ((2 * a + 1) * z)
        for k in range(a + 2, l + 1):
            nxt = ((2 * k - 1) * z * p - (k + a - 1) * prev) / (k - a)
            prev, p = p, nxt
    ratio = 1.0
    for k in range(l - a + 1, l + a + 1):
        ratio /= k
    nrm = math.sqrt((2 * l + 1) / (4.0 * math.pi) * ratio)
    if m == 0:
        ang = 1.0
    else:
        sign = 1.0 if a % 2 == 0 else -1.0
        ang = math.sqrt(2.0) * sign * (np.sin(a * phi) if m < 0 else np.cos(a * phi))
    return nrm * p * ang


# --------------------------------------------------------------------------
# Synthetic pseudopotential (crates/dft-engine/tests/common/mod.rs:146-209)
# --------------------------------------------------------------------------
@dataclass
class Projector:
    l: int
    rbeta: np.ndarray  # r * beta(r) samples (PP_BETA), full support


@dataclass
class Pseudo:
    z: float
    r: np.ndarray
    rab: np.ndarray
    local_ha: np.ndarray
    projectors: list
    dij_ha: np.ndarray  # (np, np)
    rc: float = 1.0


def synthetic_pseudo(z, rc, points, nonlocal_=True, betas=None, dij_ry=None):
    """The reference's synthetic Gaussian UPF (common/mod.rs:148-209): log mesh
    r_i = r0*expm1(i dx) to 12 Bohr, local -2 Z erf(r/rc)/r Ry, beta_i = r^(l+1)
    exp(-r^2/(i+1)) for l = 0, 0, 1 and D = [[2,.4,0],[.4,4,0],[0,0,6]] Ry. The
    UPF text round trip ({:.17e}) is exact, so samples are used directly; Ry
    to Ha halves PP_LOCAL and PP_DIJ (crates/dft-upf/src/parser.rs:735,765).

    ``betas``: optional list of (l, width) giving beta = r^(l+1) exp(-r^2/width)
    (the aluminium variant); default is the reference's (0,1),(0,2),(1,3)."""
    end, r0 = 12.0, 0.02
    dx = math.log(1.0 + end / r0) / (points - 1)
    i = np.arange(points, dtype=np.float64)
    r = r0 * np.expm1(i * dx)
    rab = dx * (r + r0)
    from scipy.special import erf
    with np.errstate(divide="ignore", invalid="ignore"):
        local = np.where(r == 0.0, -4.0 * z / (rc * math.sqrt(math.pi)), -2.0 * z * erf(r / rc) / r)
    projs = []
    if nonlocal_:
        if betas is None:
            betas = [(0, 1.0), (0, 2.0), (1, 3.0)]
        for l, width in betas:
            projs.append(Projector(l, r ** (l + 1) * np.exp(-r * r / width)))
        if dij_ry is None:
            dij_ry = np.array([[2.0, 0.4, 0.0], [0.4, 4.0, 0.0], [0.0, 0.0, 6.0]])
        dij = np.asarray(dij_ry, dtype=np.float64) / 2.0
    else:
        dij = np.zeros((0, 0))
    return Pseudo(float(z), r, rab, local / 2.0, projs, dij, rc)


def local_transform_n(pp: Pseudo, r_max=None):
    """LocalTransform::new endpoint (radial.rs:239-255): the last mesh point
    <= min(end, 10)."""
    limit = r_max if r_max is not None else min(pp.r[-1], 10.0)
    return int(np.searchsorted(pp.r, limit, side="right"))


def vloc_g(pp: Pseudo, volume, g_norms):
    """V_loc(|G|) for G>0 (radial.rs:264-280, hamiltonian.rs:256-283):
    4 pi/Omega [ int (r^2 V + Z r erf r) j0(Gr) dr - Z exp(-G^2/4)/G^2 ]."""
    from scipy.special import erf
    n = local_transform_n(pp)
    r, rab, v = pp.r[:n], pp.rab[:n], pp.local_ha[:n]
    short = r * r * v + pp.z * r * erf(r)
    q = np.asarray(g_norms, dtype=np.float64)
    uniq, inv = np.unique(q, return_inverse=True)
    vals = np.empty(len(uniq))
    chunk = 256
    for s in range(0, len(uniq), chunk):
        qq = uniq[s:s + chunk]
        integrand = (short[None, :] * bessel(0, qq[:, None] * r[None, :])) * rab[None, :]
        integ = integrate_rows(integrand)
        vals[s:s + chunk] = 4.0 * math.pi / volume * (integ - pp.z * np.exp(-qq * qq / 4.0) / (qq * qq))
    return vals[inv]


def alpha_z(pp: Pseudo, volume):
    """Coulomb-subtracted G=0 constant (radial.rs:283-287)."""
    from scipy.special import erf
    n = local_transform_n(pp)
    r, rab, v = pp.r[:n], pp.rab[:n], pp.local_ha[:n]
    short = (r * r * v + pp.z * r * erf(r)) * rab
    return 4.0 * math.pi / volume * integrate(short) + math.pi * pp.z / volume


def projector_transform(pp: Pseudo, index, q):
    """4 pi int r u(r) j_l(q r) dr with u = PP_BETA (radial.rs:318-344)."""
    p = pp.projectors[index]
    q = np.asarray(q, dtype=np.float64)
    uniq, inv = np.unique(q, return_inverse=True)
    vals = np.empty(len(uniq))
    chunk = 256
    for s in range(0, len(uniq), chunk):
        qq = uniq[s:s + chunk]
        integrand = ((pp.r[None, :] * p.rbeta[None, :]) * bessel(p.l, qq[:, None] * pp.r[None, :])) * pp.rab[None, :]
        vals[s:s + chunk] = 4.0 * math.pi * integrate_rows(integrand)
    return vals[inv]


# --------------------------------------------------------------------------
# System (one species) and its Hamiltonian ingredients
# --------------------------------------------------------------------------
@dataclass
class Species:
    pseudo: Pseudo
    frac: np.ndarray  # (na, 3) fractional positions


@dataclass
class GammaSystem:
    cell: Cell
    ecut: float
    species: list
    basis: Basis = None
    dims: tuple = None
    density: Basis = None
    half: HalfSphere = None
    fft_index: np.ndarray = None  # flat grid index of each basis G
    projectors: np.ndarray = None  # complex (n, p), basis order
    coupling: list = None  # list of (start, count, D block) in Ha
    orc_p: np.ndarray = None  # real (n, p), ORC rows
    kinetic_orc: np.ndarray = None  # (n,)
    ionic_grid: np.ndarray = None  # V_ion(r) on the full grid, Ha
    electrons: float = 0.0


def positions_cart(cell: Cell, frac):
    return np.array([cell.frac_to_cart(list(f)) for f in frac], dtype=np.float64)


def structure_factor(cart_g, positions):
    """S(G) = sum exp(-i G.tau) (hamiltonian.rs:57-75)."""
    ang = cart_g @ positions.T  # (n, na)
    return np.exp(-1j * ang).sum(axis=1)


def nonlocal_table(basis: Basis, volume, species):
    """NonlocalOperator::new (hamiltonian.rs:885-1035): columns
    (-i)^l Y_lm(q^) beta_i(|q|)/sqrt(Omega) e^{-i q.tau}, atom-major, channels
    ordered by projector then m=-l..l; per-atom coupling D over equal (l, m).
    Momenta q = k+G."""
    q = basis.cart + basis.kcart[None, :]
    q2 = basis.q2
    qn = np.sqrt(q2)
    cols = []
    coupling = []
    start = 0
    for sp in species:
        pp = sp.pseudo
        npj = len(pp.projectors)
        channels = [(i, p.l, m) for i, p in enumerate(pp.projectors) for m in range(-p.l, p.l + 1)]
        radial = [projector_transform(pp, i, qn) for i in range(npj)]
        direction = q.copy()
        direction[q2 == 0.0] = np.array([0.0, 0.0, 1.0])
        table = []
        for (i, l, m) in channels:
            table.append(((-1j) ** l) * real_ylm(l, m, direction) * radial[i] / math.sqrt(volume))
        nc = len(channels)
        dblock = np.zeros((nc, nc))
        for a, (i, l, m) in enumerate(channels):
            for b, (j, ll, mm) in enumerate(channels):
                if l == ll and m == mm:
                    dblock[a, b] = pp.dij_ha[i, j]
        cell_pos = sp.cart
        for tau in cell_pos:
            ph = np.exp(-1j * (q @ tau))
            for col in table:
                cols.append(col * ph)
            coupling.append((start, nc, dblock.copy()))
            start += nc
    if cols:
        return np.stack(cols, axis=1), coupling
    return np.zeros((len(q2), 0), dtype=np.complex128), coupling


def ionic_coefficients(density: Basis, volume, species):
    """ionic_local (hamiltonian.rs:211-340): V_ion(G) on the density sphere,
    alpha_Z at G=0, times the structure factor."""
    out = np.zeros(len(density.g2), dtype=np.complex128)
    zero = np.all(density.miller == 0, axis=1)
    gn = np.sqrt(density.g2)
    for sp in species:
        pp = sp.pseudo
        v = np.empty(len(gn))
        v[~zero] = vloc_g(pp, volume, gn[~zero])
        v[zero] = alpha_z(pp, volume)
        out += v * structure_factor(density.cart, sp.cart)
    return out


def to_grid(coeffs, miller, dims):
    """Unscaled inverse transform f(r) = sum_G f(G) e^{iG.r} of sphere
    coefficients (crates/dft-numerics/src/lib.rs:46-56); real part."""
    grid = np.zeros(dims, dtype=np.complex128)
    idx = map_miller(miller, dims)
    grid[idx[:, 0], idx[:, 1], idx[:, 2]] = coeffs
    return np.fft.ifftn(grid) * (dims[0] * dims[1] * dims[2])


def to_reciprocal_full(field_r):
    """f(G) = (1/N) sum_r f(r) e^{-iG.r} on the full grid."""
    return np.fft.fftn(field_r) / field_r.size


def build_gamma_system(cell_rows, atoms_frac, pseudo, ecut, grid_override=None):
    """Set-up of refstart.rs:536-600 for one species at Gamma: basis, density
    sphere, FFT grid, ORC, projector table (complex and ORC rows,
    distributed_real_gamma.rs:363-402), kinetic diagonal and V_ion on the grid."""
    cell = Cell.from_rows(cell_rows)
    sp = Species(pseudo, np.asarray(atoms_frac, dtype=np.float64))
    sp.cart = positions_cart(cell, sp.frac)
    basis = plane_wave_basis(cell, ecut)
    dims, dens = grid_dims(cell, ecut)
    if grid_override is not None:
        dims = tuple(grid_override)
    half = half_sphere(basis)
    idx = map_miller(basis.miller, dims)
    fft_index = flat_index(idx, dims)
    proj, coupling = nonlocal_table(basis, cell.volume, [sp])
    orc_p = orc_pack(half, proj) if proj.shape[1] else np.zeros((len(basis.g2), 0))
    kin_full = 0.5 * basis.g2
    kinetic_orc = np.empty(len(basis.g2))
    kinetic_orc[0] = kin_full[0]
    kinetic_orc[1::2] = kin_full[half.plus[1:]]
    kinetic_orc[2::2] = kin_full[half.plus[1:]]
    vion = to_grid(ionic_coefficients(dens, cell.volume, [sp]), dens.miller, dims).real
    return GammaSystem(cell=cell, ecut=ecut, species=[sp], basis=basis, dims=dims, density=dens,
                       half=half, fft_index=fft_index, projectors=proj, coupling=coupling,
                       orc_p=orc_p, kinetic_orc=kinetic_orc, ionic_grid=vion,
                       electrons=pseudo.z * len(sp.frac))


# --------------------------------------------------------------------------
# Ewald (crates/dft-physics/src/ewald.rs): point ions in a neutralising
# background. The value is independent of eta; eta = sqrt(pi)/Omega^(1/3)
# (the reference scheme, ewald.rs:4,380) with erfc(eta r).
# --------------------------------------------------------------------------
def ewald(cell: Cell, charges, positions_cart_, eta=None, tol=1e-17):
    from scipy.special import erfc
    vol = cell.volume
    if eta is None:
        eta = math.sqrt(math.pi) / vol ** (1.0 / 3.0)
    z = np.asarray(charges, dtype=np.float64)
    tau = np.asarray(positions_cart_, dtype=np.float64)
    a = np.array(cell.lattice)  # columns
    b = np.array(cell.reciprocal())
    # real-space cutoff: erfc(eta R) < tol
    rcut = 1.0
    while erfc(eta * rcut) > tol * 1e-3:
        rcut *= 1.2
    gcut = 1.0
    while math.exp(-gcut * gcut / (4 * eta * eta)) > tol * 1e-3:
        gcut *= 1.2
    # integer ranges from the dual norms
    ainv_len = [np.linalg.norm(b[:, i]) / TAU for i in range(3)]
    nr = [int(math.ceil(rcut * ainv_len[i])) + 1 for i in range(3)]
    alen = [np.linalg.norm(a[:, i]) for i in range(3)]
    ng = [int(math.ceil(gcut * alen[i] / TAU)) + 1 for i in range(3)]
    rng = [np.arange(-n, n + 1) for n in nr]
    L = np.stack(np.meshgrid(*rng, indexing="ij"), -1).reshape(-1, 3).astype(np.float64) @ a.T
    real = 0.0
    for i in range(len(z)):
        for j in range(len(z)):
            d = tau[i] - tau[j] + L
            r = np.linalg.norm(d, axis=1)
            if i == j:
                r = r[r > 0]
            r = r[r < rcut]
            real += 0.5 * z[i] * z[j] * np.sum(erfc(eta * r) / r)
    rg = [np.arange(-n, n + 1) for n in ng]
    G = np.stack(np.meshgrid(*rg, indexing="ij"), -1).reshape(-1, 3).astype(np.float64) @ b.T
    g2 = np.sum(G * G, axis=1)
    G, g2 = G[g2 > 0], g2[g2 > 0]
    sg = np.exp(1j * (G @ tau.T)) @ z
    recip = 2.0 * math.pi / vol * np.sum(np.abs(sg) ** 2 * np.exp(-g2 / (4 * eta * eta)) / g2)
    self_e = -eta / math.sqrt(math.pi) * np.sum(z * z)
    background = -math.pi * np.sum(z) ** 2 / (2.0 * vol * eta * eta)
    return dict(energy=real + recip + self_e + background, real=real, reciprocal=recip,
                self_energy=self_e, background=background, eta=eta)
