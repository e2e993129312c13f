"""The pieces that carry the path to the Si512 cell and the SCF glue, on
MI355X through the C ABI:

* the split plane pass (planes that do not fit one workgroup's LDS go
  through HBM): forced on small cells it gives the bits of the one-workgroup
  plane kernel for H·X and the band density, and matches the oracle; on the
  Si512 cell (144^3, picked automatically) H·X matches the oracle's C++
  restatement (oracle/cpu_hx.cpp, pinned to the NumPy oracle on small cells);
* the tridiagonalisation across the chip (large subspaces, up to Si512's
  2252) against LAPACK, also forced on small and degenerate matrices;
* the warm start (KeptHx): H_old X + dV X from the kept real-space fields
  equals a fresh H_new X to 1e-12;
* the SCF field kernels (grid transforms, V_ion + V_H + V_xc, residual and
  metric, mixing products and update, energy sums) against the torch
  reference of the same formulas (tests/torch_fields.py), and the field
  mixer's update against the torch mixer's over a sequence of residuals.
"""
import numpy as np
import pytest

from conftest import fcc_rows

pytestmark = pytest.mark.gpu

SI2 = (fcc_rows(10.26), [[0, 0, 0], [0.25] * 3])


def _op(system, capacity, precision=None, pot_seed=7):
    import torch
    from baandi_amd import lib as L
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    op = DistributedRealHamiltonian(system, capacity=capacity, precision=L.BD_FP64 if precision is None else precision)
    rng = np.random.default_rng(pot_seed)
    pot = rng.standard_normal(system.dims) * 0.1 - 0.3
    op.set_owned_potential(torch.from_numpy(pot.reshape(-1)).cuda())
    return op, pot


def _apply(op, x):
    import torch
    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    yd = torch.full_like(xd, float("nan"))
    op.apply_block(xd, yd)
    torch.cuda.synchronize()
    return yd


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_split_plane_pass_equals_fused(precision):
    """Si2 at 20 Ry (24^3, Z = 11) and Si8 at 6 Ha: the split pass with 1, 3
    and Z z-lines per batch against the one-workgroup plane kernel, bit for
    bit, for H·X and for the band density."""
    import torch
    from baandi_amd import lib as L
    from baandi_amd.hamiltonian import GammaSystem
    prec = L.BD_FP64 if precision == "fp64" else L.BD_MIXED
    for rows, atoms, ecut in (SI2 + (10.0,), ([[10.26, 0, 0], [0, 10.26, 0], [0, 0, 10.26]],
                                              [[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0], [.25, .25, .25],
                                               [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]], 6.0)):
        s = GammaSystem(rows, atoms, 4.0, ecut)
        op, pot = _op(s, 9, prec)
        assert op.plane_split() == 0
        Z = op.plan_info()["z_planes"]
        x = np.random.default_rng(3).standard_normal((9, s.n))
        y0 = _apply(op, x)
        xd = torch.from_numpy(x / np.linalg.norm(x, axis=1, keepdims=True)).cuda()
        w = np.linspace(2.0, 0.5, 9)
        rho0 = torch.empty(int(np.prod(s.dims)), dtype=torch.float64, device="cuda")
        op.band_density_sum(xd, w, rho0)
        for zb in (1, 3, Z):
            op.set_plane_split(zb)
            assert op.plane_split() == zb
            y1 = _apply(op, x)
            rho1 = torch.empty_like(rho0)
            op.band_density_sum(xd, w, rho1)
            torch.cuda.synchronize()
            print(f"[split] {precision} n={s.n} dims={s.dims} Z={Z} zb={zb}: H·X bitwise {torch.equal(y0, y1)}, "
                  f"density bitwise {torch.equal(rho0, rho1)}", flush=True)
            assert torch.equal(y0, y1)
            assert torch.equal(rho0, rho1)
        op.set_plane_split(0)
        assert torch.equal(_apply(op, x), y0)


def test_split_plane_pass_matches_oracle():
    from baandi_amd.hamiltonian import GammaSystem
    from oracle import physics as P
    from oracle.hamiltonian import apply_orc
    s = GammaSystem(*SI2, 4.0, 10.0)
    o = P.build_gamma_system(*SI2, P.synthetic_pseudo(4.0, 1.0, 801), 10.0)
    op, pot = _op(s, 7)
    op.set_plane_split(2)
    x = np.random.default_rng(4).standard_normal((7, s.n))
    ref = apply_orc(o, x.T, pot).T
    err = np.abs(_apply(op, x).cpu().numpy() - ref).max() / np.abs(ref).max()
    print(f"[split] oracle parity {err:.2e}", flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("cell", ["si8", "al4", "si64"])
def test_device_setup_matches_host(cell):
    """bd_system_create_dev (the set-up's atom loops on the GPU) against the
    host set-up: projector rows, Davidson diagonal and V_ion(G) to 1e-13
    relative, every other array equal."""
    from baandi_amd import workloads as W
    wl = {"si8": W.si8(ecut_ha=6.0), "al4": W.al4(), "si64": W.si64()}[cell]
    host, dev = W.build(wl), W.build(wl, device=True)
    for name in ("projectors_orc", "diagonal_static"):
        a, b = host.get(name), dev.get(name)
        err = np.abs(a - b).max() / np.abs(a).max()
        print(f"[device set-up] {cell} {name}: max relative difference {err:.2e}", flush=True)
        assert err < 1e-13, name
    # V_ion(G) as a complex array (the fcc Al cell's is real up to rounding).
    a = host.get("ionic_density_re") + 1j * host.get("ionic_density_im")
    b = dev.get("ionic_density_re") + 1j * dev.get("ionic_density_im")
    err = np.abs(a - b).max() / np.abs(a).max()
    print(f"[device set-up] {cell} V_ion(G): max relative difference {err:.2e}", flush=True)
    assert err < 1e-13
    for name in ("miller", "kinetic_orc", "coupling", "plus_index", "density_index", "ewald"):
        assert np.array_equal(host.get(name), dev.get(name)), name


def test_si512_hx_matches_cpu_restatement():
    """The Si512 cell (4x4x4 cells, 30 Ry, 144^3 grid, 2560 projectors): its
    144 x Z planes exceed one workgroup's LDS in FP64, so the op runs the
    split pass; H·X on 4 columns against the oracle's C++ restatement."""
    import time
    from baandi_amd import workloads as W
    from oracle.cpu_hx import CpuHamiltonian
    t = time.perf_counter()
    s = W.build(W.si512(), device=True)
    print(f"[si512] set-up {time.perf_counter() - t:.1f}s n={s.n} dims={s.dims} nproj={s.nproj}", flush=True)
    assert tuple(s.dims) == (144, 144, 144) and s.nproj == 2560
    op, pot = _op(s, 4)
    plan = op.plan_info()
    print(f"[si512] plan={plan} split={op.plane_split()}", flush=True)
    assert op.plane_split() > 0
    x = np.random.default_rng(5).standard_normal((4, s.n))
    y = _apply(op, x).cpu().numpy()
    cpu = CpuHamiltonian(s)
    cpu.set_potential(pot)
    ref = cpu.apply(x.T).T
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print(f"[si512] H·X vs CPU restatement: max relative error {err:.2e}", flush=True)
    assert err < 1e-12


def _eigen(A, want):
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


def _check(A, want, label):
    w, Y = _eigen(A, want)
    ref = np.linalg.eigvalsh(A)[:want]
    scale = max(np.abs(A).sum(axis=0).max(), 1e-300)
    verr = np.abs(w - ref).max() / scale
    res = np.abs(A @ Y - Y * w[None, :]).max() / scale
    orth = np.abs(Y.T @ Y - np.eye(want)).max()
    print(f"[eigen multi] {label} m={A.shape[0]} want={want}: value {verr:.2e}, residual {res:.2e}, "
          f"orthogonality {orth:.2e}", flush=True)
    m = A.shape[0]
    assert verr < 1e-13 * max(m, 10)
    assert res < 1e-12 * max(m, 10)
    assert orth < 1e-12 * max(m, 10)


@pytest.fixture
def multi_from():
    """Read and restore the size from which the tridiagonalisation runs across the chip."""
    import ctypes as C
    from baandi_amd import lib as L
    lib = L.load()
    old = C.c_int32(0)
    L.check(lib.bd_sym_eigen_multi_from(C.byref(old), 0))

    def set_(v):
        L.check(lib.bd_sym_eigen_multi_from(C.byref(C.c_int32(v)), 1))
    yield set_
    set_(old.value)


@pytest.mark.parametrize("m,want", [(1026, 60), (1500, 80), (2252, 1126)])
def test_eigen_large_subspaces(m, want):
    rng = np.random.default_rng(m)
    B = rng.standard_normal((m, m))
    _check(0.5 * (B + B.T), want, "random")


def test_eigen_large_realified_pairs():
    rng = np.random.default_rng(600)
    B = rng.standard_normal((600, 600)) + 1j * rng.standard_normal((600, 600))
    A = 0.5 * (B + B.conj().T)
    _check(np.block([[A.real, -A.imag], [A.imag, A.real]]), 100, "realified")


@pytest.mark.parametrize("m,want", [(3, 3), (17, 17), (64, 40), (320, 160), (577, 288)])
def test_eigen_multi_forced_small(multi_from, m, want):
    """The across-the-chip tridiagonalisation forced on small subspaces
    (panels of 16 reflectors: a partial last panel at 17, 577): random and
    degenerate spectra."""
    multi_from(3)
    rng = np.random.default_rng(m + 1)
    B = rng.standard_normal((m, m))
    _check(0.5 * (B + B.T), want, "random")
    if m >= 17:
        vals = np.sort(rng.uniform(-1.0, 3.0, m))
        vals[3:6] = vals[3]
        vals[8:10] = vals[8]
        Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
        A = (Q * vals[None, :]) @ Q.T
        _check(0.5 * (A + A.T), max(want // 2, 1), "degenerate")


def test_kept_fields_warm_start():
    """Si64: H_old X + dV X from the fields kept by the band density of X,
    against a fresh H_new X (to 1e-12); a different X is refused."""
    import torch
    from baandi_amd import workloads as W, lib as L
    s = W.build(W.si64())
    op, pot = _op(s, 160)
    op.keep_fields(True)
    x = np.random.default_rng(6).standard_normal((160, s.n))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    xd = torch.from_numpy(x).cuda()
    rho = torch.empty(int(np.prod(s.dims)), dtype=torch.float64, device="cuda")
    op.band_density_sum(xd, np.full(160, 2.0), rho)
    hx_old = torch.empty_like(xd)
    op.apply_block(xd, hx_old)
    rng = np.random.default_rng(8)
    new = pot + 0.02 * rng.standard_normal(s.dims)
    dv = torch.from_numpy((new - pot).reshape(-1)).cuda()
    op.set_owned_potential(torch.from_numpy(new.reshape(-1)).cuda())
    warm = hx_old.clone()
    op.apply_kept(xd, dv, warm)
    fresh = torch.empty_like(xd)
    op.apply_block(xd, fresh)
    torch.cuda.synchronize()
    err = ((warm - fresh).abs().max() / fresh.abs().max()).item()
    print(f"[kept] Si64 160 columns: warm start vs fresh H_new X, max relative difference {err:.2e}", flush=True)
    assert err < 1e-12
    other = xd.clone()
    with pytest.raises(L.BaandiError):
        op.apply_kept(other, dv, warm)


def _field_pair(kat1):
    from baandi_amd.hamiltonian import GammaSystem
    from baandi_amd.field import DensityField
    from torch_fields import DensityField as TorchField
    s = GammaSystem(fcc_rows(kat1["cell_a_bohr"]), [[0, 0, 0], [0.25] * 3], 4.0, kat1["ecut_ha"])
    return s, DensityField(s), TorchField(s)


def test_field_kernels_match_torch_reference(kat1):
    import math
    import torch
    from torch_fields import lda_pz_torch, Mixer
    s, f, t = _field_pair(kat1)
    rng = np.random.default_rng(9)
    nd = s.ndensity
    x = torch.from_numpy(rng.standard_normal(nd) + 1j * rng.standard_normal(nd)).cuda() * 1e-3
    x[torch.from_numpy(t.zero_np).cuda()] = s.electrons / s.volume
    # Grid transforms.
    g_ref = t.to_grid(x)
    g = f.to_grid(x)
    e1 = ((g - g_ref).abs().max() / g_ref.abs().max()).item()
    field = torch.from_numpy(rng.standard_normal(s.dims)).cuda()
    c_ref = t.from_grid(field)
    c = f.from_grid(field)
    e2 = ((c - c_ref).abs().max() / c_ref.abs().max()).item()
    # Potential of an input density.
    vion = torch.from_numpy(rng.standard_normal(s.dims) * 0.1).cuda()
    both = t.to_grid(x + 1j * (t.coulomb * x))
    eps, vxc = lda_pz_torch(both.real.abs())
    v_ref = vion + both.imag + vxc
    v = torch.empty_like(vion)
    vmean = f.potential(x, vion, v)
    e3 = ((v - v_ref).abs().max() / v_ref.abs().max()).item()
    assert abs(vmean - float(v_ref.mean())) < 1e-12 * abs(float(v_ref.mean())) + 1e-15
    # Residual and metric norm.
    rho_out = torch.from_numpy(np.abs(rng.standard_normal(s.dims)) * 1e-2).cuda()
    y = t.from_grid(rho_out)
    y[t.zero] = s.electrons / s.volume
    fr_ref = y - x
    fr = torch.empty_like(x)
    dr2 = f.residual(rho_out, x, fr)
    e4 = ((fr - fr_ref).abs().max() / fr_ref.abs().max()).item()
    assert abs(dr2 - t.dot(fr_ref, fr_ref)) < 1e-12 * t.dot(fr_ref, fr_ref)
    # Mixing products and update.
    rows = torch.stack([fr_ref * (1.0 + 0.1 * k) + 1e-4 * k for k in range(4)])
    a_ref, c_ref2 = t.gram(rows, fr_ref)
    a, cc = f.gram(rows, fr_ref)
    assert np.abs(a - a_ref).max() < 1e-12 * np.abs(a_ref).max()
    assert np.abs(cc - c_ref2).max() < 1e-12 * np.abs(c_ref2).max()
    gamma = np.array([0.3, -0.2, 0.1, 0.05])
    dX = torch.stack([x * (0.01 * k) for k in range(1, 5)])
    for screening in (False, True):
        m = Mixer(t, 0.7, 8, "thomas_fermi" if screening else "plain", s.volume, s.electrons)
        g_t = torch.from_numpy(gamma).cuda()
        xx = x - (g_t[:, None] * dX).sum(0)
        ff = fr_ref - (g_t[:, None] * rows).sum(0)
        if screening:
            ff = ff * m.screening
        ref = xx + 0.7 * ff
        out = torch.empty_like(x)
        f.mix(dX.contiguous(), rows.contiguous(), gamma, x, fr_ref, 0.7, screening, out)
        e5 = ((out - ref).abs().max() / ref.abs().max()).item()
        assert e5 < 1e-13, (screening, e5)
    # Energy sums.
    yo = t.from_grid(rho_out)
    vh = t.to_grid(t.coulomb * yo).real
    dv = s.volume / t.size
    hartree = 0.5 * dv * float((rho_out * vh).sum())
    eps_o, _ = lda_pz_torch(rho_out.abs())
    xc = dv * float((rho_out.abs() * eps_o).sum())
    local = dv * float((rho_out * vion).sum())
    h2, x2, l2 = f.energies(rho_out, vion)
    print(f"[field] to_grid {e1:.1e} from_grid {e2:.1e} potential {e3:.1e} residual {e4:.1e} "
          f"energies {abs(h2 - hartree):.1e} {abs(x2 - xc):.1e} {abs(l2 - local):.1e}", flush=True)
    assert max(e1, e2, e3, e4) < 1e-13
    for got, want in ((h2, hartree), (x2, xc), (l2, local)):
        assert abs(got - want) < 1e-12 * max(abs(want), 1e-300) + 1e-15
    assert math.isfinite(dr2)


def test_field_mixer_matches_torch_mixer(kat1):
    """The field mixer (Gram products and update in the field kernels, the
    small eigensolve on the host) against the torch mixer over a sequence of
    residuals: history 3 so the oldest pairs fall out, both preconditionings,
    and a repeated (x, f) whose zero difference pair makes the history rank
    deficient (both mixers then drop pairs down to plain mixing)."""
    import torch
    from baandi_amd.field import Mixer as FieldMixer
    from torch_fields import Mixer
    s, f, t = _field_pair(kat1)
    rng = np.random.default_rng(11)
    nd = s.ndensity
    zero = torch.from_numpy(t.zero_np).cuda()
    for pre in ("plain", "thomas_fermi"):
        a = FieldMixer(f, 0.7, 3, pre)
        b = Mixer(t, 0.7, 3, pre, s.volume, s.electrons)
        x = torch.from_numpy(rng.standard_normal(nd) + 1j * rng.standard_normal(nd)).cuda() * 1e-3
        x[zero] = s.electrons / s.volume
        worst = 0.0
        for step in range(7):
            fr = torch.from_numpy(rng.standard_normal(nd) + 1j * rng.standard_normal(nd)).cuda() * 1e-3 / (step + 1)
            fr[zero] = 0.0
            if step == 5:
                xa = a.update(x, fr)
                xb = b.update(x, fr)
            xa = a.update(x, fr)
            xb = b.update(x, fr)
            worst = max(worst, ((xa - xb).abs().max() / xb.abs().max()).item())
            x = xb
        print(f"[mixer] {pre}: field mixer vs torch mixer, max relative difference {worst:.1e}", flush=True)
        assert worst < 1e-11
        assert len(a.pairs) == len(b.pairs)
    with pytest.raises(ValueError):
        FieldMixer(f, 0.7, 10)
