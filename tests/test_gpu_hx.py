"""GPU parity of H·X, band density and nonlocal energies against the FP64
oracle, called through the C ABI. Tolerance: 1e-12 relative (FP64 FFT
rounding; the reference's own dense-vs-FFT test uses 1e-11)."""
import numpy as np
import pytest

from conftest import fcc_rows

pytestmark = pytest.mark.gpu

CASES = {
    # name: (rows, atoms, z, ecut, columns[, grid override])
    "kat1_si2_5ha": (fcc_rows(10.261212856717933), [[0, 0, 0], [0.25] * 3], 4.0, 5.0, 8),
    "si2_20ry": (fcc_rows(10.26), [[0, 0, 0], [0.25] * 3], 4.0, 10.0, 9),
    "kat2_skew": ([[5.0, 0, 0], [0.2, 5.0, 0], [0, 0.1, 5.0]], [[0.13, 0.21, 0.32]], 2.0, 1.0, 4),
    "si8_cubic_15ha": ([[10.26, 0, 0], [0, 10.26, 0], [0, 0, 10.26]],
                       [[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                        [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]], 4.0, 15.0, 33),
    # FFT lengths through the radix-8 and radix-9 passes: 16 = 8*2, 27 = 9*3,
    # 36 = 9*4, and 72 = 9*8 on every axis.
    "kat1_grid_16_27_36": (fcc_rows(10.261212856717933), [[0, 0, 0], [0.25] * 3], 4.0, 5.0, 5, (16, 27, 36)),
    "kat1_grid_72": (fcc_rows(10.261212856717933), [[0, 0, 0], [0.25] * 3], 4.0, 5.0, 4, (72, 72, 72)),
}


def build(case):
    import torch
    from baandi_amd.hamiltonian import GammaSystem, DistributedRealHamiltonian
    from oracle import physics as P
    rows, atoms, z, ecut, m = CASES[case][:5]
    grid = CASES[case][5] if len(CASES[case]) > 5 else None
    points = 161 if z == 2.0 else 801
    sysg = GammaSystem(rows, atoms, z, ecut, mesh_points=points, grid_override=grid)
    orc = P.build_gamma_system(rows, atoms, P.synthetic_pseudo(z, 1.0, points), ecut, grid_override=grid)
    if grid is not None:
        assert tuple(sysg.dims) == tuple(grid) == tuple(orc.dims)
    assert np.array_equal(sysg.get("miller"), orc.basis.miller)
    op = DistributedRealHamiltonian(sysg, capacity=m)
    rng = np.random.default_rng(7)
    pot = orc.ionic_grid + 0.1 * rng.standard_normal(orc.dims)
    op.set_owned_potential(torch.from_numpy(pot.reshape(-1)).cuda())
    return torch, sysg, orc, op, pot, m


@pytest.mark.parametrize("case", list(CASES))
def test_apply_matches_oracle(case):
    from oracle.hamiltonian import apply_orc
    torch, sysg, orc, op, pot, m = build(case)
    print(f"[hx parity] case={case} n={sysg.n} dims={sysg.dims} m={m}", flush=True)
    x = np.random.default_rng(11).standard_normal((m, sysg.n))
    xd = torch.from_numpy(x).cuda()
    yd = torch.full_like(xd, float("nan"))
    op.apply_block(xd, yd)
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    ref = apply_orc(orc, x.T, pot).T
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err
    # Bitwise determinism run to run.
    yd2 = torch.empty_like(xd)
    op.apply_block(xd, yd2)
    torch.cuda.synchronize()
    assert torch.equal(yd, yd2)


def test_partial_blocks_and_subsets():
    """Every column count up to capacity (odd and even pair tails)."""
    from oracle.hamiltonian import apply_orc
    torch, sysg, orc, op, pot, m = build("si2_20ry")
    x = np.random.default_rng(3).standard_normal((m, sysg.n))
    ref = apply_orc(orc, x.T, pot).T
    for k in range(1, m + 1):
        xd = torch.from_numpy(np.ascontiguousarray(x[:k])).cuda()
        yd = torch.empty_like(xd)
        op.apply_block(xd, yd)
        torch.cuda.synchronize()
        assert np.abs(yd.cpu().numpy() - ref[:k]).max() < 1e-12 * np.abs(ref).max()


def test_band_density_and_nonlocal_energies():
    from oracle import fields as F
    torch, sysg, orc, op, pot, m = build("si2_20ry")
    x = np.random.default_rng(5).standard_normal((m, sysg.n))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    w = np.linspace(2.0, 0.5, m)
    xd = torch.from_numpy(x).cuda()
    rho = torch.empty(int(np.prod(sysg.dims)), dtype=torch.float64, device="cuda")
    op.band_density_sum(xd, w, rho)
    torch.cuda.synchronize()
    ref = F.orc_density(orc, x.T, w).reshape(-1)
    assert np.abs(rho.cpu().numpy() - ref).max() < 1e-12 * np.abs(ref).max()
    e = op.nonlocal_band_energies(xd)
    assert np.allclose(e, F.nonlocal_energies_orc(orc, x.T), rtol=1e-12, atol=1e-14)


def test_host_entry_point_matches_device():
    torch, sysg, orc, op, pot, m = build("kat2_skew")
    x = np.random.default_rng(1).standard_normal((sysg.n, m))
    yh = op.apply_block_host(x)
    xd = torch.from_numpy(np.ascontiguousarray(x.T)).cuda()
    yd = torch.empty_like(xd)
    op.apply_block(xd, yd)
    torch.cuda.synchronize()
    assert np.array_equal(yh, yd.cpu().numpy().T)
