"""Gamma SCF driver around the GPU H·X operator and Davidson (the caller of
the hot path: crates/dft-engine/src/refscf/driver.rs:479-757 and map.rs:200-384).

Per map: Hartree + LDA-PZ fields of the input density, V = V_ion + V_H + V_xc
installed on the operator, the reference Davidson on the GPU with the ethr
schedule (refscf/profile.rs:49-51) and the iteration-1 redo, the output
density from the GPU band-density transform, the half-sphere Hartree metric
dr2 (refscf/metric.rs) and difference-history Anderson mixing with beta 0.7,
history 8 (refscf/broyden.rs:317-445). Stops when dr2 < threshold_ry.

The density-sphere field step (grid transforms of the density and the
Hartree potential, pointwise LDA-PZ, the residual and its metric norm, the
mixer's Gram products and update, the energy sums) runs in the field kernels
(baandi_amd.field, bd_field_*), so that nothing but scalars and the small
mixing matrix crosses to the host per map. With ``kept_hx`` (on by default,
as in driver.rs:493) each map's Davidson starts from H_old X + dV X, built
from the real-space pair fields the previous band density kept
(refscf/kept_hx.rs, map.rs:267-271 and :317-327), instead of a fresh H·X of
the start block.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from . import lib as L
from .field import DensityField, Mixer, gaussian_start
from .hamiltonian import DistributedRealHamiltonian, GammaSystem
from . import ref_davidson


@dataclass
class ScfOptions:
    bands: int
    threshold_ry: float = 1e-8       # refscf/profile.rs:17-26
    beta: float = 0.7
    history: int = 8
    max_iterations: int = 100
    davidson_max_iterations: int = 20
    preconditioner: int = L.BD_PRECOND_QE
    seed: int = 1
    # BD_FP64: every map in FP64. BD_MIXED: maps whose Davidson threshold is at
    # least mixed_floor_ha run on the mixed-tier operator (FP32 FFT H·X, BF16x3
    # GEMMs, FP64 factor/eigensolve); the later maps, and the final answer, run
    # in FP64, so the fixed point is the FP64 one. BD_FP32_HX: the same with
    # FP64 GEMMs (FP32 FFT H·X only). The floor defaults to the tier's noise:
    # 5e-5 Ha for the BF16x3 Gram, 1e-6 Ha for FP32 H·X (~3e-7 Ha Ritz error).
    precision: int = L.BD_FP64
    mixed_floor_ha: float = None
    mlp: object = None  # a baandi_amd.learned.Network when preconditioner == BD_PRECOND_MLP
    allow_unconverged: bool = False  # return the last map instead of raising at max_iterations
    preconditioning: str = "plain"   # density mixing: "plain" or "thomas_fermi" (refscf/profile.rs:4-7)
    capacity: int = 0                # Davidson subspace capacity (0: 2 bands, ref_davidson.rs:38-78)
    autotune: bool = False           # tune the H·X launch shapes at set-up (bit-neutral)
    plane_shape: tuple = None        # or fix the plane kernel's (threads, z-chunk rows) instead
    # Optional fixed-point verification map (driver.rs:727-757): one more
    # Davidson at the converged potential with every band at this threshold (Ha).
    final_threshold: float = 0.0
    # Fermi-Dirac smearing width (Ha) for cells whose gap is too small for
    # fixed occupations (occupations.rs:71-232); 0 keeps the lowest
    # electrons/2 bands doubly occupied. With smearing every band is held to
    # the Davidson threshold and the energy is the free energy E - TS.
    theta: float = 0.0
    # Warm start of each map's first H·X from the previous map's kept fields
    # (refscf/kept_hx.rs; driver.rs:493 enables it). Ops on the split plane
    # pass keep no fields and take a fresh H·X.
    kept_hx: bool = True


def next_ethr_ry(previous, dr2_ry, electrons):
    """refscf/profile.rs:49-51."""
    return max(min(previous, 0.1 * dr2_ry / max(electrons, 1.0)), 1e-13)


def describe_options(options: ScfOptions):
    """(options, mixing profile) descriptions that a continuation must match."""
    fields = {k: v for k, v in vars(options).items() if k != "mlp"}
    fields["mlp"] = None if options.mlp is None else "network"
    profile = f"beta={options.beta!r} history={options.history} preconditioning={options.preconditioning}"
    return repr(sorted(fields.items())), profile


def run_scf(system: GammaSystem, options: ScfOptions, log=None, start=None, profile_kernels=False,
            stop_after=None, resume=None):
    """SCF to dr2 < threshold_ry; returns a dict with energies (Ha), eigenvalues,
    per-iteration history and timings. ``start``: optional (bands, n) CUDA tensor.

    ``stop_after=k`` ends the run at the boundary after map k and returns
    {"state": ...}, the continuation state (baandi_amd.formats encodes it);
    ``resume=state`` continues from such a state with the same system and
    options, to the same bits as the uninterrupted run."""
    import torch
    dev = torch.device("cuda")
    t_setup = time.perf_counter()
    bands = options.bands
    electrons = system.electrons
    occupied = int(round(electrons)) // 2
    occ = np.zeros(bands)
    occ[:occupied] = 2.0
    smeared = options.theta > 0.0
    converge = bands if smeared else occupied
    mu, ts = None, 0.0
    if smeared:
        from .kpoint import fermi_dirac
        if 2 * bands <= electrons:
            raise L.BaandiError(L.BD_ERR_INVALID_PARAMETER, "smearing needs more bands than electrons/2")
    field = DensityField(system)
    vion_g = torch.from_numpy(system.get("ionic_density_re") + 1j * system.get("ionic_density_im")).to(dev)
    vion = field.to_grid(vion_g).real.contiguous()
    op = DistributedRealHamiltonian(system, capacity=bands, precision=L.BD_FP64)
    if options.precision not in (L.BD_FP64, L.BD_MIXED, L.BD_FP32_HX):
        raise L.BaandiError(L.BD_ERR_INVALID_PARAMETER, "unknown precision mode")
    op_mixed = (DistributedRealHamiltonian(system, capacity=bands, precision=options.precision)
                if options.precision != L.BD_FP64 else None)
    mixed_floor = options.mixed_floor_ha
    if mixed_floor is None:
        mixed_floor = 5e-5 if options.precision == L.BD_MIXED else 1e-6
    if options.plane_shape is not None:
        for o in (op, op_mixed):
            if o is not None:
                o.set_plane_shape(*options.plane_shape)
    elif options.autotune:  # on the ionic potential (the shapes do not depend on V)
        for o in (op, op_mixed):
            if o is not None:
                o.set_owned_potential(vion.reshape(-1).contiguous())
                o.autotune(bands)
    keeping = []  # ops whose band densities keep the pair fields (warm start)
    if options.kept_hx:
        for o in (op, op_mixed):
            if o is not None:
                try:
                    o.keep_fields(True)
                    keeping.append(o)
                except L.BaandiError:  # split plane pass: no kept fields
                    pass
    diag_static = torch.from_numpy(system.get("diagonal_static")).to(dev)
    kinetic = system.get("kinetic_orc")
    ewald = float(system.get("ewald")[0])
    x = torch.from_numpy(gaussian_start(system)).to(dev)
    mixer = Mixer(field, options.beta, options.history, options.preconditioning)
    if start is None:
        g = torch.Generator(device="cpu").manual_seed(options.seed)
        start = torch.randn(bands, system.n, generator=g, dtype=torch.float64)
        start = (start / start.norm(dim=1, keepdim=True)).to(dev)
    orbitals = start.contiguous()
    xout = torch.empty_like(orbitals)
    hxout = torch.empty_like(orbitals)
    hx_prev = torch.empty_like(orbitals)  # H X of `orbitals` from the previous map (warm start)
    rho_dev = torch.empty(field.size, dtype=torch.float64, device=dev)
    pot = torch.empty(field.size, dtype=torch.float64, device=dev)
    pot_prev = torch.empty_like(pot)
    dv = torch.empty_like(pot)
    fres = torch.empty_like(x)
    if profile_kernels:
        op.set_profiling(True)
        if op_mixed is not None:
            op_mixed.set_profiling(True)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup

    options_text, profile_text = describe_options(options)
    t0 = time.perf_counter()
    ethr = 1e-2
    last_dr2 = math.inf
    history = []
    first = 1
    warm = None  # the op whose last band density kept the fields of `orbitals`
    if resume is not None:
        if resume["options"] != options_text or resume["profile"] != profile_text:
            raise L.BaandiError(L.BD_ERR_INVALID_PARAMETER, "continuation from other SCF options")
        x = torch.from_numpy(np.asarray(resume["x"], dtype=np.complex128)).to(dev)
        mixer.pairs = [(torch.from_numpy(dx).to(dev), torch.from_numpy(df).to(dev))
                       for dx, df in resume["mixer_pairs"]]
        mixer.prev = (None if resume["mixer_prev"] is None else
                      tuple(torch.from_numpy(v).to(dev) for v in resume["mixer_prev"]))
        orbitals.copy_(torch.from_numpy(resume["orbitals"]))
        ethr, last_dr2 = resume["ethr"], resume["last_dr2"]
        history = [dict(h) for h in resume["history"]]
        first = resume["completed"] + 1
        entry = resume["warm"]
        w_op = None if entry is None else (op if entry["op"] == "fp64" else op_mixed)
        if w_op is not None and w_op in keeping:
            hx_prev.copy_(torch.from_numpy(entry["hx"]))
            pot_prev.copy_(torch.from_numpy(entry["potential"]))
            # The kept fields are a function of X alone: the band density
            # recomputes them bit for bit.
            w_op.band_density_sum(orbitals, np.zeros(bands), rho_dev)
            warm = w_op
        torch.cuda.synchronize()
    totals = dict(h_columns=0, davidson_iterations=0, dense_s=0.0, apply_s=0.0, davidson_s=0.0, fields_s=0.0,
                  mixed_maps=0, mlp_fallbacks=0, eigen_retries=0, warm_starts=0)
    result = None
    for iteration in range(first, options.max_iterations + 1):
        if iteration > 1:
            ethr = next_ethr_ry(1e-2 if iteration == 2 else ethr, last_dr2, electrons)
        redo = False
        while True:
            tf = time.perf_counter()
            vmean = field.potential(x, vion, pot)  # V_ion + V_H + v_xc of the input density
            use_mixed = op_mixed is not None and ethr / 2.0 >= mixed_floor
            cur = op_mixed if use_mixed else op
            initial_hx = None
            if warm is cur:
                # H_new X = H_old X + (V_new - V_old) X (kept_hx.rs prepare).
                torch.sub(pot, pot_prev, out=dv)
                cur.apply_kept(orbitals, dv, hx_prev)
                initial_hx = hx_prev
                totals["warm_starts"] += 1
            cur.set_owned_potential(pot)
            diag = diag_static + vmean
            totals["fields_s"] += time.perf_counter() - tf
            td = time.perf_counter()
            opts = ref_davidson.Options(bands=bands, threshold=ethr / 2.0, occupied=converge,
                                        first_step=(iteration == 1 and not redo),
                                        max_iterations=options.davidson_max_iterations, return_hx=True,
                                        preconditioner=options.preconditioner, mlp=options.mlp,
                                        mlp_vbar=vmean, capacity=options.capacity)
            sol = ref_davidson.solve_with_initial_hx(cur, orbitals, diag, opts, initial_hx=initial_hx,
                                                     out=(xout, hxout))
            totals["davidson_s"] += time.perf_counter() - td
            totals["mixed_maps"] += int(use_mixed)
            totals["mlp_fallbacks"] += int(sol.record.get("mlp_fallback", -1) >= 0)
            totals["h_columns"] += sol.record["h_columns"]
            totals["eigen_retries"] += sol.record["eigen_retries"]
            totals["davidson_iterations"] += sol.record["iterations"]
            totals["dense_s"] += sol.record["seconds_dense"]
            totals["apply_s"] += sol.record["seconds_apply"]
            tf = time.perf_counter()
            if smeared:
                occ, mu, ts = fermi_dirac(sol.eigenvalues[None, :], np.ones(1), electrons, options.theta)
                occ = occ[0]
            cur.band_density_sum(xout, occ, rho_dev)  # keeps the fields of xout when cur keeps
            rho_out = rho_dev.view(field.dims)
            dr2 = field.residual(rho_out, x, fres)
            totals["fields_s"] += time.perf_counter() - tf
            # xout becomes the next start, hxout its H X, pot its potential.
            orbitals, xout = xout, orbitals
            hx_prev, hxout = hxout, hx_prev
            pot_prev, pot = pot, pot_prev
            warm = cur if cur in keeping else None
            if iteration == 1 and not redo and dr2 < ethr * max(electrons, 1.0):
                ethr = next_ethr_ry(ethr, dr2, electrons)
                redo = True
                continue
            break
        last_dr2 = dr2
        history.append(dict(iteration=iteration, dr2_ry=dr2, ethr_ry=ethr, mixed=use_mixed,
                            davidson_iterations=sol.record["iterations"], h_columns=sol.record["h_columns"],
                            converged=bool(sol.converged.all()), warm_start=initial_hx is not None))
        if log:
            log(f"scf {iteration:3d} dr2={dr2:.3e} ethr={ethr:.2e} dav_it={sol.record['iterations']} "
                f"hcols={sol.record['h_columns']}")
        if dr2 < options.threshold_ry or (options.allow_unconverged and iteration == options.max_iterations):
            result = (sol, cur, vmean)
            break
        x = mixer.update(x, fres)
        if stop_after is not None and iteration >= stop_after:
            torch.cuda.synchronize()
            host = lambda t: t.detach().cpu().numpy().copy()
            state = dict(options=options_text, profile=profile_text, completed=iteration,
                         density=host(rho_dev), x=host(x), mixer_updates=len(history),
                         mixer_prev=None if mixer.prev is None else tuple(host(v) for v in mixer.prev),
                         mixer_pairs=[(host(dx), host(df)) for dx, df in mixer.pairs], epoch=cur.epoch,
                         orbitals=host(orbitals), kept_enabled=bool(keeping), warm=None,
                         ethr=ethr, last_dr2=last_dr2, history=[dict(h) for h in history])
            if warm is not None:
                state["warm"] = dict(x=state["orbitals"], hx=host(hx_prev), potential=host(pot_prev),
                                     epoch=warm.epoch, pairs=(bands + 1) // 2,
                                     op="fp64" if warm is op else "mixed")
            return {"state": state, "history": history, "totals": totals}
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if result is None:
        raise RuntimeError(f"SCF did not reach dr2 < {options.threshold_ry} Ry in {options.max_iterations} maps")
    sol, last_op, vmean = result
    potential = pot_prev  # the converged map's potential (swapped above)
    if last_op is not op:
        op.set_owned_potential(potential)
    if options.final_threshold > 0.0:
        diag = diag_static + vmean
        opts = ref_davidson.Options(bands=bands, threshold=options.final_threshold, occupied=converge,
                                    first_step=True, max_iterations=200, return_hx=True,
                                    preconditioner=options.preconditioner, mlp=options.mlp,
                                    mlp_vbar=vmean)
        sol = ref_davidson.solve_with_initial_hx(op, orbitals, diag, opts, out=(xout, hxout))
        orbitals, xout = xout, orbitals
        if smeared:
            occ, mu, ts = fermi_dirac(sol.eigenvalues[None, :], np.ones(1), electrons, options.theta)
            occ = occ[0]
        op.band_density_sum(orbitals, occ, rho_dev)
    rho_out = rho_dev.view(field.dims)
    # Energy audit of the converged map (refscf/map.rs:466-540).
    xo = orbitals.cpu().numpy()
    kin = float(np.sum(occ[:, None] * xo ** 2 * kinetic[None, :]))
    nonlocal_ = float(np.dot(op.nonlocal_band_energies(orbitals), occ))
    hartree, xc, local = field.energies(rho_out, vion)
    # Delivered-state checks (results.json last_iteration): the largest
    # residual norm |H x - theta x| over the converged bands and the charge
    # error of the output density.
    hx_res = torch.empty_like(orbitals)
    op.apply_block(orbitals, hx_res)
    theta = torch.from_numpy(sol.eigenvalues).to(dev)
    max_residual = float((hx_res - theta[:, None] * orbitals).norm(dim=1)[:converge].max())
    charge = float(rho_out.sum()) * system.volume / field.size
    total = kin + local + nonlocal_ + hartree + xc + ewald - ts
    energy = {"kinetic": kin, "local": local, "nonlocal": nonlocal_, "hartree": hartree, "xc": xc,
              "ewald": ewald, "entropy_term": ts, "total": total, "band_sum": float(np.dot(sol.eigenvalues, occ))}
    out = dict(energy=energy, eigenvalues=sol.eigenvalues.copy(), occupations=occ.copy(), mu=mu,
               history=history, wall_s=wall,
               setup_s=t_setup, iterations=len(history), totals=totals, orbitals=orbitals, op=op,
               plan=op.plan_info(), potential=potential.cpu().numpy(), density=rho_out.cpu().numpy(),
               converged=bool(history[-1]["dr2_ry"] < options.threshold_ry), max_eigen_residual_ha=max_residual,
               relative_charge_error=abs(charge - electrons) / electrons,
               verification_attempts=int(options.final_threshold > 0.0))
    if profile_kernels:
        out["kernel_totals"] = op.stage_totals()
        if op_mixed is not None:
            out["kernel_totals_mixed"] = op_mixed.stage_totals()
    return out
