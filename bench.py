#!/usr/bin/env python3
"""Si64 Gamma SCF wall time on one MI355X (the Davidson H·ψ path end to end).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one full SCF of the Si64 cubic supercell (2x2x2 diamond cells,
a = 20.52 Bohr, Gamma, 30 Ry, 160 bands, 128 occupied, synthetic Z=4
pseudopotential, seeded N(0,1) start) to dr2 < 1e-8 Ry, with all inputs
already resident on the GPU. Prints one JSON line.

The roofline object covers the dominant H·X kernel (k_plane: y/z transforms
with the V(r) multiply), timed with HIP events on the operator's stream over
the timed region. The CPU baseline times the oracle's C++ FP64 H·X (all host
threads, and one) on a bounded sample of the same Si64 columns. Side legs
(untimed by the step clock): the subspace GEMMs, the Al 6x6x6 k-point SCF,
and the Si512 cell's set-up, 1126-column H·X and first Davidson iterations.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The k-point leg runs 16 solves side by side, one HIP stream each: give the
# process as many hardware queues (read once, when the HIP runtime starts).
os.environ["GPU_MAX_HW_QUEUES"] = "16"

HBM_PEAK_GBS = 8000.0
# Measured HBM bytes per pair of k_plane at Si64 FP64 (autotuned 512-thread shape):
# rocprofv3 FETCH_SIZE (x2.000, the 256 MiB copy calibration) + WRITE_SIZE (x1.000)
# of an 80-pair launch, 204066.0 KiB / 80 pairs (profiles/k_plane_si64_fp64.txt).
PLANE_TRAFFIC_BYTES_PER_PAIR = 204066.0 * 1024 / 80
METRIC = "Si64 SCF wall-time to 1e-8 Ha/atom: mixed+learned vs plain FP64, 1 MI355X"


def plane_bytes_model(plan, pairs, c=16, r=8):
    """Algorithmic bytes of the y and z stages of the 3-pass pencil model
    (per pair 4 c (Z n1 n0 + N) + r N for V), for `pairs` pairs."""
    n0, n1, n2, Z = plan["n0"], plan["n1"], plan["n2"], plan["z_planes"]
    N = n0 * n1 * n2
    return pairs * (4 * c * (Z * n1 * n0 + N) + r * N)


def fused_bytes(plan, pairs, c=16, r=8):
    """Compulsory HBM bytes of k_plane as built: sticks in and out, V once."""
    n0, n1, n2 = plan["n0"], plan["n1"], plan["n2"]
    return 2 * c * pairs * n0 * plan["sticks"] + r * n0 * n1 * n2


def cpu_baseline(system, columns=160, reps=3):
    """The oracle's C++ FP64 H·X (OpenMP over band pairs, pruned stick/plane
    FFTs; NumPy BLAS for P D P^T X) on the Si64 set-up: the median of `reps`
    applies to a `columns`-column block on all the host threads this process
    may use, and on one. Returns (ms all threads, ms one thread, threads)."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    from oracle.cpu_hx import CpuHamiltonian
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    h = CpuHamiltonian(system, threads=threads)
    rng = np.random.default_rng(0)
    h.set_potential(rng.standard_normal(system.dims) * 0.1)
    x = np.asfortranarray(rng.standard_normal((system.n, columns)))
    out = {}
    for t in (threads, 1):
        h.set_threads(t)
        times = []
        with threadpool_limits(t):
            h.apply(x)  # warm-up (first-touch of the per-thread grids)
            for _ in range(reps):
                t0 = time.perf_counter()
                h.apply(x)
                times.append(time.perf_counter() - t0)
        out[t] = 1e3 * float(np.median(times))
    return out[threads], out[1], threads


def hx_profile(system, wl, count):
    """Fixed-shape H·X launches for counter runs, then a plain copy kernel of
    known size (the FETCH_SIZE/WRITE_SIZE calibration)."""
    import numpy as np
    import torch
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    bands = wl["bands"]
    op = DistributedRealHamiltonian(system, capacity=bands)
    rng = np.random.default_rng(0)
    op.set_owned_potential(torch.from_numpy(rng.standard_normal(int(np.prod(system.dims))) * 0.1).cuda())
    op.autotune(bands)  # the launch shape the SCF legs use
    x = torch.from_numpy(rng.standard_normal((bands, system.n))).cuda()
    y = torch.empty_like(x)
    op.set_profiling(True)
    for _ in range(count):
        op.apply_block(x, y)
    torch.cuda.synchronize()
    st = op.stage_totals()
    a = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 256 MiB
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    plan = op.plan_info()
    print(json.dumps({"mode": "hx_only", "applies": count, "columns": bands, "plan": plan,
                      "stage_ms_per_apply": {k: v / count for k, v in st.items()
                                             if k not in ("applies", "pairs", "columns")},
                      "copy_bytes_each_way": a.numel() * 8}))


def gemm_leg(reps=10):
    """The Si64 subspace products on the matrix cores: the border Gram
    V^T E (320 x 160 x 23,847, split-K) and the rotation V C (23,847 x 160 x
    320), FP64 MFMA and BF16x3 MFMA, each timed over `reps` back-to-back
    launches between HIP events (bd_gemm_time_dev). Rates: 2MNK / t
    ("effective") and, for BF16x3, the 3 bf16 products per term the MFMA
    units execute."""
    import numpy as np
    import torch
    from baandi_amd import lib as L
    lib = L.load()
    rng = np.random.default_rng(0)
    out = {}
    for name, trans, M, N, K in (("border_gram_tn", 1, 320, 160, 23847), ("rotation_nn", 0, 23847, 160, 320)):
        A = torch.from_numpy(rng.standard_normal(M * K)).cuda()
        B = torch.from_numpy(rng.standard_normal(K * N)).cuda()
        C = torch.zeros(M * N, dtype=torch.float64, device="cuda")
        lda = K if trans else M
        for mode, label in ((0, "fp64"), (1, "bf16x3")):
            work = torch.empty(max(lib.bd_gemm_workspace(mode, trans, M, N, K), 1), dtype=torch.float64,
                               device="cuda")
            t = np.zeros(1)
            L.check(lib.bd_gemm_time_dev(mode, trans, M, N, K, L.ptr(A), lda, L.ptr(B), K, L.ptr(C), M,
                                         L.ptr(work), reps, L.ptr(t)))
            ms = float(t[0])
            tf = 2.0 * M * N * K / (ms * 1e-3) / 1e12
            rec = {"shape": [M, N, K], "ms": round(ms, 4), "effective_TFLOPs": round(tf, 2)}
            if mode == 0:
                rec["frac_fp64_peak"] = round(tf / 78.6, 4)
            else:
                rec["mfma_TFLOPs"] = round(3 * tf, 2)
                rec["frac_bf16_peak"] = round(3 * tf / 2500.0, 4)
            out[f"{name}_{label}"] = rec
    return out


def kpoint_leg(steps, threads=16, variants=True, fused_shape=None):
    """Al fcc 4-atom cell, 6x6x6 Monkhorst-Pack (108 k points), 25 Ry, 16
    bands, Fermi-Dirac 0.01 Ha: k-point SCF wall time with the complex
    Davidsons of all k points in lock step (beside it: the realified Davidson
    per k point on 16 host threads, and the block solver), and one fused
    launch of H over every column of every k point (timed with HIP events on
    the batch stream)."""
    import numpy as np
    import torch
    from baandi_amd import workloads, kpoint as K
    wl = workloads.al4()
    system = workloads.build(wl)
    kp = K.monkhorst_pack(wl["mesh"])
    opts = K.KScfOptions(bands=wl["bands"], theta=wl["theta"], threshold_ry=1e-8, threads=threads,
                         preconditioning="thomas_fermi", autotune=fused_shape is None, fused_shape=fused_shape)
    walls, r = [], None
    for k in range(steps):
        r = K.run_scf_k(system, kp, opts)
        walls.append(r["wall_s"])
        print(f"[bench] al4 k-SCF step {k}: {r['wall_s']:.3f}s maps={r['iterations']} "
              f"E={r['energy']['total']:.12f}", file=sys.stderr, flush=True)
    # The same SCF with the realified Davidson (2 bands real columns per k
    # point), and with the block iterations of all k points stepped together
    # (one fused H·W launch per step over every column of every k point, the
    # step replayed from a captured HIP graph; realified columns).
    from dataclasses import replace
    split = {key: round(r["totals"][key], 4) for key in ("davidson_s", "fields_s")}
    if not variants:
        return {"wall_s": round(sum(walls) / len(walls), 4), "scf_iterations": r["iterations"],
                "h_columns": r["totals"]["h_columns"], "energy_ha": r["energy"]["total"], "split_s": split}
    rr = K.run_scf_k(system, kp, replace(opts, columns="realified"))
    print(f"[bench] al4 k-SCF realified Davidson: {rr['wall_s']:.3f}s maps={rr['iterations']} "
          f"E={rr['energy']['total']:.12f}", file=sys.stderr, flush=True)
    realified = {"wall_s": round(rr["wall_s"], 4), "scf_iterations": rr["iterations"],
                 "h_columns_real": rr["totals"]["h_columns"],
                 "energy_diff_per_atom_ha": abs(rr["energy"]["total"] - r["energy"]["total"]) / 4,
                 "eigenvalue_max_diff_ha": float(np.abs(rr["eigenvalues"] - r["eigenvalues"]).max())}
    rb = K.run_scf_k(system, kp, replace(opts, solver="block", columns="realified"))
    print(f"[bench] al4 k-SCF block solver: {rb['wall_s']:.3f}s maps={rb['iterations']} "
          f"E={rb['energy']['total']:.12f}", file=sys.stderr, flush=True)
    block = {"wall_s": round(rb["wall_s"], 4), "scf_iterations": rb["iterations"],
             "block_steps": rb["totals"]["davidson_iterations"], "h_columns": rb["totals"]["h_columns"],
             "energy_diff_per_atom_ha": abs(rb["energy"]["total"] - r["energy"]["total"]) / 4,
             "eigenvalue_max_diff_ha": float(np.abs(rb["eigenvalues"] - r["eigenvalues"]).max())}
    # The batched launch over the realified run's 32 real columns per k point.
    ops = rr["ops"]
    xs = [o.contiguous() for o in rr["orbitals"]]
    ys = [torch.empty_like(x) for x in xs]
    batch = K.KPointBatch(ops)
    batch.apply(xs, ys)  # warm-up
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(K.L.load().bd_k_batch_stream(batch.handle))
    reps = 5
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        batch.apply(xs, ys)
    b.record(stream)
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    cols = sum(int(x.shape[0]) for x in xs)
    # Compulsory bytes of one batched launch: every column read and written
    # once (16 B per coefficient each way) and V(r) once per launch.
    nbytes = sum(int(x.numel()) * 8 * 2 for x in xs) + 8 * int(np.prod(system.dims))
    return {
        "workload": "Al fcc 4-atom cubic a=7.65 Bohr, 6x6x6 MP (108 k), 25 Ry, 16 bands, FD 0.01 Ha, "
                    "Thomas-Fermi mixing, SCF to dr2<1e-8 Ry",
        "k_points": len(kp), "nk_range": [min(o.nk for o in ops), max(o.nk for o in ops)],
        "wall_s": round(sum(walls) / len(walls), 4), "scf_iterations": r["iterations"], "split_s": split,
        "davidson": "complex Hermitian, 16 complex columns per k point, all 108 k points in lock step",
        "realified_davidson": realified,
        "block_solver": block,
        "setup_s": round(r["setup_s"], 3), "concurrent_solves": threads,
        "energy_ha": r["energy"]["total"], "entropy_term_ha": r["energy"]["entropy_term"], "mu_ha": r["mu"],
        "h_columns": r["totals"]["h_columns"], "fused_shape": list(ops[0].fused_shape()),
        "batched_apply": {"columns": cols, "ms": round(ms, 4), "columns_per_s": round(cols / (ms * 1e-3), 1),
                          "compulsory_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                          "frac_hbm": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
    }


def si128_leg(mlp_path=None, theta=0.01, variants=True):
    """Si128 (4x2x2 conventional cells, displaced N(0, 0.05 Bohr), Gamma,
    25 Ry, 288 bands): SCF to dr2 < 1e-8 Ry in FP64, then (variants) mixed,
    learned and mixed + learned. The cell is nearly gapless at Gamma under
    the synthetic pseudopotential, so every leg uses Fermi-Dirac smearing
    (theta Ha)."""
    import numpy as np
    from baandi_amd import workloads, lib as L
    from baandi_amd.scf import run_scf, ScfOptions
    from baandi_amd.learned import Network
    from baandi_amd.train_preconditioner import DEFAULT_OUT
    wl = workloads.si128()
    system = workloads.build(wl)
    net = Network.load(mlp_path or DEFAULT_OUT)
    base = dict(bands=wl["bands"], threshold_ry=1e-8, preconditioning="thomas_fermi", autotune=True, theta=theta)
    legs = {"fp64": ScfOptions(**base),
            "mixed": ScfOptions(precision=L.BD_MIXED, **base),
            "learned": ScfOptions(preconditioner=L.BD_PRECOND_MLP, mlp=net, **base),
            "mixed_learned": ScfOptions(precision=L.BD_MIXED, preconditioner=L.BD_PRECOND_MLP, mlp=net, **base),
            "fp32_hx": ScfOptions(precision=L.BD_FP32_HX, **base)}
    if not variants:
        legs = {"fp64": legs["fp64"]}
    out, ref = {"workload": f"{wl['name']}: n={system.n}, grid {list(system.dims)}, {system.nproj} projectors, "
                            f"{wl['bands']} bands, Fermi-Dirac {theta} Ha, Thomas-Fermi mixing, dr2<1e-8 Ry"}, None
    for name, opts in legs.items():
        r = run_scf(system, opts)
        ref = ref or r
        print(f"[bench] si128 {name}: {r['wall_s']:.3f}s maps={r['iterations']} E={r['energy']['total']:.12f}",
              file=sys.stderr, flush=True)
        t = r["totals"]
        out[name] = {"wall_s": round(r["wall_s"], 3), "scf_iterations": r["iterations"],
                     "h_columns": t["h_columns"], "davidson_iterations": t["davidson_iterations"],
                     "dense_s": round(t["dense_s"], 3), "apply_s": round(t["apply_s"], 3),
                     "mixed_maps": t["mixed_maps"], "mlp_fallbacks": t["mlp_fallbacks"],
                     "energy_ha": r["energy"]["total"], "entropy_term_ha": r["energy"]["entropy_term"],
                     "energy_diff_per_atom_ha": abs(r["energy"]["total"] - ref["energy"]["total"]) / 128,
                     "eigenvalue_max_diff_ha": float(np.abs(r["eigenvalues"] - ref["eigenvalues"]).max())}
    return out


def si512_leg(iterations=2, reps=2):
    """Si512 (4x4x4 conventional cells, displaced N(0, 0.03 Bohr), Gamma,
    30 Ry, 1126 bands, 1024 occupied) at the potential of the Gaussian start
    density: the set-up with its atom loops on the GPU, the H·X of the full
    1126-column block (split plane pass, 144^3 grid) in FP64 and in the mixed
    tier, and `iterations` FP64 Davidson iterations from a seeded start
    (subspace up to 2252), each timed."""
    import numpy as np
    import torch
    from baandi_amd import workloads, lib as L, ref_davidson
    from baandi_amd.field import DensityField, gaussian_start
    from baandi_amd.hamiltonian import DistributedRealHamiltonian
    wl = workloads.si512()
    t = time.perf_counter()
    system = workloads.build(wl, device=True)
    setup_s = time.perf_counter() - t
    bands = wl["bands"]
    print(f"[bench] si512: n={system.n} grid={system.dims} proj={system.nproj} device set-up {setup_s:.2f}s",
          file=sys.stderr, flush=True)
    field = DensityField(system)
    vion_g = torch.from_numpy(system.get("ionic_density_re") + 1j * system.get("ionic_density_im")).cuda()
    vion = field.to_grid(vion_g).real.contiguous()
    pot = torch.empty(field.size, dtype=torch.float64, device="cuda")
    vmean = field.potential(torch.from_numpy(gaussian_start(system)).cuda(), vion, pot)
    g = torch.Generator(device="cpu").manual_seed(1)
    x = torch.randn(bands, system.n, generator=g, dtype=torch.float64)
    x = (x / x.norm(dim=1, keepdim=True)).cuda()
    y = torch.empty_like(x)
    out = {"workload": f"{wl['name']}: n={system.n}, grid {list(system.dims)}, {system.nproj} projectors, "
                       f"{bands} bands, potential of the Gaussian start density",
           "setup_s": round(setup_s, 3)}
    ref_y = None
    for name, prec in (("fp64", L.BD_FP64), ("mixed", L.BD_MIXED)):
        op = DistributedRealHamiltonian(system, capacity=bands, precision=prec)
        op.set_owned_potential(pot)
        op.apply_block(x, y)  # first call: plans, scratch
        torch.cuda.synchronize()
        op.set_profiling(True)
        t = time.perf_counter()
        for _ in range(reps):
            op.apply_block(x, y)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t) / reps
        st = op.stage_totals()
        plan = op.plan_info()
        plane_ms = st["plane"] / reps
        c = 16 if prec == L.BD_FP64 else 8
        model = plane_bytes_model(plan, st["pairs"] / reps, c=c, r=c // 2)
        entry = {"hx_ms": round(ms, 2), "columns": bands, "plane_split_z_lines": op.plane_split(),
                 "stage_ms": {k: round(v / reps, 3) for k, v in st.items()
                              if k not in ("applies", "pairs", "columns")},
                 "plane_model_GBs": round(model / (plane_ms * 1e-3) / 1e9, 1) if plane_ms > 0 else None,
                 "plane_frac_hbm": round(model / (plane_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if plane_ms > 0 else None}
        if ref_y is None:
            ref_y = y.clone()
            entry["plan"] = plan
        else:
            entry["max_rel_diff_vs_fp64"] = float(((y - ref_y).abs().max() / ref_y.abs().max()).item())
        out[name] = entry
        print(f"[bench] si512 {name} H·X of {bands} columns: {ms:.1f} ms {entry['stage_ms']}", file=sys.stderr,
              flush=True)
        del op
    op = DistributedRealHamiltonian(system, capacity=bands)
    op.set_owned_potential(pot)
    diag = torch.from_numpy(system.get("diagonal_static")).cuda() + vmean
    opts = ref_davidson.Options(bands=bands, threshold=1e-6, occupied=1024, first_step=True,
                                max_iterations=iterations)
    torch.cuda.synchronize()
    t = time.perf_counter()
    sol = ref_davidson.solve(op, x, diag, opts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    rec = sol.record
    out["davidson"] = {"iterations": rec["iterations"], "wall_s": round(wall, 3),
                       "dense_s": round(rec["seconds_dense"], 3), "apply_s": round(rec["seconds_apply"], 3),
                       "h_columns": rec["h_columns"], "lowest_ha": float(sol.eigenvalues[0]),
                       "eigen_retries": rec["eigen_retries"]}
    print(f"[bench] si512 davidson: {out['davidson']}", file=sys.stderr, flush=True)
    return out


def si512_scf_leg(theta=0.01, variants=True):
    """Si512 SCF to dr2 < 1e-8 Ry (Fermi-Dirac theta Ha, Thomas-Fermi mixing)
    in FP64 and, with variants, in the mixed tier: wall time, maps, and the
    mixed run's energy and eigenvalue error against the FP64 run."""
    import numpy as np
    from baandi_amd import workloads, lib as L
    from baandi_amd.scf import run_scf, ScfOptions
    wl = workloads.si512()
    t = time.perf_counter()
    system = workloads.build(wl, device=True)
    setup_s = time.perf_counter() - t
    base = dict(bands=wl["bands"], threshold_ry=1e-8, preconditioning="thomas_fermi", theta=theta)
    legs = {"fp64": ScfOptions(**base)}
    if variants:
        legs["mixed"] = ScfOptions(precision=L.BD_MIXED, **base)
    out, ref = {"workload": f"{wl['name']}: n={system.n}, grid {list(system.dims)}, {system.nproj} projectors, "
                            f"{wl['bands']} bands, Fermi-Dirac {theta} Ha, Thomas-Fermi mixing, dr2<1e-8 Ry",
                "setup_s": round(setup_s, 2)}, None
    for name, opts in legs.items():
        r = run_scf(system, opts, log=lambda m, name=name: print(f"[bench] si512 {name} {m}", file=sys.stderr,
                                                                    flush=True))
        ref = ref or r
        t_ = r["totals"]
        out[name] = {"wall_s": round(r["wall_s"], 2), "scf_iterations": r["iterations"],
                     "h_columns": t_["h_columns"], "davidson_iterations": t_["davidson_iterations"],
                     "dense_s": round(t_["dense_s"], 2), "apply_s": round(t_["apply_s"], 2),
                     "mixed_maps": t_["mixed_maps"], "energy_ha": r["energy"]["total"],
                     "energy_diff_per_atom_ha": abs(r["energy"]["total"] - ref["energy"]["total"]) / 512,
                     "eigenvalue_max_diff_ha": float(np.abs(r["eigenvalues"] - ref["eigenvalues"]).max())}
        print(f"[bench] si512 {name}: {r['wall_s']:.1f}s maps={r['iterations']} E={r['energy']['total']:.10f}",
              file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed-steps", type=int, default=5,
                    help="SCFs timed with the mixed tier (loose maps in FP32/BF16x3, final maps in FP64)")
    ap.add_argument("--kpoint-steps", type=int, default=1,
                    help="Al 6x6x6 k-point SCFs timed (0 skips the k-point leg)")
    ap.add_argument("--kpoint-only", action="store_true",
                    help="profile mode: only the Al k-point leg, its JSON object printed "
                         "(--mixed-steps 0: the complex-Davidson SCF alone)")
    ap.add_argument("--si128-only", action="store_true",
                    help="profile mode: only the Si128 smeared SCF legs, their JSON object printed "
                         "(--mixed-steps 0: the FP64 leg alone)")
    ap.add_argument("--mlp", default=None,
                    help="learned-preconditioner weights (.npz) for the learned legs (default: the packaged file)")
    ap.add_argument("--si512-only", action="store_true",
                    help="profile mode: only the Si512 leg (device set-up, 1126-column H·X, Davidson iterations)")
    ap.add_argument("--si512-scf", action="store_true",
                    help="only the Si512 SCF legs (FP64 and mixed; --mixed-steps 0: FP64 alone), JSON printed")
    ap.add_argument("--si512-iterations", type=int, default=2,
                    help="FP64 Davidson iterations of the Si512 leg (0 skips the leg)")
    ap.add_argument("--fused-shape", default=None,
                    help="THREADS,XLINES,ZROWS,XPLANES: fix the k-point fused kernel's shape instead of autotuning "
                         "it (profile runs of --kpoint-only)")
    ap.add_argument("--plane-shape", default=None,
                    help="THREADS,ROWS: fix the plane kernel's shape instead of autotuning it (profile runs: "
                         "the kernel's launches are then the SCF's only)")
    ap.add_argument("--hx-only", type=int, default=0,
                    help="profile mode: N applies of H to 160 Si64 columns, then a 256 MiB copy")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X only: the library has no multi-GPU path "
              "(run with --gpus 1).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from baandi_amd import workloads
    from baandi_amd.scf import run_scf, ScfOptions
    if args.kpoint_only:
        fshape = tuple(int(v) for v in args.fused_shape.split(",")) if args.fused_shape else None
        print(json.dumps(kpoint_leg(max(args.kpoint_steps, 1), variants=args.mixed_steps > 0, fused_shape=fshape)))
        return
    if args.si128_only:
        print(json.dumps(si128_leg(args.mlp, variants=args.mixed_steps > 0)))
        return
    if args.si512_scf:
        print(json.dumps(si512_scf_leg(variants=args.mixed_steps > 0)))
        return
    if args.si512_only:
        print(json.dumps(si512_leg(max(args.si512_iterations, 1))))
        return

    wl = workloads.si64()
    t = time.perf_counter()
    system = workloads.build(wl)
    setup_s = time.perf_counter() - t
    if args.hx_only:
        hx_profile(system, wl, args.hx_only)
        return
    # Thomas-Fermi-screened Anderson mixing for every leg (the reference's
    # Preconditioning::ThomasFermi, refscf/profile.rs:4-7): the same fixed
    # point in a third of the maps of plain mixing on this cell.
    mix = "thomas_fermi"
    # The H·X launch shapes are tuned once per op at set-up (bit-neutral).
    shape = tuple(int(v) for v in args.plane_shape.split(",")) if args.plane_shape else None
    opts = ScfOptions(bands=wl["bands"], threshold_ry=1e-8, preconditioning=mix, autotune=shape is None,
                      plane_shape=shape)
    print(f"[bench] {wl['name']}: n={system.n} grid={system.dims} proj={system.nproj} setup={setup_s:.2f}s",
          file=sys.stderr, flush=True)
    for w in range(args.warmup):
        r = run_scf(system, opts)
        print(f"[bench] warmup {w}: {r['wall_s']:.3f}s iterations={r['iterations']}", file=sys.stderr, flush=True)
    walls, results = [], []
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for k in range(args.steps):
        r = run_scf(system, opts, profile_kernels=True)
        walls.append(r["wall_s"])
        results.append(r)
        print(f"[bench] step {k}: {r['wall_s']:.3f}s iterations={r['iterations']} E={r['energy']['total']:.12f} "
              f"totals={r['totals']}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_all
    last = results[-1]
    mixed = None
    if args.mixed_steps > 0:
        from baandi_amd import lib as L
        from baandi_amd.learned import Network
        from baandi_amd.train_preconditioner import DEFAULT_OUT
        net = Network.load(args.mlp or DEFAULT_OUT)
        natoms = len(system.get("atoms_cart"))
        occ = int(round(system.electrons)) // 2
        legs = {"mixed": ScfOptions(bands=wl["bands"], threshold_ry=1e-8, precision=L.BD_MIXED,
                                    preconditioning=mix, autotune=True),
                "learned": ScfOptions(bands=wl["bands"], threshold_ry=1e-8, preconditioner=L.BD_PRECOND_MLP,
                                      mlp=net, preconditioning=mix, autotune=True),
                "mixed_learned": ScfOptions(bands=wl["bands"], threshold_ry=1e-8, precision=L.BD_MIXED,
                                            preconditioner=L.BD_PRECOND_MLP, mlp=net, preconditioning=mix,
                                            autotune=True),
                # FP32 H·X with FP64 GEMMs: no BF16x3 Gram noise, so more maps run reduced
                "fp32_hx": ScfOptions(bands=wl["bands"], threshold_ry=1e-8, precision=L.BD_FP32_HX,
                                      preconditioning=mix, autotune=True),
                "fp32_hx_learned": ScfOptions(bands=wl["bands"], threshold_ry=1e-8, precision=L.BD_FP32_HX,
                                              preconditioner=L.BD_PRECOND_MLP, mlp=net, preconditioning=mix,
                                              autotune=True)}
        mixed = {}
        for name, mopts in legs.items():
            mwalls, mres = [], None
            for k in range(args.mixed_steps):
                mres = run_scf(system, mopts)
                mwalls.append(mres["wall_s"])
                print(f"[bench] {name} step {k}: {mres['wall_s']:.3f}s iterations={mres['iterations']} "
                      f"E={mres['energy']['total']:.12f} totals={mres['totals']}", file=sys.stderr, flush=True)
            mixed[name] = {
                "wall_s": round(sum(mwalls) / len(mwalls), 4), "scf_iterations": mres["iterations"],
                "h_columns": mres["totals"]["h_columns"], "davidson_iterations": mres["totals"]["davidson_iterations"],
                "mixed_maps": mres["totals"]["mixed_maps"], "mlp_fallbacks": mres["totals"]["mlp_fallbacks"],
                "energy_diff_per_atom_ha": abs(mres["energy"]["total"] - last["energy"]["total"]) / natoms,
                "eigenvalue_max_diff_ha": float(np.abs(mres["eigenvalues"][:occ] - last["eigenvalues"][:occ]).max())}
    st = last["kernel_totals"]
    plan = last["plan"]
    launches = st["applies"]
    plane_ms = st["plane"]
    achieved = plane_bytes_model(plan, st["pairs"]) / (plane_ms * 1e-3) / 1e9 if plane_ms > 0 else 0.0
    fused = fused_bytes(plan, st["pairs"]) / (plane_ms * 1e-3) / 1e9 if plane_ms > 0 else 0.0
    per_launch_bytes = plane_bytes_model(plan, st["pairs"] / max(launches, 1))
    ms_per_step = 1e3 * sum(walls) / len(walls)
    out = {
        "metric": METRIC,
        "value": round(ms_per_step / 1e3, 4),
        "unit": "s (FP64 SCF wall)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded N(0,1) start, synthetic Gaussian pseudopotential); learned weights "
                "trained on 32 displaced Si64 cells (seeds 101-132) by baandi_amd.train_preconditioner",
        "config": {"workload": "Si64 cubic 2x2x2, a=20.52 Bohr, Gamma, 30 Ry, 160 bands (128 occ), "
                               "SCF to dr2<1e-8 Ry (Thomas-Fermi mixing), plain FP64 path",
                   "n_pw": int(system.n), "fft_grid": list(system.dims), "bands": wl["bands"],
                   "scf_iterations": last["iterations"], "h_columns": last["totals"]["h_columns"],
                   "mixed_s": None if mixed is None else mixed["mixed"]["wall_s"],
                   "mixed_learned_s": None if mixed is None else mixed["mixed_learned"]["wall_s"],
                   "speedup_mixed_learned_vs_fp64": None if mixed is None else
                   round(ms_per_step / 1e3 / mixed["mixed_learned"]["wall_s"], 3),
                   "fp32_hx_s": None if mixed is None else mixed["fp32_hx"]["wall_s"],
                   "fp32_hx_learned_s": None if mixed is None else mixed["fp32_hx_learned"]["wall_s"]},
        "variants": mixed,
        "energy_ha": last["energy"]["total"],
        "breakdown_s": {k: round(v, 4) for k, v in last["totals"].items() if k.endswith("_s")},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_plane",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": round(PLANE_TRAFFIC_BYTES_PER_PAIR * st["pairs"] / max(launches, 1)),
            "model": "3-pass pencil model, y+z stages: 4*16*(Z*n1*n0+N)+8*N bytes per pair",
            "bytes_per_launch_avg": per_launch_bytes,
            "avg_launch_ms": round(plane_ms / max(launches, 1), 4),
            "launches": launches,
            "fused_compulsory_GBs": round(fused, 1),
        },
    }
    out["subspace_gemms"] = gemm_leg()
    if args.kpoint_steps > 0:
        out["kpoint"] = kpoint_leg(args.kpoint_steps)
    if args.si512_iterations > 0:
        out["si512"] = si512_leg(args.si512_iterations)
    if not args.no_cpu:
        ms_all, ms_one, threads = cpu_baseline(system)
        gpu_ms = sum(st[k] for k in ("stick_inverse", "plane", "stick_forward", "projections",
                                     "back_product")) / max(st["columns"], 1) * 160
        out["cpu_baseline"] = {
            "value": round(ms_all, 1), "unit": "ms per Si64 FP64 H·X of 160 columns", "cores": threads,
            "kind": "port",
            "sample": "oracle/cpu_hx (C++ FP64, OpenMP over band pairs, pruned stick/plane FFTs; NumPy BLAS "
                      "for P D P^T X) on one 160-column Si64 block, median of 3 applies",
            "one_core_ms": round(ms_one, 1), "nproc": os.cpu_count(),
            "gpu_ms_same_block": round(gpu_ms, 3),
            "scf_hx_only_estimate_s": round(last["totals"]["h_columns"] * ms_all / 160 / 1e3, 2),
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
