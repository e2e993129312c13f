"""Train the learned Davidson preconditioner on the GPU and save its weights.

    python -m baandi_amd.train_preconditioner [--workload si8|si64] [--cells 8] [--steps 3000] [--out PATH]

Training pairs come from the FP64 GPU Davidson on displaced diamond-silicon
cells (Gamma, 30 Ry, synthetic Z=4 pseudopotential): each cell gets atomic
displacements N(0, sigma) and a partly converged SCF potential (a few maps
of the GPU SCF), then seeded Davidson runs yield residual -> ideal-correction
pairs (baandi_amd.learned.generate_pairs). Training runs the project's
bd_mlp_* kernels (Adam, bf16 weights, FP32 masters).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "preconditioner_si.npz")


def displaced(workload, sigma_bohr, seed):
    """The workload with its atoms displaced by N(0, sigma) Bohr (cubic cells)."""
    rng = np.random.default_rng(seed)
    a = float(workload["rows"][0][0])
    atoms = np.array(workload["atoms"], dtype=np.float64)
    atoms = atoms + rng.normal(0.0, sigma_bohr / a, atoms.shape)
    out = dict(workload)
    out["atoms"] = atoms.tolist()
    out["name"] = f"{workload['name']}_displaced{seed}"
    return out


def cell_pairs(workload, seed, maps=12, log=print):
    import torch
    from . import workloads
    from .hamiltonian import DistributedRealHamiltonian
    from .learned import generate_pairs
    from .scf import run_scf, ScfOptions
    system = workloads.build(workload)
    bands = workload["bands"]
    occupied = int(round(system.electrons)) // 2
    # A screened potential: a few SCF maps (the training needs realistic
    # spectra, not a converged density).
    potential = run_scf(system, ScfOptions(bands=bands, threshold_ry=1e-3, max_iterations=maps,
                                           allow_unconverged=True))["potential"]
    op = DistributedRealHamiltonian(system, capacity=bands + 8)
    pot = torch.from_numpy(potential.reshape(-1)).cuda()
    op.set_owned_potential(pot)
    vbar = float(potential.mean())
    diag = torch.from_numpy(system.get("diagonal_static") + vbar).cuda()
    t = time.perf_counter()
    pairs = generate_pairs(op, diag, vbar, bands, occupied, seed)
    log(f"[train] {workload['name']}: {pairs.bands} pairs (n={pairs.n}) in {time.perf_counter() - t:.1f}s")
    return pairs


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("si8", "si64"), default="si8",
                    help="cell the displaced training cells are drawn from")
    ap.add_argument("--cells", type=int, default=8)
    ap.add_argument("--sigma", type=float, default=0.05, help="displacement s.d., Bohr")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args(argv)
    from . import workloads
    from .learned import Network, train, tpa_loss, loss_and_grad
    base = workloads.si8() if args.workload == "si8" else workloads.si64()
    pairs = None
    for c in range(args.cells):
        p = cell_pairs(displaced(base, args.sigma, 1000 + c), seed=c)
        pairs = p if pairs is None else pairs.extend(p)
    net = Network.initial(args.seed)
    t = time.perf_counter()
    losses = train(net, pairs, args.steps, batch=args.batch, lr=args.lr, seed=args.seed)
    every = list(range(pairs.bands))
    final, _ = loss_and_grad(net, pairs, every)
    print(f"[train] {args.steps} steps in {time.perf_counter() - t:.1f}s: loss {losses[:10].mean():.4f} -> "
          f"{final:.4f} (TPA {tpa_loss(pairs):.4f}) over {pairs.bands} pairs", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    net.save(args.out, loss=final, tpa_loss=tpa_loss(pairs), pairs=pairs.bands, cells=args.cells,
             sigma=args.sigma, steps=args.steps, workload=args.workload)
    print(f"[train] saved {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
