"""Train the learned Davidson preconditioner on the GPU and save its weights.

    python -m baandi_amd.train_preconditioner [--workload si8|si64] [--cells 32] [--steps 3000] [--out PATH]

Training pairs come from the FP64 GPU Davidson on displaced diamond-silicon
cells (Gamma, 30 Ry, synthetic Z=4 pseudopotential; by default 32 Si64 cells,
displacement seeds 101-132): each cell gets atomic
displacements N(0, sigma) and a partly converged SCF potential (a few maps
of the GPU SCF), then seeded Davidson runs yield residual -> ideal-correction
pairs (baandi_amd.learned.generate_pairs). Training runs the project's
bd_mlp_* kernels (Adam, bf16 weights, FP32 masters), starting from K = 1
(the QE correction). The loss is a proxy: every `--check-every` steps the
network runs full smeared SCFs (dr2 < 1e-8 Ry) on held-out validation cells
(displacement seeds 201-202), and the parameters kept are those with the
fewest Davidson iterations over them (ties: fewest H columns), the start
included, since each iteration pays the FP64 subspace step.
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


# Displaced cells have nearly closed gaps at Gamma: the validation SCFs use
# Fermi-Dirac smearing (Ha) so that they converge.
VALIDATION_SMEARING_HA = 0.01


def validation_score(net, systems, bands, preconditioner=None):
    """(Davidson iterations, H columns) summed over full SCFs of `systems`
    with the network (or another preconditioner); (inf, inf) if one fails."""
    from . import lib as L
    from .scf import run_scf, ScfOptions
    pre = L.BD_PRECOND_MLP if preconditioner is None else preconditioner
    its = cols = 0
    for system in systems:
        try:
            r = run_scf(system, ScfOptions(bands=bands, threshold_ry=1e-8, preconditioner=pre,
                                           mlp=net if pre == L.BD_PRECOND_MLP else None,
                                           preconditioning="thomas_fermi", theta=VALIDATION_SMEARING_HA))
        except Exception as e:  # a diverging network scores worst
            print(f"[train] validation SCF failed: {e}", flush=True)
            return float("inf"), float("inf")
        its += r["totals"]["davidson_iterations"]
        cols += r["totals"]["h_columns"]
    return its, cols


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("si8", "si64"), default="si64",
                    help="cell the displaced training cells are drawn from")
    ap.add_argument("--cells", type=int, default=32)
    ap.add_argument("--first-seed", type=int, default=101, help="displacement seed of the first cell")
    ap.add_argument("--sigma", type=float, default=0.05, help="displacement s.d., Bohr")
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--check-every", type=int, default=250, help="steps between validation SCFs")
    ap.add_argument("--validation-seeds", default="201,202")
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args(argv)
    from . import workloads
    from .learned import Network, train, qe_loss, loss_and_grad
    base = workloads.si8() if args.workload == "si8" else workloads.si64()
    pairs = None
    for c in range(args.cells):
        p = cell_pairs(displaced(base, args.sigma, args.first_seed + c), seed=c)
        pairs = p if pairs is None else pairs.extend(p)
    from . import lib as L
    vseeds = [int(v) for v in args.validation_seeds.split(",") if v]
    vsystems = [workloads.build(displaced(base, args.sigma, s)) for s in vseeds]
    print(f"[train] validation cells {vseeds}: QE scores {validation_score(None, vsystems, base['bands'], L.BD_PRECOND_QE)} "
          f"(Davidson iterations, H columns)", flush=True)
    net = Network.initial(args.seed, output_scale=0.0)
    every = list(range(pairs.bands))
    best = (validation_score(net, vsystems, base["bands"]), 0, net.parameters())
    print(f"[train] step 0: score {best[0]}", flush=True)
    t = time.perf_counter()
    done, seed = 0, args.seed
    while done < args.steps:
        chunk = min(args.check_every, args.steps - done)
        train(net, pairs, chunk, batch=args.batch, lr=args.lr, seed=seed)
        done += chunk
        seed += 1
        sc = validation_score(net, vsystems, base["bands"])
        loss, _ = loss_and_grad(net, pairs, every)
        print(f"[train] step {done}: loss {loss:.4f}, score {sc}", flush=True)
        if sc < best[0]:
            best = (sc, done, net.parameters())
    net = Network(best[2])
    final, _ = loss_and_grad(net, pairs, every)
    base_loss = qe_loss(pairs)
    print(f"[train] {args.steps} steps in {time.perf_counter() - t:.1f}s: kept step {best[1]} (score {best[0]}), "
          f"loss {final:.4f} (QE {base_loss:.4f}) over {pairs.bands} pairs", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    net.save(args.out, loss=final, qe_loss=base_loss, pairs=pairs.bands, cells=args.cells,
             first_seed=args.first_seed, sigma=args.sigma, steps=args.steps, kept_step=best[1],
             validation_iterations=best[0][0], validation_seeds=np.array(vseeds), workload=args.workload)
    print(f"[train] saved {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
