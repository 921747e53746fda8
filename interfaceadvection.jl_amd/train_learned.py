"""Train the learned Poisson correction on the two-phase pool workload.

The network is trained on the device, inside the solver: every projection of
a LEARNED-mode simulation with training switched on takes one Adam step on
the loss ‖z - A x‖² / ‖r0‖² of its own corrected warm start.  Training runs
over several bubble layouts (seeds 1000, 1001, ...) of the 3D pool case
(liquid below 0.7 L, 32 bubbles of radius U(0.03, 0.08) L, density ratio
1000, viscosity ratio 100, slip walls, gravity), a number of time steps each.

    python interfaceadvection.jl_amd/train_learned.py --N 256 --seeds 45 --steps 25 \\
        --holdout 5 --compare interfaceadvection.jl_amd/learned_pool256.f32 \\
        --out interfaceadvection.jl_amd/learned_pool256.f32

writes the parameters (IA_LEARNED_NPAR float32 values, little-endian, the
layout of ia_learned_set_weights) and a JSON record of the run beside them.

Each projection is one snapshot of the full grid; the network is a stack of
3x3x3 convolutions, so an N³ snapshot trains it as (N/128)³ snapshots of
128³ patches would (45 layouts x 25 steps x 2 projections at 256³: 18,000
patch-equivalents).  With --holdout K, the trained parameters are then run
with training off on K further layouts (the seeds after the training ones),
against REFINED on the same layouts and, with --compare, against a previous
parameter file: the residual ratio of the corrected warm start, the
flexible-CG iterations and the solve time per projection.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--seeds", type=int, default=4)
    ap.add_argument("--first-seed", type=int, default=1000)
    ap.add_argument("--steps", type=int, default=25)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--holdout", type=int, default=0)
    ap.add_argument("--compare", default=None, help="a previous parameter file to evaluate beside the new one")
    ap.add_argument("--out", default=os.path.join(HERE, "learned_pool256.f32"))
    args = ap.parse_args()
    sys.path.insert(0, ROOT)
    import importlib
    from __graft_entry__ import load_package
    ia = load_package()
    pool = importlib.import_module(ia.__name__ + ".pool")
    w = None
    log = []
    t0 = time.perf_counter()
    for k in range(args.seeds):
        seed = args.first_seed + k
        sim, _, _ = pool.build_sim(ia, args.N, "learned", 1e-6, seed=seed)
        if w is None:
            sim.learned_init(seed=args.init_seed)
        else:
            sim.learned_weights = w
        sim.learned_train(True, lr=args.lr)
        ratios, iters = [], []
        for _ in range(args.steps):
            st = ia.MPFMomStep_(sim, sim, sim)
            ratios.append(sim.learned_stats()["last_ratio"])
            iters += [st.solve[0].iters, st.solve[1].iters]
        w = sim.learned_weights
        rec = dict(seed=seed, steps=args.steps, train_steps=sim.learned_stats()["train_steps"],
                   residual_ratio_first=round(ratios[0], 4), residual_ratio_last=round(ratios[-1], 4),
                   cg_iterations_mean=float(np.mean(iters)), elapsed_s=round(time.perf_counter() - t0, 1))
        log.append(rec)
        print("[train]", rec, flush=True)
        sim.close()
    old = np.fromfile(args.compare, dtype="<f4") if args.compare else None
    w = np.array(w, dtype=np.float32)
    w.astype("<f4").tofile(args.out)
    npatch = 2 * args.steps * args.seeds * max(1, (args.N // 128) ** 3)
    held = []
    for k in range(args.holdout):
        seed = args.first_seed + args.seeds + k
        rec = dict(seed=seed)
        cases = [("refined", None), ("learned", w)] + ([("learned_previous", old)] if old is not None else [])
        for name, par in cases:
            sim, _, _ = pool.build_sim(ia, args.N, "refined" if par is None else "learned", 1e-6, seed=seed)
            if par is not None:
                sim.learned_weights = par
            ratios, iters, ms = [], [], []
            for _ in range(args.steps):
                st = ia.MPFMomStep_(sim, sim, sim)
                if par is not None:
                    ratios.append(sim.learned_stats()["last_ratio"])
                iters += [st.solve[0].iters, st.solve[1].iters]
                ms += [st.solve[0].ms, st.solve[1].ms]
            rec[name] = dict(cg_iterations_mean=float(np.mean(iters)), solve_ms_mean=round(float(np.mean(ms)), 3))
            if ratios:
                rec[name]["residual_ratio_mean"] = round(float(np.mean(ratios)), 4)
            sim.close()
        held.append(rec)
        print("[holdout]", rec, flush=True)
    with open(os.path.splitext(args.out)[0] + ".json", "w") as fh:
        json.dump(dict(N=args.N, lr=args.lr, init_seed=args.init_seed, runs=log, parameters=int(w.size),
                       training_snapshots=2 * args.steps * args.seeds, patch_equivalents_128=npatch,
                       holdout=held), fh, indent=1)
    print(f"[train] wrote {args.out} ({npatch} 128³-patch equivalents)", flush=True)


if __name__ == "__main__":
    main()
