"""Benchmark: Mcell-steps/s of the two-phase step at 256^3 on one MI355X.

Workload (the 3D two-phase configuration): 256^3 cells, FP32 stencils, slip
walls, gravity -z, density ratio 1000 (λρ = 1e-3), viscosity ratio 100
(λμ = 1e-2), a liquid pool z < 0.7L holding 32 spherical bubbles with radii
U(0.03, 0.08)L at non-overlapping centres (numpy PCG64 seed 3, >= 2-cell
clearance from walls, surface and each other), L = 256 cells, U = 1,
g0 = U^2/L, ν = UL/1e4; the pressure solve is FP32 MultiLevelPoisson
V-cycles as the preconditioner of an FP64 flexible conjugate gradient, run to
||z - A x||_2 / ||z||_2 <= 1e-6 (FP64 residual and pressure accumulation).

A step is one full MPFMomStep! (predictor + corrector: 6 VOF sweeps, 6
momentum sweeps, 2 viscous/RK updates, 2 projections, CFL).  Inputs are
resident in HBM before the timed region; the timed region brackets K steps
with device synchronisations.  By default K = 500, the configuration's own
length (after 2 warm-up steps), so the rate covers the bubbles' whole rise
and not only the quiet start, where the solves need fewer iterations.

Prints one JSON line (see the driver contract): value = whole-job
Mcell-steps/s, roofline of the dominant kernel (algorithmic bytes per launch
/ average launch time from HIP events recorded around that kernel during the
timed region), and the CPU oracle timed on the host cores over a bounded
sample of the same workload.  The other kernel classes are timed by events
over the warm-up steps (the dominant class is the one with the largest total
there); recording every class inside the timed steps costs about 3% of a step
(`--events all` does so).
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
from __graft_entry__ import load_package  # noqa: E402

_ia = load_package()
pool = importlib.import_module(_ia.__name__ + ".pool")
bubbles, pool_sdf_samples, make_case, build_sim = pool.bubbles, pool.pool_sdf_samples, pool.make_case, pool.build_sim

METRIC = "Mcell-steps/s at 256^3 two-phase; Poisson time-to-1e-6 learned vs plain MG"
HBM_PEAK = 8000.0  # GB/s, MI355X HBM3E
WEIGHTS = os.path.join(ROOT, "interfaceadvection.jl_amd", "learned_pool256.f32")


# algorithmic bytes per interior cell (FP32, 3D): each distinct array a kernel
# must read plus each it must write, once (halo re-reads excluded)
BYTES_PER_CELL = {
    "vof": 4 + 4 + 4 + 4 + 1 + 4,                    # f in/out, u_d, u⁰_d, c̄, ρuf_d
    "momentum": 12 + 12 + 4 + 4 + 4 + 4 + 1 + 12,    # ρu in/out, f_old, ρuf_d, u_d, u⁰_d, c̄, uOld
    "visc_update": 12 + 4 + 4 + 12 + 12 + 12 + 12,   # u, f, f', u⁰, ρu -> u, L
    "pcg_Ae": 4 + 12 + 4,                            # ϵ, L -> z
    "pcg_update": 20 + 12,                           # x, ϵ, r, z, iD -> x, r, z
    "increment": 4 + 4 + 12 + 4 + 8,                 # r, ϵ, L, x -> r, x
    "load_inc": 8 + 4 + 12 + 8,                      # r64, iD, L -> r, x
    "prolong_inc": 4 + 4 + 12 + 4 + 0.5 + 12,        # r, x, L, iD, coarse x -> r, x, ϵ
    "pcg_last": 4 + 4 + 8 + 8 + 4,                   # x, ϵ, r64, q64 -> x
    "cheb2_dots": 4 + 12 + 4 + 4 + 4 + 8 + 8 + 4,    # ϵ, L, r, x, iD, r64, q64 -> x
    "fcg_Ap": 4 + 8 + 12 + 16,                       # w, p, L -> p', q
    "fcg_update": 32 + 16,                           # p', q, x64, r64 -> x64, r64
    "redist": (12 + 16 + 16) / 3,                    # ϕ, ϕini (+ ϕ⁰ in stages 2, 3) -> ϕ', per stage
    # single-reduction CG step on coarse level l (per cell of that level):
    # r, w, s, iD, p, x, L -> p, s', x, r', w'
    "cg1_l1": 36 + 20, "cg1_l2": 36 + 20, "cg1_l3": 36 + 20,
}
# coarse level of a class (its cells are (N >> level)^3); 0 for the fine-level classes
LEVEL = {"cg1_l1": 1, "cg1_l2": 2, "cg1_l3": 3}
# kernel classes timed with HIP events in the timed region; the one with the
# largest total time is reported as the dominant kernel
PROFILED = ["load_inc", "prolong_inc", "cheb2_dots", "fcg_Ap", "fcg_update", "vof", "momentum",
            "visc_update", "cg1_l1", "cg1_l2", "cg1_l3"]


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs (copy-calibrated), when such a measurement is committed,
    and where it comes from (file, commands, the commit it was measured at)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None, None
    src = d.get("_source", {})
    return d.get(kernel), dict(file="profiles/pmc_traffic.json", measured_at_commit=src.get("commit"),
                               commands=src.get("commands"), calibration=d.get("_calib", {}).get("how"))


def error_vs_fp64_oracle(ia, O, N=64, steps=5, report=(1, 5)):
    """The benchmarked FP32 step against the FP64 oracle from identical inputs:
    the pool case at N^3 (the same generator: 32 bubbles, seed 3), f from the
    device's applyVOF! handed to the FP64 oracle, both runs on the device's
    Δt sequence.  Field errors after the steps in `report`."""
    bl, kw = make_case(N, "refined")
    g, _, _ = build_sim(ia, N, "refined", 1e-6)
    okw = dict(kw)
    okw["T"] = np.float64
    o = O.OracleSim((N,) * 3, (0.0, 0.0, 0.0), float(N), poisson="refined", poisson_rtol=1e-6, **okw)
    o.set(O.F, g.intf.f.astype(np.float64))
    o.set(O.F0, g.intf.f0.astype(np.float64))
    o.Δt = [np.float64(g.flow.Δt[-1])]
    V0 = float(O.inside(g.intf.f).astype(np.float64).sum())
    out = {}
    for n in range(1, steps + 1):
        st = ia.sim_step_(g)
        o.sim_step()
        o.Δt[-1] = np.float64(g.flow.Δt[-1])
        if n in report:
            fg, fo = g.intf.f.astype(np.float64), o.f
            ug, uo = g.flow.u.astype(np.float64), o.u
            # p is defined up to a constant: compare interior values about their interior means
            pg, po = O.inside(g.flow.p).astype(np.float64), O.inside(o.p).astype(np.float64)
            pg, po = pg - pg.mean(), po - po.mean()
            out[f"step_{n}"] = dict(
                f_max_abs=float(np.abs(fg - fo).max()),
                u_max_abs_rel=float(np.abs(ug - uo).max() / max(np.abs(uo).max(), 1e-30)),
                p_max_abs_rel=float(np.abs(pg - po).max() / max(np.abs(po).max(), 1e-30)),
                volume_drift_rel=float((st.vol - V0) / V0),
                f_range=[st.fmin, st.fmax],
                solve_rel_residual=[st.solve[0].rel, st.solve[1].rel])
    g.close()
    out["case"] = (f"pool {N}^3 (32 bubbles, seed 3), FP32 device (REFINED, 1e-6) vs the FP64 oracle from the "
                   f"same f, same Δt sequence")
    return out


def cpu_baseline(N_cpu, poisson, rtol, steps, warmup=1, ia=None):
    """The CPU oracle (C + OpenMP, same algorithm: FP32 stencils, FP64 flexible CG
    around the FP32 V-cycle to the same relative residual) on the same case at
    a bounded size, timed after `warmup` untimed steps.  With `ia`, the same leg
    also reports the device step's error against the FP64 oracle at 64^3."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    bl, kw = make_case(N_cpu, poisson)
    o = O.OracleSim((N_cpu,) * 3, (0.0, 0.0, 0.0), float(N_cpu), poisson="refined", poisson_rtol=rtol,
                    smooth_it=2, **kw)
    s = pool_sdf_samples(N_cpu, bl, 0.7)
    o.lib.oc_apply_vof(o.ctx, O._ptr(s))
    o.Δt[-1] = min(o.Δt[-1], o.T.type(o.cfl()))
    O.lib(o.T).oc_set_threads(threads)
    for _ in range(warmup):
        o.sim_step()
    t0 = time.perf_counter()
    its = []
    for _ in range(steps):
        o.sim_step()
        its += [o.lib.oc_last_iters(o.ctx, 0), o.lib.oc_last_iters(o.ctx, 1)]
    t = time.perf_counter() - t0
    rec = dict(value=round(N_cpu ** 3 * steps / t / 1e6, 4), unit="Mcell-steps/s", cores=threads, kind="port",
               sample=f"CPU oracle (C, OpenMP, FP32 stencils, FP64 flexible CG + FP32 V-cycle to {rtol:g}), "
                      f"the same pool+bubbles case at {N_cpu}^3, {steps} step(s) after {warmup} warm-up, "
                      f"{t:.1f} s, CG iterations per solve {its}")
    if ia is not None:
        rec["error_vs_fp64_oracle"] = error_vs_fp64_oracle(ia, O)
    return rec


def poisson_compare(ia, N, rtol, warm=8, steps=3, lr=1e-3, case=None):
    """Time-to-rtol of the pressure solve for the plain multigrid, the flexible
    CG and the learned-correction modes on the same case and the same step
    range (the learned network trains during the warm-up steps, then is frozen)."""
    out = {}
    for mode in ("plain", "refined", "learned"):
        sim, _, _ = build_sim(ia, N, mode, rtol, **(case or {}))
        if mode == "learned":
            if os.path.exists(WEIGHTS):  # trained by interfaceadvection.jl_amd/train_learned.py on other layouts
                sim.learned_weights = np.fromfile(WEIGHTS, dtype="<f4")
            else:
                sim.learned_init(seed=0)
            sim.learned_train(True, lr=lr)
        for _ in range(warm):
            ia.sim_step_(sim)
        if mode == "learned":
            sim.learned_train(False)
            trained = sim.learned_stats()
        ms, its, rel = [], [], []
        for _ in range(steps):
            st = ia.MPFMomStep_(sim, sim, sim)
            for q in range(2):
                ms.append(st.solve[q].ms)
                its.append(st.solve[q].iters)
                rel.append(st.solve[q].rel)
        rec = dict(ms=round(float(np.mean(ms)), 3), iterations=float(np.mean(its)), rel_max=float(np.max(rel)))
        if mode == "learned":
            s = sim.learned_stats()
            rec.update(train_steps=s["train_steps"], accepted=s["accepted"] - trained["accepted"],
                       timed_solves=len(ms), last_residual_ratio=round(s["last_ratio"], 4))
        out[mode] = rec
        sim.close()
        print(f"[bench] poisson {mode}: {rec}", file=sys.stderr, flush=True)
    out["learned_vs_plain_speedup"] = round(out["plain"]["ms"] / out["learned"]["ms"], 3)
    out["refined_vs_plain_speedup"] = round(out["plain"]["ms"] / out["refined"]["ms"], 3)
    out["steps"] = f"{warm} warm-up (learned: training) + {steps} timed steps, mean over the timed solves"
    return out


def redistance_rate(ia, sim, cells, d=2.0, dτ=0.5):
    """LevelSet(sim) + redistaning!(ls; d, dτ) on the state left by the timed
    steps: round(d/dτ) SSP-RK3 steps of three fused stage launches, timed with
    HIP events around each launch (after one untimed step)."""
    ls = ia.LevelSet(sim)
    ia.redistaning_(ls, d=dτ, dτ=dτ)
    ls = ia.LevelSet(sim)
    sim._call("ia_profile_enable", ia.PROFILE["redist"], 1)
    ia.redistaning_(ls, d=d, dτ=dτ)
    sim._call("ia_synchronize")
    nl, ms = ia.C.c_int(), ia.C.c_double()
    sim._call("ia_profile_read", ia.PROFILE["redist"], ia.C.byref(nl), ia.C.byref(ms))
    sim._call("ia_profile_enable", ia.PROFILE["redist"], 0)
    avg = ms.value / nl.value
    return dict(d=d, dτ=dτ, launches=nl.value, total_ms=round(ms.value, 3), avg_ms=round(avg, 4),
                GBps=round(BYTES_PER_CELL["redist"] * cells / (avg * 1e-3) / 1e9, 1),
                bytes_per_launch=int(BYTES_PER_CELL["redist"] * cells))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--case", default="pool", choices=sorted(pool.CASES),
                    help="pool: 256^3, 32 bubbles (the benchmark case); slab: 512^3, 256 bubbles")
    ap.add_argument("--N", type=int, default=None, help="grid side (default: the case's)")
    ap.add_argument("--poisson", default="refined", choices=["refined", "plain", "reference", "learned"])
    ap.add_argument("--rtol", type=float, default=1e-6)
    ap.add_argument("--profile", default="auto", help="kernel class reported in `roofline` (default: the "
                    "event-timed class with the largest total time)")
    ap.add_argument("--cpu-N", type=int, default=192)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the plain / refined / learned solver comparison")
    ap.add_argument("--no-events", action="store_true", help="time the steps without the per-launch HIP events")
    ap.add_argument("--events", choices=["dominant", "all"], default="dominant",
                    help="classes timed by HIP events inside the timed steps: the dominant one (the others are timed "
                         "over the warm-up steps), or all of them")
    ap.add_argument("--compare-warm", type=int, default=8)
    ap.add_argument("--compare-steps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only (the path would shard by z-slabs across GPUs, "
              "but no multi-GPU path is built)", file=sys.stderr)
        sys.exit(2)
    ia = load_package()
    case = dict(pool.CASES[args.case])
    N = args.N or case.pop("N")
    case.pop("N", None)
    t0 = time.perf_counter()
    sim, bl, kw = build_sim(ia, N, args.poisson, args.rtol, **case)
    if args.poisson == "learned":  # the shipped parameters, frozen
        if os.path.exists(WEIGHTS):
            sim.learned_weights = np.fromfile(WEIGHTS, dtype="<f4")
        else:
            sim.learned_init(seed=0)
    print(f"[bench] setup {N}^3 with {len(bl)} bubbles in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    classes = [] if args.no_events else (PROFILED if args.profile == "auto" else [args.profile])
    cells = N ** 3

    def read_classes(names):
        out = {}
        for k in names:
            nl, ms = ia.C.c_int(), ia.C.c_double()
            sim._call("ia_profile_read", ia.PROFILE[k], ia.C.byref(nl), ia.C.byref(ms))
            sim._call("ia_profile_enable", ia.PROFILE[k], 0)
            if nl.value:
                avg = ms.value / nl.value
                kc = (N >> LEVEL.get(k, 0)) ** 3
                out[k] = dict(launches=nl.value, total_ms=round(ms.value, 3), avg_ms=round(avg, 4),
                              GBps=round(BYTES_PER_CELL[k] * kc / (avg * 1e-3) / 1e9, 1), cells=kc)
        return out

    # Every class is timed over the warm-up steps; inside the timed steps only
    # the dominant one (largest total over the warm-up) keeps its events, as
    # the events of every class cost about 3% of the step.
    warm_classes = classes if (args.events == "dominant" and args.warmup > 0) else []
    for k in warm_classes:
        sim._call("ia_profile_enable", ia.PROFILE[k], 1)
    for _ in range(args.warmup):
        ia.sim_step_(sim)
    sim._call("ia_synchronize")
    warm = read_classes(warm_classes)
    if warm:
        classes = [max(warm, key=lambda k: warm[k]["total_ms"])]
    for k in classes:
        sim._call("ia_profile_enable", ia.PROFILE[k], 1)
    iters, solve_ms, rel = [], [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = ia.MPFMomStep_(sim, sim, sim)
        for q in range(2):
            iters.append(st.solve[q].iters)
            solve_ms.append(st.solve[q].ms)
            rel.append(st.solve[q].rel)
    sim._call("ia_synchronize")
    t = time.perf_counter() - t0
    timed = read_classes(classes)
    dom = max(timed, key=lambda k: timed[k]["total_ms"]) if timed else None
    its = iters if len(iters) <= 40 else f"mean {np.mean(iters):.2f}, min {min(iters)}, max {max(iters)}"
    print(f"[bench] {args.steps} steps in {t:.3f} s; CG iterations per solve {its}", file=sys.stderr, flush=True)
    value = cells * args.steps / t / 1e6
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3D two-phase pool (liquid below {case['zmax']} L) + {len(bl)} bubbles of radius "
                               f"U({case['rmin']}, {case['rmax']}) L (seed {case['seed']}), {N}^3, slip walls, "
                               f"gravity, density ratio 1000, viscosity ratio 100",
                   "grid": [N, N, N], "poisson": args.poisson, "poisson_rtol": args.rtol,
                   "cg_iterations_per_solve_mean": float(np.mean(iters)),
                   "poisson_time_to_1e-6_ms": round(float(np.mean(solve_ms)), 3),
                   "poisson_rel_residual_max": float(np.max(rel)) if rel else None},
        "kernels": timed if not warm else warm,
        "kernels_timed_over": "the timed steps" if not warm else f"the {args.warmup} warm-up steps "
                              f"(the roofline kernel: the timed steps)",
    }
    if dom:
        traffic, traffic_src = pmc_traffic(dom) if (args.case, N) == ("pool", 256) else (None, None)
        avg_ms = timed[dom]["avg_ms"]
        bpl = BYTES_PER_CELL[dom] * timed[dom]["cells"]
        achieved = bpl / (avg_ms * 1e-3) / 1e9
        out["roofline"] = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                           "traffic": traffic,
                           "traffic_source": traffic_src,
                           "bytes_per_launch": int(bpl), "avg_launch_ms": avg_ms,
                           "launches": timed[dom]["launches"]}
    out["levelset_redistance"] = redistance_rate(ia, sim, cells)
    sim.close()
    if not args.no_compare:
        out["config"]["poisson_compare"] = poisson_compare(ia, N, args.rtol, warm=args.compare_warm,
                                                           steps=args.compare_steps, case=case)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_N, args.poisson, args.rtol, args.cpu_steps, ia=ia)
        out["error_vs_fp64_oracle"] = out["cpu_baseline"].pop("error_vs_fp64_oracle")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
