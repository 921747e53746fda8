"""The pool (32 bubbles of radius U(0.03, 0.08) L, seed 3; slip walls,
gravity, density ratio 1000, viscosity ratio 100) over many steps: REFINED
steps on the device against the FP64 oracle run over the same Δt sequence
from the same initial f.  The step is MPFMomStep! (flow.jl:60-109); the
integrals are metrics.jl:15-50 (kinetic energy, momentum, potential energy,
enstrophy, Σf).  The grid side is 256 (the benchmark case), or C3_N.

Not collected by pytest: the FP64 oracle takes about 30 s per 256³ step on
eight cores, and several times that where the solves get hard, so the
oracle leg of a long run takes hours.  Three stages, the first on a GPU,
the others on any CPU:

    python tests/c3_trajectory.py device DEV.npz [--steps 500 --every 50]
    python tests/c3_trajectory.py oracle DEV.npz ORC.npz
    python tests/c3_trajectory.py compare DEV.npz ORC.npz

`device` writes the initial and final f (ghost-inclusive), the Δt of every
step, per-step volume, f range, solver iterations and residuals, and the
integrals every `every` steps.  `oracle` replays the Δt sequence on the
FP64 oracle and records the same.  `compare` prints the integrals side by
side and matches the gas bubbles (connected regions of f < ½ below the free
surface) of the two final states by nearest centroid."""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

N = int(os.environ.get("C3_N", "256"))  # a smaller side only for trying the stages out
INTEGRALS = ("kinetic_energy", "momentum_x", "momentum_y", "momentum_z", "potential_energy", "enstrophy",
             "volume")


def say(*a):
    print("[c3]", *a, flush=True)


def _pool():
    import importlib.util as ilu
    from __graft_entry__ import load_package
    ia = load_package()
    spec = ilu.spec_from_file_location("ia_pool_c3", os.path.join(os.path.dirname(ia.__file__), "pool.py"))
    pool = ilu.module_from_spec(spec)
    spec.loader.exec_module(pool)
    return ia, pool


def device(out, steps, every):
    ia, pool = _pool()
    sim, bl, _ = pool.build_sim(ia, N, "refined", 1e-6)
    f_init = sim.intf.f.copy()
    rec = dict(dt=[], vol=[], fmin=[], fmax=[], iters=[], rel=[], metric_steps=[], metrics=[])

    def snap(n):
        m = sim.metrics(U=(1.0, 1.0, 1.0))
        rec["metric_steps"].append(n)
        rec["metrics"].append([m["kinetic_energy"], *m["momentum"], m["potential_energy"], m["enstrophy"],
                               m["volume"]])

    snap(0)
    t0 = time.perf_counter()
    for n in range(1, steps + 1):
        rec["dt"].append(float(sim.flow.Δt[-1]))
        st = ia.sim_step_(sim)
        if st.status not in (0, 3):
            raise RuntimeError(f"step {n}: status {st.status}")
        rec["vol"].append(st.vol)
        rec["fmin"].append(st.fmin)
        rec["fmax"].append(st.fmax)
        rec["iters"].append([st.solve[0].iters, st.solve[1].iters])
        rec["rel"].append([st.solve[0].rel, st.solve[1].rel])
        if n % every == 0 or n == steps:
            snap(n)
            say(f"device step {n}: {time.perf_counter() - t0:.1f} s, Σf {st.vol:.10e}, FCG its {rec['iters'][-1]}")
    np.savez_compressed(out, f_init=f_init, f_final=sim.intf.f, **{k: np.array(v) for k, v in rec.items()})
    sim.close()


def oracle(dev, out, threads):
    from oracle import oracle as O
    _, pool = _pool()
    d = np.load(dev)
    _, kw = pool.make_case(N, "refined")
    kw["T"] = np.float64
    O.lib(np.float64).oc_set_threads(int(threads))
    o = O.OracleSim((N,) * 3, (0.0, 0.0, 0.0), float(N), poisson="refined", poisson_rtol=1e-6, **kw)
    f = d["f_init"].astype(np.float64)
    o.set(O.F, f)
    o.set(O.F0, f)
    dts = d["dt"]
    every_steps = set(int(s) for s in d["metric_steps"])
    rec = dict(vol=[], iters=[], metric_steps=[], metrics=[])

    def snap(n):
        m = o.metrics(U=(1.0, 1.0, 1.0))
        rec["metric_steps"].append(n)
        rec["metrics"].append([m[0], m[1], m[2], m[3], m[4], m[5], m[6]])

    snap(0)
    t0 = time.perf_counter()
    for n in range(1, len(dts) + 1):
        o.Δt[-1] = np.float64(dts[n - 1])
        o.sim_step()
        rec["vol"].append(float(O.inside(o.f).sum()))
        rec["iters"].append([o.lib.oc_last_iters(o.ctx, 0), o.lib.oc_last_iters(o.ctx, 1)])
        if n in every_steps:
            snap(n)
            # the run so far, so that an interrupted run still compares up to here
            np.savez_compressed(out, f_final=o.f, last_step=n, **{k: np.array(v) for k, v in rec.items()})
        if n % 10 == 0 or n in every_steps:
            say(f"oracle step {n}/{len(dts)}: {time.perf_counter() - t0:.0f} s, FCG its {rec['iters'][-1]}")
    np.savez_compressed(out, f_final=o.f, last_step=len(dts), **{k: np.array(v) for k, v in rec.items()})


def bubbles_of(f):
    """Connected gas regions (f < ½) that do not touch the top ghost-free
    layer: (volume in cells, centroid) of each."""
    from scipy import ndimage
    g = f[1:-1, 1:-1, 1:-1] < 0.5
    lab, n = ndimage.label(g)
    top = set(np.unique(lab[:, :, -1])) - {0}
    out = []
    for k in range(1, n + 1):
        if k in top:
            continue
        idx = np.argwhere(lab == k)
        out.append((len(idx), idx.mean(axis=0) + 0.5))
    return out


def compare(dev, orc):
    d, o = np.load(dev), np.load(orc)
    last = int(o["last_step"]) if "last_step" in o else len(d["dt"])
    if last < len(d["dt"]):
        print(f"the oracle run stopped at step {last}: the final fields are compared with the device's end")
    print("step | " + " | ".join(INTEGRALS) + "  (device / oracle - 1)")
    for k, n in enumerate(d["metric_steps"][:len(o["metric_steps"])]):
        a, b = d["metrics"][k], o["metrics"][k]
        rel = [(x - y) / (abs(y) if abs(y) > 0 else 1.0) for x, y in zip(a, b)]
        print(f"{int(n):4d} | " + " | ".join(f"{x:.3e}/{r:+.1e}" for x, r in zip(b, rel)))
    if len(o["iters"]):
        di, oi = d["iters"][:len(o["iters"])], o["iters"]
        print(f"flexible-CG iterations per solve: device {di.mean():.2f} (max {di.max()}), "
              f"oracle {oi.mean():.2f} (max {oi.max()})")
    vd, vo = d["vol"][:len(o["vol"])], o["vol"]
    print(f"Σf drift over the run: device {(vd[-1] - vd[0]) / vd[0]:+.2e}, oracle {(vo[-1] - vo[0]) / vo[0]:+.2e}")
    fd, fo = d["f_final"].astype(np.float64), o["f_final"]
    diff = np.abs(fd - fo)[1:-1, 1:-1, 1:-1]
    print(f"final f: max |Δf| {diff.max():.3e}, Σ|Δf| {diff.sum():.3e} cells "
          f"({np.count_nonzero(diff > 1e-3)} cells differ by more than 1e-3)")
    bd, bo = bubbles_of(fd), bubbles_of(fo)
    print(f"gas bubbles below the surface: device {len(bd)}, oracle {len(bo)}")
    worst = 0.0
    for vol, c in sorted(bo, key=lambda t: -t[0]):
        if not bd:
            break
        j = int(np.argmin([np.linalg.norm(c - cd) for _, cd in bd]))
        dist = float(np.linalg.norm(c - bd[j][1]))
        worst = max(worst, dist)
        print(f"  oracle bubble {vol:6d} cells at {np.round(c, 2)}: device {bd[j][0]:6d} cells, centroid "
              f"{dist:.3f} cells away")
    print(f"largest centroid distance: {worst:.3f} cells")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("stage", choices=["device", "oracle", "compare"])
    ap.add_argument("files", nargs="+")
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    if a.stage == "device":
        device(a.files[0], a.steps, a.every)
    elif a.stage == "oracle":
        oracle(a.files[0], a.files[1], a.threads)
    else:
        compare(a.files[0], a.files[1])


if __name__ == "__main__":
    main()
