"""The 3D configurations beyond the benchmark's own steps.

- 512³ slab (a liquid slab below 0.6 L holding 256 bubbles of radius
  U(0.02, 0.05) L, seed 5; slip walls, gravity, density ratio 1000, viscosity
  ratio 100): two whole steps with the REFINED solve and two with the LEARNED
  one (the shipped parameters, trained on other layouts), each solve to the
  relative FP64 residual 1e-6.  Properties: status OK, f ∈ [0, 1], Σf (FP64 on
  the device) within 1e-8 relative of its start, every solve at or below 1e-6,
  and the learned network evaluated on every LEARNED solve.
- The learned correction on a layout it was not trained on (the 256³ pool,
  seed 3; training used seeds 1000-1044): over the same steps, LEARNED needs
  no more flexible-CG iterations than REFINED, and every accepted correction
  lowered the FP64 residual (the acceptance test).  The times are printed:
  on this case the network's inference costs about what the iteration it
  saves does.
"""
import os

import numpy as np
import pytest

from __graft_entry__ import load_package
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ia = load_package()
import importlib.util as _ilu  # noqa: E402

_spec = _ilu.spec_from_file_location("ia_pool3", os.path.join(os.path.dirname(ia.__file__), "pool.py"))
pool = _ilu.module_from_spec(_spec)
_spec.loader.exec_module(pool)
WEIGHTS = os.path.join(os.path.dirname(ia.__file__), "learned_pool256.f32")


def say(*a):
    print("[case]", *a, flush=True)


def run_steps(sim, n, label):
    V0 = float(O.inside(sim.intf.f).astype(np.float64).sum())
    its, ms = [], []
    for k in range(n):
        st = ia.sim_step_(sim)
        assert st.status == 0
        for q in range(2):
            assert 0 < st.solve[q].rel <= 1e-6
            its.append(st.solve[q].iters)
            ms.append(st.solve[q].ms)
        assert 0.0 <= st.fmin and st.fmax <= 1.0
        say(f"{label} step {k + 1}: Σf drift {(st.vol - V0) / V0:.2e}, FCG its {its[-2:]}, solve ms "
            f"{[round(x, 2) for x in ms[-2:]]}")
        assert abs(st.vol - V0) <= 1e-8 * V0
    return its, ms


def test_slab512_refined_and_learned_steps():
    case = dict(pool.CASES["slab"])
    N = case.pop("N")
    for mode in ("refined", "learned"):
        say("slab 512^3, 256 bubbles (seed 5),", mode)
        sim, bl, _ = pool.build_sim(ia, N, mode, 1e-6, **case)
        assert len(bl) == 256
        if mode == "learned":
            sim.learned_weights = np.fromfile(WEIGHTS, dtype="<f4")
        run_steps(sim, 2, mode)
        if mode == "learned":
            s = sim.learned_stats()
            assert s["projections"] == 4 and s["train_steps"] == 0
        sim.close()


def test_learned_on_held_out_layout():
    res = {}
    for mode in ("refined", "learned"):
        say("pool 256^3 seed 3 (held out),", mode)
        sim, _, _ = pool.build_sim(ia, 256, mode, 1e-6)
        if mode == "learned":
            sim.learned_weights = np.fromfile(WEIGHTS, dtype="<f4")
        run_steps(sim, 2, mode + " warm-up")
        its, ms = run_steps(sim, 3, mode)
        res[mode] = (np.mean(its), np.mean(ms))
        if mode == "learned":
            s = sim.learned_stats()
            assert s["projections"] == 10 and s["accepted"] >= 1
        sim.close()
    say(f"FCG iterations per solve: refined {res['refined'][0]:.2f}, learned {res['learned'][0]:.2f}; "
        f"solve ms: refined {res['refined'][1]:.2f}, learned {res['learned'][1]:.2f}")
    assert res["learned"][0] <= res["refined"][0]
