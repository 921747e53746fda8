"""Whole steps of the benchmarked path: the 3D pool workload (a liquid pool
below 0.7 L holding spherical bubbles, slip walls, gravity -z, density ratio
1000, viscosity ratio 100) advanced by `sim_step!` → `MPFMomStep!`
(InterfaceAdvection.jl:100-103, flow.jl:60-109) with the REFINED pressure
solve (FP64 flexible CG preconditioned by FP32 V-cycles, relative FP64
residual 1e-6), the mode the benchmark times.

- 64³, the 32-bubble generator at that size (radii U(0.03, 0.08)·64 cells):
  three whole steps on the GPU and on the FP32 oracle from identical inputs.
  Both run the same solver (the oracle's oc_project_fcg mirrors the device's
  V-cycle, coarse iteration rule and fine smoother); their dot products sum in
  different orders, so each solve stops at a slightly different iterate, both
  below the tolerance.  Bars: f within 1e-6 (absolute, f ∈ [0, 1]; measured
  6e-8), u within 1e-5 of max|u| (measured 5e-8: the solves' stopping points
  differ by up to their 1e-6 relative residual, which the operator's density
  ratio of 1000 can amplify in u), every solve on both sides at or below
  1e-6, and the flexible-CG iteration counts within one of each other
  (measured equal).  The same with the fine smoother switched to two pcg!
  iterations.
- 64³ against the FP64 oracle from the same FP32 inputs (the error of the
  FP32 stencils): f within 2e-5 (measured 1.7e-6 and 2.9e-6) and u within
  1e-3 of max|u| (measured 1.1e-4 and 1.2e-4) after 1 and 5 steps.
- 128³: one step with the coarse CG step's workgroups at the default sizes
  and at three other sizes per level, each against the FP32 oracle.
- 256³, the benchmark case itself (32 bubbles, seed 3): five steps with
  property checks: status OK, f ∈ [0, 1], Σf (FP64 on the device) drifting by
  at most 1e-8 relative (measured 1.3e-10 after 5 steps), every solve at or
  below 1e-6.
"""
import os

import numpy as np
import pytest

from __graft_entry__ import load_package
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ia = load_package()
import importlib.util as _ilu  # noqa: E402

_spec = _ilu.spec_from_file_location("ia_pool", os.path.join(os.path.dirname(ia.__file__), "pool.py"))
pool = _ilu.module_from_spec(_spec)
_spec.loader.exec_module(pool)


def say(*a):
    print("[case]", *a, flush=True)


def pool_pair(N, T_oracle=np.float32, smooth_it=None):
    """The pool case at N³ on the GPU (FP32, REFINED) and on the oracle of
    type T_oracle, with the oracle's f, f⁰ and Δt set from the GPU's."""
    env_old = os.environ.get("IA_FCG_SMOOTH")
    if smooth_it is not None:
        os.environ["IA_FCG_SMOOTH"] = str(smooth_it)
    try:
        g, bl, kw = pool.build_sim(ia, N, "refined", 1e-6)
    finally:
        if smooth_it is not None:
            if env_old is None:
                del os.environ["IA_FCG_SMOOTH"]
            else:
                os.environ["IA_FCG_SMOOTH"] = env_old
    okw = dict(kw)
    okw["T"] = T_oracle
    o = O.OracleSim((N, N, N), (0.0, 0.0, 0.0), float(N), poisson="refined", poisson_rtol=1e-6, **okw)
    if smooth_it is not None:
        o.set_fine_cheb(False)  # the oracle's smooth_it (2) pcg! iterations
    f = g.intf.f
    o.set(O.F, f.astype(T_oracle))
    o.set(O.F0, g.intf.f0.astype(T_oracle))
    o.Δt = [o.T.type(g.flow.Δt[-1])]
    return g, o


def rel_u(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("smooth", ["chebyshev", "pcg2"])
def test_pool64_refined_steps_match_oracle(smooth):
    say("pool 64^3 REFINED, 3 steps vs FP32 oracle, fine smoother", smooth)
    g, o = pool_pair(64, np.float32, smooth_it=2 if smooth == "pcg2" else None)
    V0 = O.inside(g.intf.f).astype(np.float64).sum()
    for n in range(3):
        st = ia.sim_step_(g)
        o.sim_step()
        gi = [st.solve[0].iters, st.solve[1].iters]
        oi = [o.lib.oc_last_iters(o.ctx, 0), o.lib.oc_last_iters(o.ctx, 1)]
        df = float(np.abs(g.intf.f.astype(np.float64) - o.f.astype(np.float64)).max())
        du = rel_u(g.flow.u, o.u)
        say(f"step {n + 1}: max|Δf| {df:.3e}, max|Δu|/max|u| {du:.3e}, FCG its gpu {gi} oracle {oi}, "
            f"rel {st.solve[0].rel:.2e} {st.solve[1].rel:.2e}, dt {g.flow.Δt[-1]:.6g} / {o.Δt[-1]:.6g}")
        assert st.status == 0
        assert df <= 1e-6
        assert du <= 1e-5
        for q in range(2):
            assert 0 < st.solve[q].rel <= 1e-6
            assert abs(gi[q] - oi[q]) <= 1
        assert abs(float(g.flow.Δt[-1]) - float(o.Δt[-1])) <= 1e-4 * float(o.Δt[-1])
    # Σf from the step statistics (FP64 on the device) equals the host sum
    fi = O.inside(g.intf.f).astype(np.float64)
    assert abs(st.vol - fi.sum()) <= 1e-9 * fi.sum()
    assert abs(st.vol - V0) <= 1e-5 * V0
    g.close()


def test_refined_steps_non_cubic_match_oracle():
    """The benchmarked REFINED step on a grid of unequal sides, 96 x 48 x 64:
    the level-0 tiles are half filled along x, the column tilings differ per
    axis and level, and the coarse levels reach the 32-wide rows and the tail
    at different sizes per axis.  Three bubbles under a free surface at 0.7 of
    the height, slip walls, gravity -z, density ratio 1000; two whole steps
    against the FP32 oracle, with the 64³ pool test's bars."""
    dims = (96, 48, 64)
    say("REFINED whole steps on", dims)
    H = float(dims[2])
    bub = [(30.0, 20.0, 20.0, 7.0), (62.0, 28.0, 30.0, 5.5), (75.0, 15.0, 14.0, 4.0)]

    def sdf(x):
        s = x[2] - 0.7 * H
        for (cx, cy, cz, r) in bub:
            s = np.maximum(s, r - np.sqrt((x[0] - cx) ** 2 + (x[1] - cy) ** 2 + (x[2] - cz) ** 2))
        return s

    kw = dict(T=np.float32, λρ=1e-3, λμ=1e-2, ν=H / 1e4, g=(0.0, 0.0, -1.0 / H), InterfaceSDF=sdf)
    g = ia.TwoPhaseSimulation(dims, (0.0, 0.0, 0.0), H, poisson="refined", poisson_rtol=1e-6,
                              poisson_max_outer=200, **kw)
    o = O.OracleSim(dims, (0.0, 0.0, 0.0), H, poisson="refined", poisson_rtol=1e-6, **kw)
    assert float(np.abs(g.intf.f.astype(np.float64) - o.f.astype(np.float64)).max()) <= 1e-6
    o.set(O.F, g.intf.f)
    o.set(O.F0, g.intf.f0)
    o.Δt = [o.T.type(g.flow.Δt[-1])]
    for n in range(2):
        st = ia.sim_step_(g)
        o.sim_step()
        gi = [st.solve[0].iters, st.solve[1].iters]
        oi = [o.lib.oc_last_iters(o.ctx, 0), o.lib.oc_last_iters(o.ctx, 1)]
        df = float(np.abs(g.intf.f.astype(np.float64) - o.f.astype(np.float64)).max())
        du = rel_u(g.flow.u, o.u)
        say(f"step {n + 1}: max|Δf| {df:.3e}, max|Δu|/max|u| {du:.3e}, FCG its gpu {gi} oracle {oi}")
        assert st.status == 0
        assert df <= 1e-6
        assert du <= 1e-5
        for q in range(2):
            assert 0 < st.solve[q].rel <= 1e-6
            assert abs(gi[q] - oi[q]) <= 1
    g.close()


def test_pool64_fp32_error_vs_fp64_oracle():
    say("pool 64^3 REFINED FP32 GPU vs FP64 oracle, 5 steps")
    g, o = pool_pair(64, np.float64)
    errs = []
    for n in range(5):
        ia.sim_step_(g)
        o.sim_step()
        # keep the two runs on the same step sizes (the FP32 Δt)
        o.Δt[-1] = np.float64(g.flow.Δt[-1])
        df = float(np.abs(g.intf.f.astype(np.float64) - o.f).max())
        du = rel_u(g.flow.u, o.u)
        errs.append((df, du))
        say(f"step {n + 1}: max|Δf| {df:.3e}, max|Δu|/max|u| {du:.3e}")
        if n in (0, 4):
            assert df <= 2e-5
            assert du <= 1e-3
    g.close()


def test_pool128_coarse_cg_workgroup_sizes_match_oracle():
    """The coarse CG step's workgroup size per level (IA_CG1_NB: 512 threads
    on levels above 100k cells, else 256, by default) changes only the order
    in which its partial sums add.  At 128³ the 64³ and 32³ levels run the
    one-cell-per-thread step; one whole REFINED step with the default sizes
    and with 1024/128, 256/256 and 128/1024 threads on those levels, each
    against the same FP32 oracle step, with the 64³ test's bars."""
    ref = None
    for nb in (None, "1024,128", "256,256", "128,1024"):  # levels 1 (64³) and 2 (32³)
        say("pool 128^3 REFINED, one step, IA_CG1_NB =", nb or "default")
        env_old = os.environ.get("IA_CG1_NB")
        if nb is not None:
            os.environ["IA_CG1_NB"] = nb
        try:
            g, _, kw = pool.build_sim(ia, 128, "refined", 1e-6)
        finally:
            if env_old is None:
                os.environ.pop("IA_CG1_NB", None)
            else:
                os.environ["IA_CG1_NB"] = env_old
        if ref is None:
            okw = dict(kw)
            okw["T"] = np.float32
            o = O.OracleSim((128,) * 3, (0.0, 0.0, 0.0), 128.0, poisson="refined", poisson_rtol=1e-6, **okw)
            f_in = g.intf.f.copy()
            o.set(O.F, f_in)
            o.set(O.F0, g.intf.f0)
            o.Δt = [o.T.type(g.flow.Δt[-1])]
            o.sim_step()
            ref = (o.f.astype(np.float64), np.asarray(o.u, dtype=np.float64),
                   [o.lib.oc_last_iters(o.ctx, 0), o.lib.oc_last_iters(o.ctx, 1)])
            del o
        else:
            assert np.array_equal(g.intf.f, f_in)
        st = ia.sim_step_(g)
        gi = [st.solve[0].iters, st.solve[1].iters]
        df = float(np.abs(g.intf.f.astype(np.float64) - ref[0]).max())
        du = rel_u(g.flow.u, ref[1])
        say(f"max|Δf| {df:.3e}, max|Δu|/max|u| {du:.3e}, FCG its gpu {gi} oracle {ref[2]}")
        assert st.status == 0
        assert df <= 1e-6
        assert du <= 1e-5
        for q in range(2):
            assert 0 < st.solve[q].rel <= 1e-6
            assert abs(gi[q] - ref[2][q]) <= 1
        g.close()


def test_pool256_refined_properties():
    say("pool 256^3 (32 bubbles, seed 3) REFINED, 5 steps, properties")
    g, _, _ = pool.build_sim(ia, 256, "refined", 1e-6)
    V0 = None
    for n in range(5):
        st = ia.sim_step_(g)
        assert st.status == 0
        for q in range(2):
            assert 0 < st.solve[q].rel <= 1e-6
        assert 0.0 <= st.fmin and st.fmax <= 1.0
        if V0 is None:
            f = O.inside(g.intf.f0).astype(np.float64)  # f before the first step
            V0 = f.sum()
            del f
        say(f"step {n + 1}: Σf {st.vol:.10e} (drift {(st.vol - V0) / V0:.2e}), f in [{st.fmin}, {st.fmax}], "
            f"FCG its {st.solve[0].iters}/{st.solve[1].iters}")
        assert abs(st.vol - V0) <= 1e-8 * V0
    f = g.intf.f
    assert np.isfinite(f).all() and f.min() >= 0 and f.max() <= 1
    assert np.isfinite(g.flow.u).all()
    g.close()
