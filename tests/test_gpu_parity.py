"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on
identical inputs, stage by stage and for whole steps.

Bars (stated per assertion):
- FP64: every field within 1e-12 (VOF) / 1e-10 (velocity, pressure after the
  iterative solver, whose dot products are summed in a different order);
- FP32: VOF within 1e-5 of the FP32 oracle (same operation order, no FMA
  contraction on either side; the transcendental functions of the PLIC
  inverse may differ by an ulp), velocity/pressure as stated per test.
Integer/byte fields (c̄) are bit-exact.
"""
import numpy as np
import pytest

from __graft_entry__ import load_package
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ia = load_package()


def say(*a):
    print("[case]", *a, flush=True)


# FP64 f after whole steps: the GPU and the oracle sum the Poisson solver's
# dot products in different orders, so their velocities differ at the solver's
# round-off and f, advected by them, at about 1e-11.  A wrong flux, normal or
# limiter moves f by far more.
F_STEP_TOL = 1e-10
# FP64 u after whole steps whose solves run WaterLily's solver! to an
# unreachable Σr² (200 V-cycles, stagnating at 1e-12 to 1e-11): pcg!'s early
# exits (|α| < 1e-2, ρ < 10eps) are discrete decisions that round-off flips,
# so two summation orders of the same solver end up to 2.3e-5 of max|u| apart
# after one step (measured: the GPU against the oracle on these cases).  f
# stays within F_STEP_TOL and carries the plug-ins' effect from the second
# step on.
U_STAG_TOL = 1e-4

def maxdiff(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max())


def tgv_droplet(N=16, T=np.float64, λμ=0.1, λρ=0.1, **kw):
    """helper.jl:6-13 (TGV droplet, periodic)."""
    κ = 2 * np.pi / N
    R = N / 4
    u0 = lambda i, x: (-np.sin(x[0] * κ) * np.cos(x[1] * κ)) if i == 1 else (np.cos(x[0] * κ) * np.sin(x[1] * κ))
    sdf = lambda x: np.sqrt((x[0] - N / 2) ** 2 + (x[1] - N / 2) ** 2) - R
    return dict(dims=(N, N), u_BC=(0.0, 0.0), L=float(N), kw=dict(T=T, λμ=λμ, λρ=λρ, InterfaceSDF=sdf, u0=u0,
                                                                   perdir=(1, 2), **kw))


def quiescent(N=8, T=np.float64):
    """helper.jl:17-18."""
    sdf = lambda x: np.sqrt((x[0] - N / 2) ** 2 + (x[1] - N / 2) ** 2) - N / 4
    return dict(dims=(N, N), u_BC=(0.0, 0.0), L=float(N), kw=dict(T=T, InterfaceSDF=sdf, perdir=(1, 2)))


def bubble2d(N=64, T=np.float32, ν=None, g0=None):
    """Config-2 shape at a small size: free surface with a cosine bump and one bubble, slip walls."""
    L = float(N)
    ν = L / 1e4 if ν is None else ν
    g0 = 1.0 / L if g0 is None else g0

    def sdf(x):
        h = 0.5 * L + 0.05 * L * np.cos(2 * np.pi * x[0] / L)
        return np.maximum(x[1] - h, 0.1 * L - np.sqrt((x[0] - 0.5 * L) ** 2 + (x[1] - 0.25 * L) ** 2))

    return dict(dims=(N, N), u_BC=(0.0, 0.0), L=L, kw=dict(T=T, λρ=1e-3, λμ=1e-2, ν=ν, g=(0.0, -g0),
                                                          InterfaceSDF=sdf))


def bubbles3d(N=32, T=np.float32, seed=3, nb=4):
    L = float(N)
    rng = np.random.default_rng(seed)
    cs = [(rng.uniform(0.25, 0.75) * L, rng.uniform(0.25, 0.75) * L, rng.uniform(0.2, 0.45) * L,
           rng.uniform(0.06, 0.1) * L) for _ in range(nb)]

    def sdf(x):
        s = x[2] - 0.7 * L
        for (cx, cy, cz, r) in cs:
            s = np.maximum(s, r - np.sqrt((x[0] - cx) ** 2 + (x[1] - cy) ** 2 + (x[2] - cz) ** 2))
        return s

    return dict(dims=(N, N, N), u_BC=(0.0, 0.0, 0.0), L=L, kw=dict(T=T, λρ=1e-3, λμ=1e-2, ν=L / 1e4,
                                                                     g=(0.0, 0.0, -1.0 / L), InterfaceSDF=sdf))


def pair(case, **extra):
    kw = dict(case["kw"])
    kw.update(extra)
    okw = {k: v for k, v in kw.items() if k not in ("poisson", "poisson_rtol", "poisson_max_outer")}
    gsim = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **kw)
    osim = O.OracleSim(case["dims"], case["u_BC"], case["L"], **okw)
    return gsim, osim


# --------------------------------------------------------------------------- construction
@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("D", [2, 3])
def test_apply_vof_and_cfl(T, D):
    say("applyVOF!+MPCFL", T.__name__, f"{D}D")
    case = bubble2d(32, T) if D == 2 else bubbles3d(16, T)
    g, o = pair(case)
    tol = 1e-12 if T == np.float64 else 1e-6
    assert maxdiff(g.intf.f, o.f) <= tol
    assert (g.intf.f0 == g.intf.f).all()
    assert g.flow.Δt[-1] == o.Δt[-1]
    g.close()


# --------------------------------------------------------------------------- the VOF sweep
def _swirl(case, T, amp=0.6, seed=0):
    """A discretely divergence-free face field (corner streamfunction in the x-y plane)."""
    dims = case["dims"]
    D = len(dims)
    Ng = [n + 2 for n in dims]
    x = np.arange(Ng[0]) - 1.0
    y = np.arange(Ng[1]) - 1.0
    X, Y = np.meshgrid(x, y, indexing="ij")
    psi = amp * np.sin(np.pi * X / dims[0]) ** 2 * np.sin(np.pi * Y / dims[1]) ** 2 * dims[0] / np.pi
    u = np.zeros(tuple(Ng) + (D,))
    ux = np.zeros((Ng[0], Ng[1]))
    uy = np.zeros((Ng[0], Ng[1]))
    ux[:, :-1] = -(psi[:, 1:] - psi[:, :-1])
    uy[:-1, :] = psi[1:, :] - psi[:-1, :]
    if D == 2:
        u[..., 0], u[..., 1] = ux, uy
    else:
        zz = 1.0 + 0.3 * np.cos(np.pi * (np.arange(Ng[2]) - 0.5) / dims[2])
        u[..., 0] = ux[..., None] * zz
        u[..., 1] = uy[..., None] * zz
        u[..., 2] = 0.05 * np.sin(2 * np.pi * (X[..., None] + 0.5) / dims[0]) * np.ones(Ng[2])
    return np.asfortranarray(u, dtype=T)


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("per", [False, True])
def test_vof_sweep(T, D, per):
    say("advectVOF1d!", T.__name__, f"{D}D", "periodic" if per else "walls")
    case = bubble2d(32, T) if D == 2 else bubbles3d(16, T)
    if per:
        case["kw"]["perdir"] = tuple(range(1, D + 1))
    g, o = pair(case)
    u = _swirl(case, T)
    u0 = np.asfortranarray(u * T(0.9))
    g.flow.u = u
    g.flow.u0 = u0
    o.set(O.U, u)
    o.set(O.U0, u0)
    tol = 1e-12 if T == np.float64 else 1e-5
    for d in range(1, D + 1):
        rc = g._call("ia_vof_sweep", d, 0.5, None)
        ro = o.lib.oc_vof_sweep(o.ctx, d, 0.5)
        assert rc == ro
        assert maxdiff(g.intf.f, o.f) <= tol, d
        assert (g.intf.cbar == o.get(O.CBAR)).all()
        rg, ro_ = g.intf.rhouf[..., d - 1], o.get(O.RHOUF)[..., d - 1]
        inner = tuple(slice(1, None if a == d - 1 else -1) for a in range(D))
        assert maxdiff(rg[inner], ro_[inner]) <= tol * 10
    g.close()


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_advect_single_vortex(T):
    """advect! (pure cVOF, the config-1 call) for 24 steps on a 64² periodic grid."""
    say("advect! single vortex 64^2", T.__name__)
    N = 64
    sdf = lambda x: np.sqrt((x[0] - 0.5 * N) ** 2 + (x[1] - 0.75 * N) ** 2) - 0.15 * N
    case = dict(dims=(N, N), u_BC=(0.0, 0.0), L=float(N), kw=dict(T=T, InterfaceSDF=sdf, perdir=(1, 2)))
    g, o = pair(case)
    u = _swirl(case, T, amp=1.0)
    g.flow.u = u
    g.flow.u0 = u
    o.set(O.U, u)
    o.set(O.U0, u)
    V0 = O.inside(o.f).astype(np.float64).sum()
    for n in range(1, 25):
        st = ia.advect_(g, dt=0.25)
        g.flow.Δt.append(g.T.type(0.25))
        o.advect(0.25, len(g.flow.Δt) - 1)
    fg = g.intf.f
    tol = 1e-12 if T == np.float64 else 1e-5
    assert maxdiff(fg, o.f) <= tol
    assert fg.min() >= 0 and fg.max() <= 1
    assert abs(O.inside(fg).astype(np.float64).sum() - V0) <= (1e-12 if T == np.float64 else 1e-5) * V0
    # the call's statistics: FP64 Σf and the extrema of the advected field
    fi = O.inside(fg).astype(np.float64)
    assert abs(st.vol - fi.sum()) <= 1e-12 * fi.size
    assert st.fmin == fi.min() and st.fmax == fi.max()
    g.close()


def channel(N=(48, 24), T=np.float64, exitBC=True):
    """A channel of the dark fluid with inflow u_BC = (1, 0) at x = 0, the exit
    at the upper x face (exitBC), slip walls in y (and z), gravity across it,
    density ratio 1000: a light bubble near the exit, carried out through it."""
    D = len(N)
    L = float(N[1])
    c = [0.85 * N[0], 0.5 * N[1], 0.5 * N[2] if D == 3 else 0.0]

    def sdf(x):  # dark outside the bubble
        return 0.3 * L - np.sqrt(sum((x[a] - c[a]) ** 2 for a in range(D)))

    uλ = lambda i, x: np.ones_like(x[0]) if i == 1 else np.zeros_like(x[0])
    g = (0.0, -0.5 / L) + ((0.0,) if D == 3 else ())
    return dict(dims=N, u_BC=(1.0,) + (0.0,) * (D - 1), L=L,
                kw=dict(T=T, λρ=1e-3, λμ=1e-2, ν=L / 1e3, g=g, InterfaceSDF=sdf, u0=uλ, exitBC=exitBC))


@pytest.mark.parametrize("D", [2, 3])
def test_exitbc_steps_match_oracle(D):
    """exitBC (WaterLily BC! leaving the exit face of u_x alone; flow.jl:69,
    :79, :82, :197, :207): whole FP64 steps of a light bubble leaving a
    channel, on the device and the oracle (REFERENCE solve).  The exit face
    keeps u_x = 1 (only gravity's x component, 0 here, would update it), the
    inflow faces stay at u_BC, and the oracle's run without exitBC differs by
    far more than the parity bar (there the exit's mass flux is reset to u_BC,
    which the light fluid leaving cannot carry), so the exit path is
    exercised."""
    say("exitBC channel", f"{D}D")
    N = (48, 24) if D == 2 else (32, 16, 16)
    case = channel(N)
    g, o = pair(case)
    plain = O.OracleSim(case["dims"], case["u_BC"], case["L"], **dict(case["kw"], exitBC=False))
    nst = 8 if D == 2 else 4
    for n in range(nst):
        ia.sim_step_(g)
        o.sim_step()
        plain.sim_step()
        assert maxdiff(g.intf.f, o.f) <= F_STEP_TOL
        assert maxdiff(g.flow.u, o.u) <= U_STAG_TOL * max(1.0, np.abs(o.u).max())
        assert g.flow.Δt[-1] == pytest.approx(o.Δt[-1], rel=1e-12)
    u = g.flow.u
    assert np.all(u[-1, ..., 0] == 1.0)          # the exit face of u_x
    assert np.all(u[0, ..., 0] == 1.0) and np.all(u[1, ..., 0] == 1.0)   # the inflow
    assert np.abs(u).max() < 1.2
    assert maxdiff(o.f, plain.f) > 100 * F_STEP_TOL
    g.close()


def test_exitbc_rejects_gravity_along_x():
    case = channel((32, 16))
    kw = dict(case["kw"], g=(0.1, -0.01))
    with pytest.raises(ValueError):
        ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_divergent_squeeze_explodes(T):
    """reportFillError's explosion (advection.jl:160): a column of f = 0.4
    between full columns, squeezed by u_x = ±U from both sides with δl = 0.4.
    f reaches 1.2 where |∇·u⁰| + |∇·u| = 4U: U = 3 (12 > 10) raises the
    divergence error (status 2), U = 2.5 (exactly 10) only warns (status 3),
    as on the oracle."""
    N = 16
    for U, want in ((3.0, 2), (2.5, 3)):
        say("explosion squeeze", T.__name__, U)
        sim = ia.TwoPhaseSimulation((N, N), (0.0, 0.0), float(N), T=T, perdir=(1, 2))
        o = O.OracleSim((N, N), (0.0, 0.0), float(N), T=T, perdir=(1, 2))
        f = np.ones(sim.Ng, dtype=T, order="F")
        f[8, :] = 0.4
        u = np.zeros(sim.Ng + (2,), dtype=T, order="F")
        u[8, :, 0] = U
        u[9, :, 0] = -U
        for a in (sim, ):
            a.intf.f = f
            a.flow.u = u
            a.flow.u0 = u
        o.set(O.F, f)
        o.set(O.U, u)
        o.set(O.U0, u)
        dt = 0.4 / U
        assert o.advect(dt, 1) == want
        if want == 2:
            with pytest.raises(ia.IAError) as e:
                ia.advect_(sim, dt=dt)
            assert e.value.code == 2 and "exploding" in str(e.value)
        else:
            st = ia.advect_(sim, dt=dt)
            assert st.status == 3 and st.fill_warnings == 1
            assert maxdiff(sim.intf.f, o.f) <= (1e-12 if T == np.float64 else 1e-6)
        sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_advect_named_fields(T):
    """advect!(a, c, f, u¹, u², dt) with f = c.f⁰ and the velocities taken from
    Flow.f and cVOF.ρu (advection.jl:17): the same sweeps as the oracle's
    advect! of that f with those velocities; c.f is left untouched."""
    sdf = lambda x: np.sqrt((x[0] - 12.3) ** 2 + (x[1] - 19.1) ** 2) - 6.2
    case = dict(dims=(32, 32), u_BC=(0.0, 0.0), L=32.0, kw=dict(T=T, InterfaceSDF=sdf, perdir=(1, 2)))
    g, o = pair(case)
    u = _swirl(case, T, amp=1.0, seed=3)
    w = _swirl(case, T, amp=0.7, seed=4)
    f_before = g.intf.f
    g._set("force", u)
    g._set("rhou", w)
    g.intf.f0 = f_before  # the field to advect
    o.set(O.U0, u)
    o.set(O.U, w)
    for n in range(1, 6):
        g.flow.Δt.append(g.T.type(0.25))
        st = ia.advect_(g, f="f0", u1="force", u2="rhou", dt=0.25)
        o.advect(0.25, len(g.flow.Δt))
    tol = 1e-12 if T == np.float64 else 1e-5
    assert maxdiff(g.intf.f0, o.f) <= tol
    assert np.array_equal(g.intf.f, f_before)
    fi = O.inside(g.intf.f0).astype(np.float64)
    assert abs(st.vol - fi.sum()) <= 1e-12 * fi.size
    assert st.fmin == fi.min() and st.fmax == fi.max()
    with pytest.raises(ia.IAError):
        g._call("ia_advect", 12, 9, 10, 0.25, 1, None)  # p is not an advectable field
    g.close()


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("D", [2, 3])
def test_reconstruct(T, D):
    say("reconstructInterface!", T.__name__, f"{D}D")
    case = bubble2d(32, T) if D == 2 else bubbles3d(16, T)
    case["kw"]["perdir"] = (1,)
    g, o = pair(case)
    g._call("ia_reconstruct")
    o.lib.oc_reconstruct(o.ctx)
    tol = 1e-12 if T == np.float64 else 2e-5
    assert maxdiff(O.inside(g.intf.alpha), O.inside(o.get(O.ALPHA))) <= tol
    assert maxdiff(O.inside(g.intf.nhat, D), O.inside(o.get(O.NHAT), D)) <= tol
    g.close()


# --------------------------------------------------------------------------- viscosity, coefficients, solve
@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("D", [2, 3])
def test_visc_and_coefficients(T, D):
    say("visc!/updateL!/update!", T.__name__, f"{D}D")
    case = bubble2d(32, T, ν=0.05) if D == 2 else bubbles3d(16, T)
    g, o = pair(case)
    u = _swirl(case, T)
    g.flow.u = u
    o.set(O.U, u)
    g._call("ia_visc", 0)
    o.lib.oc_visc(o.ctx, 0)
    Fg, Fo = O.inside(g.flow.force, D), O.inside(o.get(O.FORCE), D)
    scale = max(1e-30, np.abs(Fo).max())
    assert maxdiff(Fg, Fo) <= (1e-12 if T == np.float64 else 2e-5) * scale
    g._call("ia_update_L", 0)
    o.lib.oc_update_L(o.ctx, 0)
    assert maxdiff(g.flow.mu0, o.get(O.MU0)) <= (1e-12 if T == np.float64 else 1e-6) * 1000
    g.close()


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("per", [(), (1,)])
def test_project_reference(T, per):
    say("myproject! (WaterLily solver!)", T.__name__, per)
    case = bubble2d(32, T)
    case["kw"]["perdir"] = per
    g, o = pair(case)
    rng = np.random.default_rng(7)
    u = np.zeros(g.Ng + (2,), dtype=T, order="F")
    u[1:-1, 1:-1] = rng.standard_normal((32, 32, 2)) * 0.1
    g._call("ia_set_initial_velocity", ia._ptr(u), u.nbytes)
    o.lib.oc_set_initial_velocity(o.ctx, O._ptr(u))
    for s in ("ia_update_L",):
        g._call(s, 0)
    g._call("ia_mg_update")
    o.lib.oc_update_L(o.ctx, 0)
    o.lib.oc_mg_update(o.ctx)
    st = ia.IASolveStats()
    g._call("ia_project", 0.5, ia.C.byref(st))
    it, r2 = o.project(0.5)
    assert abs(st.iters - it) <= 1
    tol = 1e-9 if T == np.float64 else 2e-3
    pg, po = O.inside(g.flow.p), O.inside(o.p)
    assert maxdiff(pg - pg.mean(), po - po.mean()) <= tol * max(1.0, np.abs(po).max())
    assert maxdiff(g.flow.u, o.u) <= tol
    g.close()


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("D, per", [(2, ()), (2, (1, 2)), (3, ()), (3, (1, 2, 3))])
def test_project_psolver(T, D, per):
    """inproject!(::Poisson) + psolver! (flow.jl:300-341): the single-level
    Jacobi-PCG to Σr² <= 50eps(T), against the oracle's.  The dot products sum
    in another order on each side, so the iteration counts may differ by a few
    (FP64: hundreds of iterations to 1e-14); both stop at the tolerance, and p
    and u agree to the converged solution's round-off (FP64: 1e-9 relative to
    max|p|, FP32: 2e-3)."""
    say("psolver!", T.__name__, f"{D}D", per)
    case = bubble2d(32, T) if D == 2 else bubbles3d(16, T)
    case["kw"]["perdir"] = per
    g = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], poisson="pcg", **case["kw"])
    o = O.OracleSim(case["dims"], case["u_BC"], case["L"], poisson="pcg", **case["kw"])
    rng = np.random.default_rng(8)
    N = case["dims"]
    u = np.zeros(g.Ng + (D,), dtype=T, order="F")
    u[(slice(1, -1),) * D] = rng.standard_normal(tuple(N) + (D,)) * 0.1
    g._call("ia_set_initial_velocity", ia._ptr(u), u.nbytes)
    o.lib.oc_set_initial_velocity(o.ctx, O._ptr(u))
    g._call("ia_update_L", 0)
    g._call("ia_mg_update")
    o.lib.oc_update_L(o.ctx, 0)
    o.lib.oc_mg_update(o.ctx)
    st = ia.IASolveStats()
    g._call("ia_project", 0.5, ia.C.byref(st))
    it, r2 = o.project(0.5)
    eps = np.finfo(T).eps
    say(f"iterations gpu {st.iters} oracle {it}, Σr² gpu {st.r2:.3e} oracle {r2:.3e}")
    assert st.r2 <= 50 * eps and r2 <= 50 * eps
    assert abs(st.iters - it) <= max(3, it // 20)
    tol = 1e-9 if T == np.float64 else 2e-3
    pg, po = O.inside(g.flow.p), O.inside(o.p)
    assert maxdiff(pg - pg.mean(), po - po.mean()) <= tol * max(1.0, np.abs(po).max())
    assert maxdiff(g.flow.u, o.u) <= tol
    g.close()


def test_psolver_steps_match_oracle():
    """Whole MPFMomStep!s with the single-level Poisson (FP64 TGV droplet)."""
    say("MPFMomStep! with psolver!, TGV droplet FP64")
    case = tgv_droplet(16, np.float64)
    g = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], poisson="pcg", **case["kw"])
    o = O.OracleSim(case["dims"], case["u_BC"], case["L"], poisson="pcg", **case["kw"])
    for _ in range(3):
        ia.sim_step_(g)
        o.sim_step()
        assert maxdiff(g.intf.f, o.f) <= F_STEP_TOL
        assert maxdiff(g.flow.u, o.u) <= 1e-9 * max(1.0, np.abs(o.u).max())
        # psolver! stops at Σr² ≤ 50eps, where the two sides' dot products
        # (summed in different orders) differ in the last bits: u agrees to
        # 1e-9 relative, and so does the CFL Δt computed from it.
        assert g.flow.Δt[-1] == pytest.approx(o.Δt[-1], rel=1e-9)
    g.close()


@pytest.mark.parametrize("D", [2, 3])
def test_psolver_walls_steps_match_oracle(D):
    """psolver! with slip walls, gravity and density ratio 1000 (FP64): two
    whole steps of bubbles under a free surface.  psolver! stops at
    Σr² ≤ 50eps, where one iteration more or less (the two sides sum their dot
    products in different orders) moves u by about 1e-9 of max|u| and f by
    about 1e-9 (measured on the oracle alone, perturbing the TGV droplet's
    initial u by 1e-15 relative): bars 1e-8."""
    say("MPFMomStep! with psolver!, walls + gravity FP64", f"{D}D")
    case = bubble2d(24, np.float64) if D == 2 else bubbles3d(16, np.float64)
    g = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], poisson="pcg", **case["kw"])
    o = O.OracleSim(case["dims"], case["u_BC"], case["L"], poisson="pcg", **case["kw"])
    for _ in range(2):
        ia.sim_step_(g)
        o.sim_step()
        assert maxdiff(g.intf.f, o.f) <= 1e-8
        assert maxdiff(g.flow.u, o.u) <= 1e-8 * max(1.0, np.abs(o.u).max())
        assert g.flow.Δt[-1] == pytest.approx(o.Δt[-1], rel=1e-8)
    assert np.abs(o.u).max() > 0  # gravity has moved the fluid
    g.close()


@pytest.mark.parametrize("per", [(), (1, 2, 3)])
def test_project_reference_3d_tiled(per):
    """WaterLily's solver! in FP64 at 96^3: level 0 takes the LDS-tiled
    kernels (4-cell columns; the second x tile is half filled), which the FP64
    projection otherwise never reaches.  Velocity scale 0.05 ends both solves
    at Σr² ≈ 9.25e-5 (35 and 31 V-cycles at about 0.93 per cycle), well clear
    of the 1e-4 stop, so round-off cannot move the iteration count."""
    N = 96
    say(f"myproject! 3D {N}^3 float64", "periodic" if per else "walls")
    case = bubbles3d(N, np.float64)
    if per:
        case["kw"]["perdir"] = per
        case["kw"]["g"] = None
    g, o = pair(case)
    rng = np.random.default_rng(11)
    u = np.zeros(g.Ng + (3,), dtype=np.float64, order="F")
    u[1:-1, 1:-1, 1:-1] = rng.standard_normal((N, N, N, 3)) * 0.05
    g._call("ia_set_initial_velocity", ia._ptr(u), u.nbytes)
    o.lib.oc_set_initial_velocity(o.ctx, O._ptr(u))
    g._call("ia_update_L", 0)
    g._call("ia_mg_update")
    o.lib.oc_update_L(o.ctx, 0)
    o.lib.oc_mg_update(o.ctx)
    st = ia.IASolveStats()
    g._call("ia_project", 0.5, ia.C.byref(st))
    it, r2 = o.project(0.5)
    print(f"gpu {st.iters} V-cycles, oracle {it} (Σr² {r2:.3e})")
    assert st.iters == it
    pg, po = O.inside(g.flow.p), O.inside(o.p)
    assert maxdiff(pg - pg.mean(), po - po.mean()) <= 1e-9 * max(1.0, np.abs(po).max())
    assert maxdiff(g.flow.u, o.u) <= 1e-9
    g.close()


@pytest.mark.parametrize("N", [64, 256])
def test_project_refined_reaches_1e6(N):
    """At 256^2 the 128^2 level runs the single-reduction CG (above the tail)."""
    say(f"FP64 flexible CG + FP32 V-cycle, {N}^2")
    case = bubble2d(N, np.float32)
    g, o = pair(case, poisson="refined", poisson_rtol=1e-6, poisson_max_outer=100)
    rng = np.random.default_rng(8)
    u = np.zeros(g.Ng + (2,), dtype=np.float32, order="F")
    u[1:-1, 1:-1] = rng.standard_normal((N, N, 2))
    g._call("ia_set_initial_velocity", ia._ptr(u), u.nbytes)
    o.lib.oc_set_initial_velocity(o.ctx, O._ptr(u))
    g._call("ia_update_L", 0)
    g._call("ia_mg_update")
    o.lib.oc_update_L(o.ctx, 0)
    o.lib.oc_mg_update(o.ctx)
    st = ia.IASolveStats()
    g._call("ia_project", 1.0, ia.C.byref(st))
    outer, rel = o.project_fcg(1.0, 1e-6, 100, 2)
    assert st.rel <= 1e-6 and rel <= 1e-6
    assert abs(st.iters - outer) <= 3
    pg, po = O.inside(g.flow.p).astype(np.float64), O.inside(o.p).astype(np.float64)
    assert maxdiff(pg - pg.mean(), po - po.mean()) <= 1e-4 * np.abs(po).max()
    g.close()


@pytest.mark.parametrize("per,N", [(False, 32), (True, 32), (False, 128), (True, 128)])
def test_project_refined_3d(per, N):
    """3D takes the LDS-tiled solver kernels; walls and fully periodic.  At 128^3
    the preconditioner's coarse levels of 64^3 and 32^3 run 48 and 64 iterations
    of the single-reduction CG outside the one-workgroup tail (16^3 and below,
    32 pcg! iterations each)."""
    say(f"FCG + V-cycle 3D {N}^3", "periodic" if per else "walls")
    case = bubbles3d(N, np.float32)
    if per:
        case["kw"]["perdir"] = (1, 2, 3)
        case["kw"]["g"] = None
    g, o = pair(case, poisson="refined", poisson_rtol=1e-6, poisson_max_outer=100)
    rng = np.random.default_rng(10)
    u = np.zeros(g.Ng + (3,), dtype=np.float32, order="F")
    u[1:-1, 1:-1, 1:-1] = rng.standard_normal((N, N, N, 3))
    g._call("ia_set_initial_velocity", ia._ptr(u), u.nbytes)
    o.lib.oc_set_initial_velocity(o.ctx, O._ptr(u))
    g._call("ia_update_L", 0)
    g._call("ia_mg_update")
    o.lib.oc_update_L(o.ctx, 0)
    o.lib.oc_mg_update(o.ctx)
    st = ia.IASolveStats()
    g._call("ia_project", 1.0, ia.C.byref(st))
    outer, rel = o.project_fcg(1.0, 1e-6, 100, 2)
    print(f"{N}^3: gpu {st.iters} iterations (rel {st.rel:.2e}), oracle {outer} (rel {rel:.2e})")
    assert st.rel <= 1e-6 and rel <= 1e-6
    assert abs(st.iters - outer) <= 3
    pg, po = O.inside(g.flow.p).astype(np.float64), O.inside(o.p).astype(np.float64)
    assert maxdiff(pg - pg.mean(), po - po.mean()) <= 1e-4 * np.abs(po).max()
    ug, uo = g.flow.u.astype(np.float64), o.get(O.U).astype(np.float64)
    assert maxdiff(O.inside(ug, 3), O.inside(uo, 3)) <= 1e-4 * np.abs(uo).max()
    g.close()


def test_project_plain_multigrid_reaches_1e6():
    say("plain multigrid (FP32 V-cycles in FP64 refinement), 64^2")
    case = bubble2d(64, np.float32)
    g, o = pair(case, poisson="plain", poisson_rtol=1e-6, poisson_max_outer=200)
    rng = np.random.default_rng(9)
    u = np.zeros(g.Ng + (2,), dtype=np.float32, order="F")
    u[1:-1, 1:-1] = rng.standard_normal((64, 64, 2))
    g._call("ia_set_initial_velocity", ia._ptr(u), u.nbytes)
    o.lib.oc_set_initial_velocity(o.ctx, O._ptr(u))
    g._call("ia_update_L", 0)
    g._call("ia_mg_update")
    o.lib.oc_update_L(o.ctx, 0)
    o.lib.oc_mg_update(o.ctx)
    st = ia.IASolveStats()
    g._call("ia_project", 1.0, ia.C.byref(st))
    outer, rel = o.project_refined(1.0, 1e-6, 200)
    assert st.rel <= 1e-6 and rel <= 1e-6
    assert abs(st.iters - outer) <= 3
    pg, po = O.inside(g.flow.p).astype(np.float64), O.inside(o.p).astype(np.float64)
    assert maxdiff(pg - pg.mean(), po - po.mean()) <= 1e-4 * np.abs(po).max()
    g.close()


# --------------------------------------------------------------------------- whole steps
def test_quiescent_step_is_noop():
    """maintests.jl:193-202 on the GPU path."""
    say("quiescent droplet 8^2 FP64, sim_step!")
    case = quiescent()
    g = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **case["kw"])
    f0 = g.intf.f
    ia.sim_step_(g)
    assert (g.flow.u == 0).all()
    assert (g.intf.f == f0).all()
    assert np.isfinite(g.flow.p).all()
    g.close()


def test_tgv_droplet_volume():
    """maintests.jl:204-215 on the GPU path."""
    say("TGV droplet 16^2 FP64, 4 steps")
    case = tgv_droplet()
    g = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **case["kw"])
    ia.sim_step_(g)
    V0 = ia.inside(g.intf.f).sum()
    for _ in range(3):
        ia.sim_step_(g)
    assert np.isfinite(g.flow.u).all() and np.isfinite(g.intf.f).all()
    assert abs(ia.inside(g.intf.f).sum() - V0) <= 1e-4 * abs(V0)
    g.close()


def test_tgv_droplet_matches_oracle():
    say("TGV droplet 16^2 FP64 vs oracle, 4 steps")
    g, o = pair(tgv_droplet())
    for n in range(4):
        ia.sim_step_(g)
        o.sim_step()
        assert maxdiff(g.intf.f, o.f) <= F_STEP_TOL, n
        assert maxdiff(g.flow.u, o.u) <= 1e-9, n
        assert abs(g.flow.Δt[-1] - o.Δt[-1]) <= 1e-12
    pg, po = O.inside(g.flow.p), O.inside(o.p)
    assert maxdiff(pg - pg.mean(), po - po.mean()) <= 1e-8 * max(1, np.abs(po).max())
    g.close()


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_bubble2d_steps_match_oracle(T):
    say("free surface + bubble 64^2 vs oracle, 3 steps", T.__name__)
    g, o = pair(bubble2d(64, T))
    for n in range(3):
        ia.sim_step_(g)
        o.sim_step()
        ftol = F_STEP_TOL if T == np.float64 else 1e-5
        assert maxdiff(g.intf.f, o.f) <= ftol, n
        umax = max(1e-12, np.abs(o.u).max())
        assert maxdiff(g.flow.u, o.u) <= (1e-8 if T == np.float64 else 2e-3) * umax, n
    g.close()


def test_bubbles3d_step_matches_oracle():
    say("3D pool + 4 bubbles 32^3 FP32 vs oracle, 2 steps")
    g, o = pair(bubbles3d(32, np.float32))
    for n in range(2):
        ia.sim_step_(g)
        o.sim_step()
        assert maxdiff(g.intf.f, o.f) <= 1e-5, n
        umax = max(1e-12, np.abs(o.u).max())
        assert maxdiff(g.flow.u, o.u) <= 2e-3 * umax, n
    g.close()


def test_uniform_flow_stays_uniform():
    """CMOM consistency on the GPU: uniform u across a 1000:1 droplet stays uniform."""
    say("uniform flow 16^2 FP64")
    N = 16
    Uv = (0.37, -0.21)
    sdf = lambda x: np.sqrt((x[0] - 8.2) ** 2 + (x[1] - 7.7) ** 2) - 4.1
    g = ia.TwoPhaseSimulation((N, N), (0.0, 0.0), T=np.float64, InterfaceSDF=sdf,
                              u0=lambda i, x: np.full_like(x[0], Uv[i - 1]), perdir=(1, 2), λρ=1e-3)
    for _ in range(3):
        ia.sim_step_(g)
    u = ia.inside(g.flow.u, 2)
    for i in range(2):
        assert np.abs(u[..., i] - Uv[i]).max() < 1e-12
    g.close()


def test_nan_is_reported():
    """reportFillError: a NaN in f raises (advection.jl:148)."""
    say("NaN injection 16^2")
    case = tgv_droplet()
    g = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **case["kw"])
    f = g.intf.f
    f[5, 5] = np.nan
    g.intf.f = f
    with pytest.raises(ia.IAError) as e:
        ia.sim_step_(g)
    assert e.value.code == 1
    g.close()


@pytest.mark.parametrize("D", [2, 3])
def test_fill_report_extremes_and_ties(D):
    """reportFillError (advection.jl:145-189): the sweep reports the largest and
    smallest f before cleanWisp! and, on ties, the first cell in memory order.
    Still fluid, so the sweep leaves f as it was; two cells share the maximum
    and two the minimum, each pair straddling wavefronts and blocks."""
    say("fill report", f"{D}D")
    case = bubble2d(32, np.float64) if D == 2 else bubbles3d(16, np.float64)
    g = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **case["kw"])
    g.flow.u = np.zeros_like(g.flow.u)
    g.flow.u0 = np.zeros_like(g.flow.u0)
    f = g.intf.f
    # interior cells, the second of each pair first in memory
    hi = [(3, 7), (20, 2)] if D == 2 else [(3, 7, 5), (12, 2, 3)]  # both 1.5
    lo = [(30, 30), (9, 4)] if D == 2 else [(14, 14, 3), (9, 4, 12)]  # both -0.25
    for c in hi:
        f[c] = 1.5
    for c in lo:
        f[c] = -0.25
    g.intf.f = f
    flat = lambda c: int(np.ravel_multi_index(c, f.shape, order="F"))
    rc = g._call("ia_vof_sweep", 1, 0.5, None)
    assert rc == 3
    msg = g._lib.ia_last_error(g._ctx).decode()
    assert f"max 1.5 @ {min(flat(c) for c in hi)}," in msg, msg
    assert f"min -0.25 @ {min(flat(c) for c in lo)}" in msg, msg
    assert min(flat(c) for c in hi) == flat(hi[1]) and len({flat(c) // 64 for c in hi}) == 2
    g.close()


# --------------------------------------------------------------------------- plug-ins
def _pair_plugin(case, scheme="WH", lim="Koren"):
    """GPU and oracle with one plug-in; the pressure solves run to round-off
    (Σr² tolerance 1e-20, 200 V-cycles) so the velocities compare tightly."""
    kw = dict(case["kw"])
    g = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], normalScheme=scheme, λ=lim, poisson_tol=1e-20,
                              **kw)
    o = O.OracleSim(case["dims"], case["u_BC"], case["L"], normal=scheme, limiter=lim, mg_tol=1e-20, **kw)
    return g, o


@pytest.mark.parametrize("scheme", ["Column", "WY", "MYC", "PCD", "Youngs", "SLIC"])
@pytest.mark.parametrize("D", [2, 3])
def test_reconstruct_normal_schemes(scheme, D):
    """normalEstimation.jl:35-68 (WY), :75-97 (Column), :161-176 (PCD, SLIC),
    :185-257 (MYC, Youngs) on the GPU."""
    say("reconstructInterface!", scheme, f"{D}D FP64")
    case = bubble2d(32, np.float64) if D == 2 else bubbles3d(16, np.float64)
    case["kw"]["perdir"] = (1,)
    g, o = _pair_plugin(case, scheme=scheme)
    g._call("ia_reconstruct")
    o.lib.oc_reconstruct(o.ctx)
    assert maxdiff(O.inside(g.intf.alpha), O.inside(o.get(O.ALPHA))) <= 1e-12
    assert maxdiff(O.inside(g.intf.nhat, D), O.inside(o.get(O.NHAT), D)) <= 1e-12
    g.close()


@pytest.mark.parametrize("scheme", ["WY", "MYC", "Youngs", "SLIC"])
def test_steps_with_normal_schemes(scheme):
    say("free surface + bubble 32^2 FP64, 3 steps, normal", scheme)
    g, o = _pair_plugin(bubble2d(32, np.float64), scheme=scheme)
    for n in range(3):
        ia.sim_step_(g)
        o.sim_step()
        assert maxdiff(g.intf.f, o.f) <= F_STEP_TOL, n
        assert maxdiff(g.flow.u, o.u) <= U_STAG_TOL * max(1e-12, np.abs(o.u).max()), n
    g.close()


@pytest.mark.parametrize("lim", ["minmod", "upwind", "vanAlbada1", "superbee", "TVDcen", "TVDdown", "Sweby"])
def test_steps_with_limiters(lim):
    """flow.jl:5-15 plug-ins in the SynDRoM face values of the momentum sweep."""
    say("free surface + bubble 32^2 FP64, 3 steps, limiter", lim)
    case = bubble2d(32, np.float64)
    g, o = _pair_plugin(case, lim=lim)
    kw = dict(case["kw"])
    ok = O.OracleSim(case["dims"], case["u_BC"], case["L"], limiter="Koren", mg_tol=1e-20, **kw)
    for n in range(3):
        ia.sim_step_(g)
        o.sim_step()
        ok.sim_step()
        assert maxdiff(g.intf.f, o.f) <= F_STEP_TOL, n
        assert maxdiff(g.flow.u, o.u) <= U_STAG_TOL * max(1e-12, np.abs(o.u).max()), n
    # the f check tells this limiter from the default one by orders of magnitude
    assert maxdiff(ok.f, o.f) >= 100 * F_STEP_TOL
    g.close()


# --------------------------------------------------------------------------- surface tension
def droplet(N=32, D=2, T=np.float64, η=0.05, per=()):
    """An elliptic droplet (dark fluid inside) off the grid's symmetry lines, so
    no interface cell has two equally dominant normal components (a tie would
    let round-off pick the height direction); density and viscosity ratio 0.1,
    surface tension η."""
    c = (0.5 * N + 0.37, 0.5 * N - 0.21, 0.5 * N + 0.13)
    ax = (0.27 * N, 0.22 * N, 0.25 * N)
    sdf = lambda x: 0.24 * N * (np.sqrt(sum(((x[a] - c[a]) / ax[a]) ** 2 for a in range(D))) - 1)
    return dict(dims=(N,) * D, u_BC=(0.0,) * D, L=float(N),
                kw=dict(T=T, λρ=0.1, λμ=0.1, ν=0.01, η=η, InterfaceSDF=sdf, perdir=per))


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("D", [2, 3])
def test_surface_tension_force(T, D):
    """viscSurfTenρu! with η set (surfaceTension.jl:7-124): the force field of the
    GPU's surfTen! (face-averaged f, WY normal, height-function curvature) plus
    the viscous stress, against the oracle's."""
    say("surfTen! + visc!", T.__name__, f"{D}D")
    case = droplet(32 if D == 2 else 16, D, T)
    g, o = pair(case)
    u = _swirl(case, T, amp=0.2)
    g.flow.u = u
    o.set(O.U, u)
    for use_f0 in (0, 1):
        g._call("ia_visc", use_f0)
        o.lib.oc_visc(o.ctx, use_f0)
        Fg, Fo = O.inside(g.flow.force, D), O.inside(o.get(O.FORCE), D)
        scale = max(1e-30, np.abs(Fo).max())
        assert np.abs(Fo).max() > 0
        assert maxdiff(Fg, Fo) <= (1e-12 if T == np.float64 else 1e-5) * scale
    assert g.MPCFL() == o.cfl()
    g.close()


@pytest.mark.parametrize("D", [2, 3])
def test_droplet_steps_with_surface_tension(D):
    """Whole sim_step!s of a droplet with surface tension against the oracle."""
    say("droplet with surface tension, FP64, 3 steps", f"{D}D")
    case = droplet(32 if D == 2 else 16, D, np.float64, per=tuple(range(1, D + 1)))
    kw = dict(case["kw"])
    g = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], poisson_tol=1e-20, **kw)
    o = O.OracleSim(case["dims"], case["u_BC"], case["L"], mg_tol=1e-20, **kw)
    for n in range(3):
        ia.sim_step_(g)
        o.sim_step()
        assert maxdiff(g.intf.f, o.f) <= F_STEP_TOL, n
        assert maxdiff(g.flow.u, o.u) <= 1e-9 * max(1e-12, np.abs(o.u).max()), n
        assert g.flow.Δt[-1] == o.Δt[-1]
    assert np.abs(g.flow.u).max() > 0  # the surface tension drives a flow
    g.close()


# --------------------------------------------------------------------------- diagnostics
@pytest.mark.parametrize("D", [2, 3])
def test_metrics_integrals(D):
    """metrics.jl:15-50 integrated over the interior, GPU against the oracle."""
    say("metrics integrals", f"{D}D")
    case = bubble2d(32, np.float64) if D == 2 else bubbles3d(16, np.float64)
    g, o = pair(case)
    u = _swirl(case, np.float64, amp=0.3)
    g.flow.u = u
    o.set(O.U, u)
    for U, wl in ((None, None), ((0.1, -0.2, 0.3)[:D], (0.0, 8.0, 4.0)[:D])):
        mg = g.metrics(U, wl)
        mo = o.metrics(U, wl)
        got = [mg["kinetic_energy"], *mg["momentum"], *([0.0] * (3 - D)), mg["potential_energy"], mg["enstrophy"],
               mg["volume"]]
        for a, b in zip(got, mo):
            assert abs(a - b) <= 1e-10 * max(1.0, abs(b)), (got, mo)
    assert mg["kinetic_energy"] > 0 and mg["enstrophy"] > 0
    g.close()

