"""Pins of the oracle's step functions that no reference test covers, each by
what the mathematics fixes (CPU only, FP64 oracle):

- visc! (flow.jl:113-153): uniform u gives no force, whatever the phases;
  single phase (λμ = λρ = 1, so the bounded viscosity is μ), a quadratic u
  gives exactly μ(∇²u_i + ∂_i ∇·u) at every interior face, wall slices
  included; on periodic axes, Fourier modes give their exact discrete
  eigenvalue 2cos(k) - 2 (the pure-shear mode once, the compression mode
  twice, through the transposed gradient).
- MPCFL (flow.jl:262-281): hand-computed cases in which each of its five
  limits is the active one.
- ϕq (flow.jl:37-57): hand-computed face values in the donor regime
  (|Ψ|δt > mOld: Ψ·c), in the blend regime, and the l2 → 0 limit (Ψ times
  the limiter's face value).
- updateU! with gravity (flow.jl:244-252) through a whole MPFMomStep!: a flat
  interface at rest stays at rest, and the pressure is hydrostatic,
  p_j - p_{j-1} = ρ_face g_y on every vertical face.
- updateU! in a periodic box (flow.jl:244-252, :74-77): free fall, u = gΔt,
  and the interface translated by the predictor's gΔt²/2.
- WaterLily BC!(a, A, exitBC, perdir) (as flow.jl:69, :79, :82, :197, :207
  call it): with exitBC the exit face of u_x (the upper x face)
  keeps its value, and the tangential Neumann images at that face copy it;
  everything else is as without.
- reportFillError (advection.jl:145-189): a strongly divergent velocity at the
  interface (|∇·u⁰| + |∇·u| > 10 where f leaves [0, 1]) returns the explosion
  status, exactly 10 only the fill warning.
- MPFMomStep! trajectories (flow.jl:60-109) on a periodic grid: three steps
  commute with a translation by whole cells, and with the reflection of a
  periodic axis.
"""
import numpy as np
import pytest

from oracle import oracle as O

T = np.float64


def _sim(dims, **kw):
    return O.OracleSim(dims, (0.0,) * len(dims), float(dims[0]), T=T, **kw)


def _face_coords(Ng, i):
    """Coordinates of the staggered component i (0-based) at every index,
    x_d = q_d - 0.5 δ(i, d)."""
    D = len(Ng)
    ax = [np.arange(n, dtype=np.float64) - (0.5 if d == i else 0.0) for d, n in enumerate(Ng)]
    return np.meshgrid(*ax, indexing="ij")


# ----------------------------------------------------------------------------- visc!
@pytest.mark.parametrize("dims", [(8, 8, 8), (16, 8)])
def test_visc_uniform_velocity_gives_no_force(dims):
    rng = np.random.default_rng(0)
    o = _sim(dims, ν=0.3, λμ=1e-2, λρ=1e-3)
    D = len(dims)
    Ng = o.Ng
    f = np.asfortranarray(rng.uniform(0, 1, Ng))
    o.set(O.F, f)
    u = np.zeros(Ng + (D,), order="F")
    for i in range(D):
        u[..., i] = 0.37 * (i + 1) - 0.5
    o.set(O.U, u)
    o.lib.oc_visc(o.ctx, 0)
    r = o.get(O.FORCE)
    assert np.abs(r).max() == 0.0


@pytest.mark.parametrize("dims", [(8, 8, 8), (16, 8)])
def test_visc_single_phase_quadratic_is_exact(dims):
    """r_i = μ (tr H_i + Σ_j (H_j)_{ij}) for u_i with Hessian H_i (constant)."""
    rng = np.random.default_rng(1)
    D = len(dims)
    μ = 0.05
    o = _sim(dims, ν=μ, λμ=1.0, λρ=1.0)
    Ng = o.Ng
    o.set(O.F, np.asfortranarray(rng.uniform(0, 1, Ng)))  # any f: single phase
    H = [None] * D
    u = np.zeros(Ng + (D,), order="F")
    for i in range(D):
        A = rng.uniform(-1, 1, (D, D))
        H[i] = (A + A.T) / 10  # Hessian of u_i
        b = rng.uniform(-1, 1, D)
        X = _face_coords(Ng, i)
        val = np.full(Ng, 0.3)
        for a in range(D):
            val += b[a] * X[a]
            for c in range(D):
                val += 0.5 * H[i][a, c] * X[a] * X[c]
        u[..., i] = val
    o.set(O.U, u)
    o.lib.oc_visc(o.ctx, 0)
    r = o.get(O.FORCE)
    inner = tuple(slice(1, n + 1) for n in dims)
    for i in range(D):
        expect = μ * (np.trace(H[i]) + sum(H[j][i, j] for j in range(D)))
        got = r[inner + (i,)]
        assert np.abs(got - expect).max() <= 1e-11 * max(1.0, abs(expect)), (i, expect, got.min(), got.max())


@pytest.mark.parametrize("mode", ["shear", "compression"])
def test_visc_periodic_fourier_modes(mode):
    """Periodic in x and y (2D, N = 16): u_x = cos(k y) (shear) gives
    μ(2cos k - 2) cos(k y); u_x = cos(k x) (compression, ∇·u ≠ 0) gives twice
    that, through the transposed gradient."""
    N = 16
    μ = 0.2
    o = _sim((N, N), ν=μ, λμ=1.0, λρ=1.0, perdir=(1, 2))
    Ng = o.Ng
    o.set(O.F, np.ones(Ng, order="F"))
    k = 2 * np.pi / N
    X, Y = _face_coords(Ng, 0)
    u = np.zeros(Ng + (2,), order="F")
    u[..., 0] = np.cos(k * Y) if mode == "shear" else np.cos(k * X)
    o.set(O.U, u)
    o.lib.oc_visc(o.ctx, 0)
    r = o.get(O.FORCE)
    lam = 2 * np.cos(k) - 2
    inner = (slice(1, N + 1), slice(1, N + 1))
    expect = μ * lam * (np.cos(k * Y) if mode == "shear" else 2 * np.cos(k * X))
    assert np.abs(r[inner + (0,)] - expect[inner]).max() <= 1e-12
    assert np.abs(r[inner + (1,)]).max() <= 1e-12


# ----------------------------------------------------------------------------- MPCFL
def _cfl_case(u_scale=1.0, **kw):
    """2D 8x8 walls; cell (2,2) (0-based, ghosts included) has u_x = 0.5 on its
    upper x face and u_y = -0.25 on its upper y face, every other face at rest.
    By hand: maxTotalFlux is 0.75 at cell (2,2) (0.5 + 0.25), 0.5 at (3,2) and
    0.25 at (2,3); flux_out is 0.5 at (2,2) (through the x face), 0 at (3,2)
    (inflow) and 0.25 at (2,3) (the downward flow leaves it)."""
    o = _sim((8, 8), **kw)
    u = np.zeros(o.Ng + (2,), order="F")
    u[3, 2, 0] = 0.5 * u_scale
    u[2, 3, 1] = -0.25 * u_scale
    o.set(O.U, u)
    return o.cfl()


def test_mpcfl_each_limit_active():
    # Δt_cVOF = 1/(2·0.75) is the smallest: 0.8/1.5
    assert _cfl_case() == pytest.approx(0.8 / 1.5, rel=1e-14)
    # Δt_Adv = 1/(0.5 + 5ν) = 0.5 with ν = 0.3, λμ = λρ (Δt_Visc = 3/4.2 = 0.714)
    assert _cfl_case(ν=0.3, λμ=1.0, λρ=1.0) == pytest.approx(0.8 * 0.5, rel=1e-14)
    # Δt_Grav = 1/(2|g|) = 0.25 with g = (0, -2)
    assert _cfl_case(g=(0.0, -2.0)) == pytest.approx(0.8 * 0.25, rel=1e-14)
    # |g| = √(3² + 4²) = 5: 0.1
    assert _cfl_case(g=(3.0, -4.0)) == pytest.approx(0.8 * 0.1, rel=1e-14)
    # Δt_Visc = 3/(14 ν max(1, λμ/λρ)) = 3/7 with ν = 0.05, λμ/λρ = 10 (Δt_Adv = 1/0.75)
    assert _cfl_case(ν=0.05, λμ=1e-2, λρ=1e-3) == pytest.approx(0.8 * 3 / 7, rel=1e-14)
    # ... and max(1, λμ/λρ) = 1 when λμ < λρ: ν = 0.3, u scaled by 1/10: Δt_Visc = 3/4.2 = 0.714
    # (it would be 0.0714 with the ratio inverted), Δt_Adv = 1/(0.05 + 1.5) = 0.645,
    # Δt_cVOF = 1/0.15 = 6.667: Adv is active
    assert _cfl_case(u_scale=0.1, ν=0.3, λμ=1e-3, λρ=1e-2) == pytest.approx(0.8 / 1.55, rel=1e-14)
    # Δt_SurfT = √((1 + λρ)/(8π η)) with η = 1, λρ = 1e-3
    assert _cfl_case(η=1.0, λρ=1e-3) == pytest.approx(0.8 * np.sqrt(1.001 / (8 * np.pi)), rel=1e-14)
    # nothing moves, no physics: every limit is Δt_max = 1 or infinite
    o = _sim((8, 8))
    assert o.cfl() == pytest.approx(0.8, rel=1e-14)


# ----------------------------------------------------------------------------- ϕq
def _phiq(Psi, fd, dt, lr, u, c, d, lim="Koren"):
    return O.lib(T).oc_phiq_value(Psi, fd, dt, lr, O.LIMITER[lim], u, c, d)


def test_phiq_regimes():
    # Koren(0, 1, 3) = median((7 + 3)/6, 1, median(2, 1, 3)) = 5/3, so va = 2 - 5/3 = 1/3
    # donor regime: |Ψ|δt = 2 > mOld = lin(0, 0.5) = 0.5 -> Ψ·c = 2
    assert _phiq(2.0, 0.0, 1.0, 0.5, 0.0, 1.0, 3.0) == 2.0
    # the same with Ψ < 0: -2·c
    assert _phiq(-2.0, 0.0, 1.0, 0.5, 0.0, 1.0, 3.0) == -2.0
    # blend: Ψ = 0.5, δt = 1, fd = 1 -> mOld = 1, l2 = 0.5, vb = (1/3 + 5/3)/2 = 1,
    # Ψ(vb + vd)/2 = 0.5 (1 + 5/3)/2 = 2/3
    assert _phiq(0.5, 1.0, 1.0, 0.5, 0.0, 1.0, 3.0) == pytest.approx(2 / 3, rel=1e-15)
    # blend with a light donor: fd = 0.2, λρ = 0.5 -> mOld = 0.6; Ψ = 0.3, δt = 1 -> l2 = 1/2 again
    assert _phiq(0.3, 0.2, 1.0, 0.5, 0.0, 1.0, 3.0) == pytest.approx(0.3 * (1 + 5 / 3) / 2, rel=1e-14)
    # l2 -> 0: Ψ times the limiter's face value 5/3
    v = _phiq(1e-9, 1.0, 1.0, 0.5, 0.0, 1.0, 3.0)
    assert v / 1e-9 == pytest.approx(5 / 3, rel=1e-8)
    # minmod(0, 1, 3) = median(3/2, 1, 2) = 3/2, l2 = 1/4: vb = (1/4)(1/2) + (3/4)(3/2) = 5/4,
    # Ψ(vb + vd)/2 = 0.25 (5/4 + 3/2)/2 = 11/32
    assert _phiq(0.25, 1.0, 1.0, 0.5, 0.0, 1.0, 3.0, "minmod") == pytest.approx(11 / 32, rel=1e-15)
    # mOut == mOld is not the donor branch (strict >): l2 = 1 -> Ψ(va + vd)/2 = Ψ·c
    assert _phiq(1.0, 1.0, 1.0, 0.5, 0.0, 1.0, 3.0) == pytest.approx(1.0, rel=1e-15)


# ----------------------------------------------------------------------------- gravity
@pytest.mark.parametrize("yint", [8.0, 8.3])
def test_flat_interface_under_gravity_is_hydrostatic(yint):
    N = 16
    g = -0.05
    o = _sim((N, N), λρ=1e-3, λμ=1e-2, g=(0.0, g), InterfaceSDF=lambda x: x[1] - yint, poisson="refined",
             poisson_rtol=1e-12)
    f0 = o.f
    for _ in range(2):
        o.sim_step()
    assert np.abs(o.f - f0).max() <= 1e-12  # nothing moves (u is 0 to the solver's tolerance)
    assert np.abs(o.u).max() <= 1e-10 * abs(g) * float(o.Δt[-2])
    p = o.p
    L = o.get(O.MU0)
    # every vertical face between interior cells: p_j - p_{j-1} = g / L_y = ρ_face g
    dp = p[1:N + 1, 2:N + 1] - p[1:N + 1, 1:N]
    expect = g / L[1:N + 1, 2:N + 1, 1]
    assert np.abs(dp - expect).max() <= 1e-9 * abs(g)
    # and ρ_face is the face-averaged density of updateL!: 1 in the liquid, λρ in the gas
    assert expect.min() == pytest.approx(g, rel=1e-12)
    assert expect.max() == pytest.approx(g * 1e-3, rel=1e-12)


def test_periodic_free_fall():
    """Periodic in x and y, two phases at rest with a flat interface, gravity
    -y: nothing resists the fall, so one step gives u = g Δt everywhere
    (exactly: uniform u is divergence-free and CMOM keeps it uniform), and the
    corrector advects f with the predictor's u = g Δt/2, which translates the
    planar interface by g Δt²/2 (exact for a planar interface in its cell)."""
    N = 16
    g = -0.05
    dt = 0.5
    o = _sim((N, N), λρ=1e-3, λμ=1e-2, g=(0.0, g), InterfaceSDF=lambda x: x[1] - 8.3, perdir=(1, 2), Δt=dt)
    assert float(o.Δt[-1]) == dt
    f0 = o.f
    o.sim_step()
    u = o.u
    assert np.abs(u[..., 1] - g * dt).max() <= 1e-14
    assert np.abs(u[..., 0]).max() <= 1e-14
    f1 = o.f
    rows = [j for j in range(1, N + 1) if 0 < f0[5, j] < 1]
    assert len(rows) == 1
    j = rows[0]
    assert np.abs(f1[1:N + 1, j] - (f0[1:N + 1, j] + g * dt * dt / 2)).max() <= 1e-14
    # the periodic wrap is a second, cell-aligned interface (liquid row 1 under
    # gas row N): the top row receives the fluid leaving row 1 downwards
    assert np.abs(f1[1:N + 1, N] - (-g * dt * dt / 2)).max() <= 1e-14
    others = [q for q in range(1, N) if q != j]
    assert np.array_equal(f1[1:N + 1, others], f0[1:N + 1, others])


# ----------------------------------------------------------------------------- reportFillError
def _squeeze(U, dt):
    """A column of f = 0.4 between full columns, squeezed by u_x = ±U (u⁰ = u)
    from both sides: δl = U δt of fluid enters through each face."""
    N = 16
    o = _sim((N, N), perdir=(1, 2))
    f = np.ones(o.Ng, order="F")
    f[8, :] = 0.4
    o.set(O.F, f)
    u = np.zeros(o.Ng + (2,), order="F")
    u[8, :, 0] = U    # lower x face of column 8: inflow
    u[9, :, 0] = -U   # upper x face: inflow
    o.set(O.U, u)
    o.set(O.U0, u)
    return o.advect(dt, 1)


def test_divergent_flow_at_the_interface_explodes():
    """δl = 0.4 from each side takes f to 1.2.  |∇·u⁰| + |∇·u| is 4U there:
    U = 3 gives 12 > 10, the explosion (advection.jl:160, status 2); U = 2.5
    gives exactly 10, only the printed fill error (status 3); a tenth of the
    inflow stays in [0, 1]."""
    assert _squeeze(3.0, 0.4 / 3.0) == 2
    assert _squeeze(2.5, 0.4 / 2.5) == 3
    assert _squeeze(3.0, 0.04 / 3.0) == 0


# ----------------------------------------------------------------------------- BC! with exitBC
@pytest.mark.parametrize("D", [2, 3])
def test_bc_exit_face_is_left_alone(D):
    rng = np.random.default_rng(12)
    Ng = (10, 8, 6)[:D]
    a = np.asfortranarray(rng.standard_normal(Ng + (D,)))
    A = np.array([0.7, -0.2, 0.3])
    L = O.lib(np.float64)
    plain, kept = a.copy(order="F"), a.copy(order="F")
    L.oc_bc_vel(D, O._ints(Ng), O._ptr(plain), O._ptr(A), 0, 0)
    L.oc_bc_vel(D, O._ints(Ng), O._ptr(kept), O._ptr(A), 0, 1)
    nx = Ng[0]
    # without exitBC: every normal face on slices 1, 2 and N is A
    assert np.all(plain[0, ..., 0] == A[0]) and np.all(plain[1, ..., 0] == A[0]) and np.all(plain[-1, ..., 0] == A[0])
    # with it: the exit slice of u_x keeps its interior-row values, and its
    # tangential ghosts copy their interior neighbours
    ex = a[-1, ..., 0].copy()
    inner = tuple(slice(1, n - 1) for n in Ng[1:])
    assert np.array_equal(kept[(-1,) + inner + (0,)], ex[inner])
    assert np.array_equal(kept[-1, 0, ..., 0], kept[-1, 1, ..., 0])
    assert np.array_equal(kept[-1, -1, ..., 0], kept[-1, -2, ..., 0])
    # and everything else is as without exitBC
    mask = np.ones(kept.shape, bool)
    mask[-1, ..., 0] = False
    assert np.array_equal(kept[mask], plain[mask])


# ----------------------------------------------------------------------------- translation
def _drop_case(shift, poisson):
    """A droplet (radius 5.2 cells, density and viscosity ratios 0.1) in a
    Taylor-Green field plus a uniform drift, fully periodic 32², with the
    droplet and the field moved by `shift` cells."""
    N = 32
    k = 2 * np.pi / N
    cx, cy = 10.3 + shift[0], 9.6 + shift[1]

    def sdf(x):
        return np.sqrt((x[0] - cx) ** 2 + (x[1] - cy) ** 2) - 5.2

    def u0(i, x):
        a, b = k * (x[0] - shift[0]), k * (x[1] - shift[1])
        return -np.sin(a) * np.cos(b) + 0.3 if i == 1 else np.cos(a) * np.sin(b) + 0.2

    kw = dict(poisson_rtol=1e-12) if poisson == "refined" else dict(mg_tol=1e-12, mg_itmx=400)
    return O.OracleSim((N, N), (0.0, 0.0), float(N), T=T, λμ=0.1, λρ=0.1, ν=0.02, InterfaceSDF=sdf, u0=u0,
                       perdir=(1, 2), poisson=poisson, **kw)


@pytest.mark.parametrize("shift,poisson,tol", [((16, 16), "reference", 1e-12), ((5, 11), "refined", 1e-11),
                                               ((8, 16), "refined", 1e-11)])
def test_steps_are_translation_equivariant(shift, poisson, tol):
    """A whole MPFMomStep! (flow.jl:60-109) on a periodic grid commutes with
    a translation by whole cells: three steps of the moved case equal the
    moved result of the original's, f and u to `tol`, p about its mean, and
    the Δt sequence.  WaterLily's solver (a few V-cycles here) depends on
    how the coarse levels aggregate cells: a shift by 16 keeps every level of
    32² (down to 2², each coarse cell 16 fine ones wide) aligned, so only the
    order of the solver's sums changes (measured 4e-15); by (8, 16) or
    (5, 11) its results move by 2e-9 and 3e-8.  The refined solve, to 1e-12,
    removes that: under any shift the steps agree to round-off (measured
    4e-15 in f, 3e-14 in u).  A stencil offset, a ghost image or a wrap that
    depends on the position in the grid breaks it."""
    a, b = _drop_case((0, 0), poisson), _drop_case(shift, poisson)
    for _ in range(3):
        a.sim_step()
        b.sim_step()
    assert np.allclose(a.Δt, b.Δt, rtol=1e-12, atol=0)

    def cells(x):
        return x[1:-1, 1:-1]

    def moved(x):
        return np.roll(cells(x), shift, axis=(0, 1))

    fa, fb = moved(a.f), cells(b.f)
    assert 0.05 < fa.sum() / fa.size < 0.5
    assert np.abs(fa - fb).max() <= tol
    ua, ub = moved(a.u), cells(b.u)
    assert np.abs(ua - ub).max() <= tol * np.abs(ua).max()
    pa, pb = moved(a.p), cells(b.p)
    pa, pb = pa - pa.mean(), pb - pb.mean()
    assert np.abs(pa - pb).max() <= tol * np.abs(pa).max() * 10


# ----------------------------------------------------------------------------- mirror
def _mirror_case(mirror, limiter):
    """A bubble below a wavy free surface, x periodic (32 cells), slip walls
    in y, gravity -y, density ratio 1000, viscosity ratio 100; with `mirror`
    every field reflected x -> N - x (u_x negated)."""
    N = 32
    k = 2 * np.pi / N
    X = (lambda x: N - x) if mirror else (lambda x: x)
    sx = -1.0 if mirror else 1.0

    def sdf(x):
        xx = X(x[0])
        bub = 5.2 - np.sqrt((xx - 11.3) ** 2 + (x[1] - 12.6) ** 2)
        return np.maximum(x[1] - 22.4 - 1.5 * np.sin(k * xx + 0.4), bub)

    def u0(i, x):
        xx = X(x[0])
        if i == 1:
            return sx * (0.3 * np.sin(k * xx) * np.cos(np.pi * x[1] / N) + 0.05 * np.cos(2 * k * xx + 0.3))
        return -0.2 * np.cos(k * xx) * np.sin(np.pi * x[1] / N) + 0.04 * np.sin(k * xx + 0.7)

    return O.OracleSim((N, N), (0.0, 0.0), float(N), T=T, λμ=1e-2, λρ=1e-3, ν=0.02, g=(0.0, -0.01),
                       InterfaceSDF=sdf, u0=u0, perdir=(1,), limiter=limiter, mg_tol=1e-12, mg_itmx=400)


@pytest.mark.parametrize("limiter", ["Koren", "minmod", "superbee"])
def test_steps_are_mirror_symmetric_along_a_periodic_axis(limiter):
    """Three whole steps commute with the reflection x -> N - x of a periodic
    axis (f and u_y mirrored, u_x mirrored and negated: x-face q maps to
    face N + 2 - q, face 1 to itself): every upwind choice (VOF donor, the
    momentum face triples, ϕq's donor cell, the limiter's arguments) has its
    mirror image in the other sign.  Measured 1e-14.  Along a wall axis the
    reference itself is not symmetric: its lower-boundary momentum flux
    (ϕuL) sits at the ghost cell's centre, so the normal component's flux
    at the first interior cell takes the ghost face as its far-upwind value,
    while ϕuR extrapolates at the last interior cell (flow.jl:233-236); with
    the upwind limiter, which ignores that value, walls are symmetric too."""
    a, b = _mirror_case(False, limiter), _mirror_case(True, limiter)
    N = 32
    for _ in range(3):
        a.sim_step()
        b.sim_step()
    assert np.allclose(a.Δt, b.Δt, rtol=1e-13, atol=0)
    fa, fb = a.f[1:-1, 1:-1], b.f[1:-1, 1:-1][::-1, :]
    assert 0.3 < fa.mean() < 0.8
    assert np.abs(fa - fb).max() <= 1e-11
    q = np.arange(1, N + 1)
    qm = (N + 2 - q - 1) % N + 1  # x-face q -> N + 2 - q, face N + 1 is face 1
    uxa, uxb = a.u[q, 1:-1, 0], -b.u[qm, 1:-1, 0]
    scale = np.abs(a.u).max()
    assert np.abs(uxa - uxb).max() <= 1e-11 * scale
    uya, uyb = a.u[1:-1, 1:-1, 1], b.u[1:-1, 1:-1, 1][::-1, :]
    assert np.abs(uya - uyb).max() <= 1e-11 * scale
    pa, pb = a.p[1:-1, 1:-1], b.p[1:-1, 1:-1][::-1, :]
    assert np.abs((pa - pa.mean()) - (pb - pb.mean())).max() <= 1e-10 * np.abs(pa).max()
