"""Level-set reinitialisation (redistaning.jl): LevelSet(sim) and redistaning!.

CPU: the oracle's restatement against the reference's own testset
(maintests.jl:264-319: the exact ramp, the diagonal plane, the planar front),
the ENO contribution 𝛁ϕᵢ² against closed forms (exact one-sided derivatives of
quadratics, the minmod pick on a cubic, the upwind choice at kinks), the
periodic building block against the inner stencil on a tripled array, and a
disc whose redistanced ϕ must approach the signed distance with |∇ϕ| = 1.

GPU: ia_levelset / ia_redistance through the C ABI against the oracle on the
reference's planar case and on 2D / 3D bubble fields in FP32 and FP64,
periodic and not.  Both sides evaluate every expression in the same order with
contraction off and a correctly rounded sqrt, so the bar is equality to 0 ulp.
"""
import numpy as np
import pytest

from oracle import oracle as O

F = lambda a, T=np.float64: np.asfortranarray(np.asarray(a, dtype=T))


# --------------------------------------------------------------------------- the reference's testset
@pytest.mark.parametrize("perdir", [(), (2,)])
def test_computeL_vanishes_on_an_exact_ramp(perdir):
    """maintests.jl:267-278: ϕ = i - 2.5 - 8 (|∇ϕ| = 1 everywhere, edges included)."""
    N = (18, 18)
    phi = F([[i - 2.5 - 8 for j in range(1, N[1] + 1)] for i in range(1, N[0] + 1)])
    ini = phi.copy(order="F")
    L = np.zeros_like(phi)
    O.computeL(L, phi, ini, perdir=perdir)
    assert np.abs(O.inside(L)).max() < 1e-12


def test_stage_sequence_on_a_diagonal_plane():
    """maintests.jl:280-293: three stages (α = 0, 3/4, 1/3) with ϕini aliasing ϕ."""
    phi = F([[i + j for j in range(1, 7)] for i in range(1, 7)])
    dτ, Lval = 0.01, 1 - np.sqrt(2)
    Lb = np.zeros_like(phi)
    phi0 = phi.copy(order="F")
    for α in (0.0, 0.75, 1 / 3):
        O.redistaning_stage(phi, phi0, phi, Lb, dτ, α)
    assert np.abs(O.inside(phi) - (O.inside(phi0) + dτ * Lval)).max() < 0.05


def planar_front(N=16, T=np.float64):
    """applyVOF! of x -> x[1] - 8 on N x N (perdir=(2,)): the plane lies on a face,
    so the fractions are exactly 1 (x < 8) and 0; ghosts by BCf!."""
    f = np.zeros((N + 2, N + 2), dtype=T, order="F")
    f[1:9, :] = 1
    f[0, :] = f[1, :]
    return f


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_planar_front_becomes_a_distance(T):
    """maintests.jl:304-317 (d = 4, dτ = 0.05, perdir = (2,))."""
    f = planar_front(T=T)
    phi = F(2 * f - 1, T)
    ini, phi0, L = phi.copy(order="F"), np.zeros_like(phi), np.zeros_like(phi)
    O.redistaning(phi, phi0, ini, L, d=4, dτ=0.05, perdir=(2,))
    assert np.isfinite(phi).all()
    col = phi[1:-1, 8]  # ϕ[2:end-1, 9]
    assert (np.diff(col) <= 0).all()
    assert (phi[:, 0] == phi[:, -2]).all() and (phi[:, -1] == phi[:, 1]).all()
    for ix in range(8, 12):
        assert abs(phi[ix - 1, 8] - (-(ix - 1.5 - 8))) <= 0.01, ix


# --------------------------------------------------------------------------- closed forms and invariants
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_eno_contribution_closed_forms(T):
    g = lambda *a: O.ls_grad2(*a, T=T)
    # linear ramp of slope 1: both one-sided derivatives are 1; the upwind one is used
    assert g(0, 1, 2, 3, 4, 1) == 1 and g(0, 1, 2, 3, 4, -1) == 1
    # ϕ = x² about x = 0 and x = 1: second-order one-sided differences are exact (0 and 2)
    assert g(4, 1, 0, 1, 4, 1) == 0 and g(4, 1, 0, 1, 4, -1) == 0
    assert g(1, 0, 1, 4, 9, 1) == 4 and g(1, 0, 1, 4, 9, -1) == 4
    # ϕ = x³ about 0: minmod picks the zero second difference -> both derivatives 1
    assert g(-8, -1, 0, 1, 8, 1) == 1
    # a V (expansion) and its mirror Λ carry no upwind information: 0
    assert g(2, 1, 0, 1, 2, 1) == 0 and g(-2, -1, 0, -1, -2, -1) == 0
    # slope -2 on the left, +1 on the right: for s = -1 information comes from the left,
    # dL = -2 + minmod(3, -1)/2 = -2.5
    assert g(3, 2, 0, 1, 2, -1) == 6.25 and g(3, 2, 0, 1, 2, 1) == 0
    # mirror image of the last case (x -> -x swaps the one-sided roles)
    assert g(2, 1, 0, 2, 3, -1) == 6.25


@pytest.mark.parametrize("axis", [1, 2])
def test_periodic_block_equals_the_inner_stencil_on_a_tripled_array(axis):
    """A field periodic along `axis`: its periodic computeL! must equal the inner
    stencil's result on the middle copy of the field repeated three times."""
    rng = np.random.default_rng(axis)
    n, m = 7, 6
    core = rng.standard_normal((n, m) if axis == 1 else (m, n))
    rep = [3, 1] if axis == 1 else [1, 3]

    def embed(c):
        a = np.zeros((c.shape[0] + 2, c.shape[1] + 2))
        a[1:-1, 1:-1] = c
        return F(a)

    phi, big = embed(core), embed(np.tile(core, rep))
    for a in (phi, big):  # BCf! ghosts: periodic along `axis` for the small one, Neumann otherwise
        O.lib(np.float64).oc_bc_scalar(2, O._ints(a.shape), O._ptr(a), O.perdir_mask((axis,) if a is phi else ()), 0)
    ini_s, ini_b = F(np.sign(phi) * 0.7), F(np.sign(big) * 0.7)
    Ls, Lb = np.zeros_like(phi), np.zeros_like(big)
    O.computeL(Ls, phi, ini_s, perdir=(axis,))
    O.computeL(Lb, big, ini_b)
    mid = (slice(1 + n, 1 + 2 * n), slice(1, -1)) if axis == 1 else (slice(1, -1), slice(1 + n, 1 + 2 * n))
    assert (Lb[mid] == O.inside(Ls)).all()


@pytest.mark.parametrize("axis", [1, 2])
def test_neumann_block_equals_the_inner_stencil_past_a_repeated_ghost(axis):
    """computeL!'s Neumann block takes the point two cells out equal to the ghost
    one cell out (whatever the ghost holds): on the array padded along `axis` by a
    copy of each ghost layer, the inner stencil must give the same L."""
    rng = np.random.default_rng(10 + axis)
    phi = F(rng.standard_normal((8, 7)))  # raw ghosts: not set by BCf!
    ini = F(rng.standard_normal((8, 7)))
    pad = [(1, 1) if a == axis - 1 else (0, 0) for a in range(2)]
    big, ini_b = F(np.pad(phi, pad, mode="edge")), F(np.pad(ini, pad, mode="edge"))
    Ls, Lb = np.zeros_like(phi), np.zeros_like(big)
    O.computeL(Ls, phi, ini)
    O.computeL(Lb, big, ini_b)
    inner = tuple(slice(2, -2) if a == axis - 1 else slice(1, -1) for a in range(2))
    assert (Lb[inner] == O.inside(Ls)).all()


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("nit", [1, 2])
def test_one_rk3_step_on_a_scaled_ramp_is_exact(T, nit):
    """With ϕini = ψ = x - x₀ (|∇ψ| = 1) and ϕ = a ψ, the ENO gradient of the linear
    field is exact, so ∂τϕ = ψ (1 - a) keeps ϕ = a(τ) ψ with the linear ODE
    a' = 1 - a.  SSP-RK3 then gives 1 - a₁ = (1 - a₀) R(-dτ) per step with
    R(z) = 1 + z + z²/2 + z³/6.  Cells within 8 of the x walls (reached by
    BCf!'s edits in 3 stages of a 5-point stencil) are left out."""
    n, a0, dτ = 40, 0.5, 0.3
    psi = np.array([[i - 20.5 for j in range(6)] for i in range(n + 2)])
    phi, ini = F(a0 * psi, T), F(psi, T)
    phi0, L = np.zeros_like(phi), np.zeros_like(phi)
    O.redistaning(phi, phi0, ini, L, d=nit * dτ, dτ=dτ, perdir=(2,))
    z = -dτ
    a = 1 - (1 - a0) * (1 + z + z * z / 2 + z ** 3 / 6) ** nit
    tol = 1e-13 if T == np.float64 else 2e-6
    got = phi[9:n - 7, 1:-1]
    assert np.abs(got - a * psi[9:n - 7, 1:-1]).max() <= tol * np.abs(psi).max()


def disc_levelset(N, R, c, T=np.float64, sub=16):
    """ϕ = 2f - 1 of a disc (f sub-sampled sub² times in cells near the circle)."""
    x = np.arange(N) + 0.5
    X, Y = np.meshgrid(x, x, indexing="ij")
    d = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2)
    f = (d <= R).astype(np.float64)
    s = (np.arange(sub) + 0.5) / sub - 0.5
    for i, j in zip(*np.nonzero(np.abs(d - R) < 1)):
        px, py = np.meshgrid(x[i] + s, x[j] + s, indexing="ij")
        f[i, j] = (((px - c[0]) ** 2 + (py - c[1]) ** 2) <= R * R).mean()
    g = np.zeros((N + 2, N + 2))
    g[1:-1, 1:-1] = f
    a = F(g, T)
    O.lib(T).oc_bc_scalar(2, O._ints(a.shape), O._ptr(a), 0, 0)
    return F(2 * a - 1, T), d


def test_disc_redistances_to_the_signed_distance():
    """Far from a steady state ϕ = 2f - 1 is ±1; after a pseudo-time of 8 the band
    within 4 cells of the circle holds the signed distance R - |x - c| to within
    0.25 cell and |∇ϕ| = 1 to within 5% (central differences, cells 1 to 4 out;
    flipping the upwind choice or the sign of L diverges or stays at ±1)."""
    N, R, c = 64, 17.3, (31.7, 32.4)
    phi, d = disc_levelset(N, R, c)
    ini, phi0, L = phi.copy(order="F"), np.zeros_like(phi), np.zeros_like(phi)
    O.redistaning(phi, phi0, ini, L, d=8, dτ=0.25)
    p = O.inside(phi)
    sd = R - d
    band = np.abs(sd) <= 4
    assert np.abs(p - sd)[band].max() <= 0.25
    gx = (p[2:, 1:-1] - p[:-2, 1:-1]) / 2
    gy = (p[1:-1, 2:] - p[1:-1, :-2]) / 2
    gn = np.sqrt(gx ** 2 + gy ** 2)
    ring = (np.abs(sd[1:-1, 1:-1]) >= 1) & (np.abs(sd[1:-1, 1:-1]) <= 4)
    assert np.abs(gn - 1)[ring].max() <= 0.05


# --------------------------------------------------------------------------- GPU parity
def _oracle_redistance(f, d, dτ, perdir):
    T = f.dtype
    phi = F(T.type(2) * f - T.type(1), T)
    ini, phi0, L = phi.copy(order="F"), np.zeros_like(phi), np.zeros_like(phi)
    O.redistaning(phi, phi0, ini, L, d=d, dτ=dτ, perdir=perdir)
    return phi


def _ulp_diff(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.spacing(np.maximum(np.abs(a), np.abs(b)).astype(a.dtype))
    return float((np.abs(a.astype(np.float64) - b.astype(np.float64)) / scale.astype(np.float64)).max())


@pytest.mark.gpu
def test_levelset_construction_matches_2f_minus_1():
    """maintests.jl:295-302: LevelSet(sim) on (8,8), SDF x -> x[1] - 4, perdir (2,)."""
    from test_gpu_parity import ia, say
    say("LevelSet(sim) on 8^2")
    sim = ia.TwoPhaseSimulation((8, 8), (0.0, 0.0), 8.0, T=np.float64, InterfaceSDF=lambda x: x[0] - 4,
                                perdir=(2,))
    f0 = sim.intf.f
    ls = ia.LevelSet(sim)
    assert (ls.ϕ == 2 * f0 - 1).all()
    assert (sim.intf.f0 == ls.ϕ).all()
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_planar_front_on_the_device(T):
    """maintests.jl:304-317 through the C ABI, and equal to the oracle."""
    from test_gpu_parity import ia, say
    say(f"redistaning! planar front 16^2 {np.dtype(T).name}")
    sim = ia.TwoPhaseSimulation((16, 16), (0.0, 0.0), 16.0, T=T, InterfaceSDF=lambda x: x[0] - 8, perdir=(2,))
    f = sim.intf.f
    assert (f == planar_front(T=T)).all()
    ls = ia.LevelSet(sim)
    ia.redistaning_(ls, d=4, dτ=0.05, perdir=(2,))
    phi = ls.ϕ
    ref = _oracle_redistance(f, 4, 0.05, (2,))
    print(f"planar front {np.dtype(T).name}: max |gpu - oracle| = {np.abs(phi - ref).max():.3e}")
    assert (phi == ref).all()
    assert np.isfinite(phi).all() and (np.diff(phi[1:-1, 8]) <= 0).all()
    assert (phi[:, 0] == phi[:, -2]).all() and (phi[:, -1] == phi[:, 1]).all()
    for ix in range(8, 12):
        assert abs(phi[ix - 1, 8] - (-(ix - 1.5 - 8))) <= 0.01
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("D,T,perdir,d,dτ,N", [
    (2, np.float32, (), 5, 0.5, 48),
    (2, np.float64, (1,), 3, 0.25, 48),
    (3, np.float32, (1, 3), 4, 0.5, 32),
    (3, np.float64, (), 2, 0.5, 32),
    # 96^3 takes the LDS-tiled stage (4-cell columns, a half-filled second x tile)
    (3, np.float32, (1, 2, 3), 2, 0.5, 96),
    (3, np.float64, (2,), 2, 0.5, 96),
])
def test_redistance_matches_the_oracle(D, T, perdir, d, dτ, N):
    from test_gpu_parity import bubble2d, bubbles3d, ia, say
    say(f"redistaning! {D}D {N} {np.dtype(T).name} perdir={perdir} d={d} dτ={dτ}")
    case = bubble2d(N, T) if D == 2 else bubbles3d(N, T)
    sim = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **case["kw"])
    f = sim.intf.f
    ls = ia.LevelSet(sim)
    ia.redistaning_(ls, d=d, dτ=dτ, perdir=perdir)
    phi = ls.ϕ
    ref = _oracle_redistance(f, d, dτ, perdir)
    print(f"{D}D {np.dtype(T).name}: max |gpu - oracle| = {np.abs(phi - ref).max():.3e} "
          f"({_ulp_diff(phi, ref):.1f} ulp)")
    assert np.isfinite(phi).all()
    assert (phi == ref).all()
    # a level set that starts at ±1 spreads: the far field reaches beyond 1 after pseudo-time d > 1
    assert np.abs(O.inside(phi)).max() > 1.5
    sim.close()


@pytest.mark.gpu
def test_redistance_requires_a_fresh_levelset():
    from test_gpu_parity import bubble2d, ia, say
    say("ia_redistance argument checks")
    case = bubble2d(32, np.float32)
    sim = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **case["kw"])
    with pytest.raises(ia.IAError):
        sim._call("ia_redistance", 1.0, 0.5, 0)
    ls = ia.LevelSet(sim)
    with pytest.raises(ia.IAError):
        ia.redistaning_(ls, d=1.0, dτ=0.0)
    with pytest.raises(ia.IAError):
        ia.redistaning_(ls, perdir=(3,))
    ia.redistaning_(ls, d=1.0, dτ=0.5)
    ia.sim_step_(sim)
    with pytest.raises(ia.IAError):
        ia.redistaning_(ls, d=1.0, dτ=0.5)
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_scaled_ramp_rk3_step_on_the_device(T):
    """The closed form above through the C ABI: f = (ψ + 1)/2 makes LevelSet's
    ϕini = 2f - 1 = ψ exactly; ϕ is then set to a₀ψ through f⁰."""
    from test_gpu_parity import ia, say
    say(f"redistaning! scaled ramp 40x8 {np.dtype(T).name}")
    n, a0, dτ = 40, 0.5, 0.3
    psi = np.array([[i - 20.5 for j in range(10)] for i in range(n + 2)])
    sim = ia.TwoPhaseSimulation((n, 8), (0.0, 0.0), float(n), T=T, perdir=(2,))
    sim.intf.f = F((psi + 1) / 2, T)
    ls = ia.LevelSet(sim)
    assert (ls.ϕ == F(psi, T)).all()
    ls.ϕ = F(a0 * psi, T)
    ia.redistaning_(ls, d=dτ, dτ=dτ, perdir=(2,))
    phi = ls.ϕ
    ref, ini = F(a0 * psi, T), F(psi, T)
    O.redistaning(ref, np.zeros_like(ref), ini, np.zeros_like(ref), d=dτ, dτ=dτ, perdir=(2,))
    assert (phi == ref).all()
    z = -dτ
    a = 1 - (1 - a0) * (1 + z + z * z / 2 + z ** 3 / 6)
    tol = 1e-13 if T == np.float64 else 2e-6
    assert np.abs(phi[9:n - 7, 1:-1] - a * psi[9:n - 7, 1:-1]).max() <= tol * np.abs(psi).max()
    sim.close()
