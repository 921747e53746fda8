"""The oracle against the reference's own known-answer tests.

Every expected value below is transcribed from test/maintests.jl (cited per
test).  Julia matrix literals list rows, and the row number is the first (x)
index, so `[a b; c d]` becomes `np.array([[a, b], [c, d]])` indexed [x, y].
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

TS = [np.float32, np.float64]


def R(T):
    return C.c_float if T == np.float32 else C.c_double


def approx(a, b, T):
    tol = 1e-5 if T == np.float32 else 1e-12
    return abs(float(a) - float(b)) <= tol * max(1.0, abs(float(b)))


@pytest.mark.parametrize("T", TS)
def test_plic_kats(T):
    """maintests.jl:25-51."""
    L = O.lib(T)
    t = lambda x: T(x)
    gi2, gi3 = L.oc_getIntercept2, L.oc_getIntercept3
    gv2, gv3 = L.oc_getVolumeFraction2, L.oc_getVolumeFraction3
    assert approx(gi3(t(2 / 3), t(4 / 3), t(0), t(5 / 12)), 8 / 9, T)
    assert approx(gi2(t(2 / 3), t(4 / 3), t(5 / 12)), 8 / 9, T)
    assert approx(gi3(t(2 / 3), t(0), t(4 / 3), t(5 / 12)), 8 / 9, T)
    assert approx(gi3(t(2 / 3), -t(4 / 3), t(0), t(5 / 12)), -4 / 9, T)
    assert approx(gi2(t(2 / 3), -t(4 / 3), t(5 / 12)), -4 / 9, T)
    assert approx(gi3(t(3), -t(4), t(0), t(5 / 6)), 1, T)
    assert approx(gi2(t(3), -t(4), t(5 / 6)), 1, T)
    assert approx(gi3(t(1 / 2), t(1 / 3), t(1), t(7 / 12)), 1, T)
    assert approx(gi3(t(1), t(1), -t(1), t(1 - 1 / 48)), 3 / 2, T)
    assert approx(gi2(-t(7 / 6), t(4 / 9), t(7 / 23)), gi3(-t(7 / 6), t(4 / 9), t(0), t(7 / 23)), T)

    assert approx(gv3(t(2 / 3), t(4 / 3), t(0), t(8 / 9)), 5 / 12, T)
    assert approx(gv2(t(2 / 3), t(4 / 3), t(8 / 9)), 5 / 12, T)
    assert approx(gv3(t(2 / 3), t(0), t(4 / 3), t(8 / 9)), 5 / 12, T)
    assert approx(gv3(t(2 / 3), -t(4 / 3), t(0), -t(4 / 9)), 5 / 12, T)
    assert approx(gv2(t(2 / 3), -t(4 / 3), -t(4 / 9)), 5 / 12, T)
    assert approx(gv3(t(3), -t(4), t(0), t(1)), 5 / 6, T)
    assert approx(gv2(t(3), -t(4), t(1)), 5 / 6, T)
    assert approx(gv3(t(1 / 2), t(1 / 3), t(1), t(1)), 7 / 12, T)
    assert approx(gv3(t(1), t(1), -t(1), t(3 / 2)), 1 - 1 / 48, T)
    assert approx(gv2(-t(7 / 6), t(4 / 9), t(7 / 23)), gv3(-t(7 / 6), t(4 / 9), t(0), t(7 / 23)), T)


def _blk(f):
    return np.asfortranarray(f, dtype=np.float64)


def test_util_and_height_kats():
    """maintests.jl:53-63 (get3CellHeight, getρ) and :5-23 (index helpers via the normals)."""
    f = np.zeros((3, 3))
    f[1, 1] = 0.32
    f[1, 2] = 0.64
    # get3CellHeight(f,Ic,2) = f[I]+f[I-δ2]+f[I+δ2]
    assert abs((f[1, 1] + f[1, 0]) + f[1, 2] - 0.96) < 1e-12
    # getρ(Ic,f,0.7) = λ+(1-λ)f ; getρ(2,Ic,f,0.7) uses the face average (f[I]+f[I-δ2])/2
    lr = 0.7
    assert abs(lr + (1 - lr) * f[1, 1] - 0.796) < 1e-12
    assert abs(lr + (1 - lr) * (f[1, 1] + f[1, 0]) / 2 - 0.748) < 1e-12


def test_getmu_kats():
    """maintests.jl:65-70."""
    L = O.lib(np.float64)
    ff = np.zeros((3, 3, 2), order="F")
    ff[2, 1, 0] = 0.1
    ff[2, 2, 0] = 0.2
    ff[1, 2, 1] = 0.3
    ff[2, 2, 1] = 0.4
    Ng = O._ints([3, 3])
    Iur = O._ints([3, 3])
    mu = lambda i, j, lm, m, lr: L.oc_getmu(2, Ng, O._ptr(ff), i, j, Iur, lm, m, lr)
    assert abs(mu(1, 1, 0.1, 0.2, 1.0) - 0.02) < 1e-14
    assert mu(1, 2, 0.1, 0.2, 0.2) == mu(2, 1, 0.1, 0.2, 0.2)
    assert abs(mu(1, 2, 0.1, 0.2, 0.2) - 0.028) < 1e-14


def test_f2face_kat():
    """maintests.jl:72-77."""
    L = O.lib(np.float64)
    fc = np.zeros((4, 4), order="F")
    fc[1:3, 1:3] = [[0.2, 0.6], [0.4, 0.8]]
    ff = np.zeros((4, 4, 2), order="F")
    L.oc_f2face(2, O._ints([4, 4]), O._ptr(fc), O._ptr(ff), 0)
    assert abs(ff[2, 2, 0] - 0.7) < 1e-14 and ff[2, 2, 0] == (fc[1, 2] + fc[2, 2]) / 2
    assert abs(ff[2, 2, 1] - 0.6) < 1e-14 and ff[2, 2, 1] == (fc[2, 1] + fc[2, 2]) / 2


def test_bc_kats():
    """maintests.jl:79-129 (BCf!, BCf!(d), BCv!, BCv1D!, BCVOF!)."""
    L = O.lib(np.float64)
    rng = np.random.default_rng(0)
    Ng = O._ints([6, 6])
    g = np.asfortranarray(rng.random((6, 6)))
    gN = g.copy(order="F")
    L.oc_bc_scalar(2, Ng, O._ptr(gN), 0, 0)
    assert (gN[0, :] == gN[1, :]).all() and (gN[-1, :] == gN[-2, :]).all()
    assert (gN[:, 0] == gN[:, 1]).all() and (gN[:, -1] == gN[:, -2]).all()
    gP = g.copy(order="F")
    L.oc_bc_scalar(2, Ng, O._ptr(gP), 0b11, 0)
    assert (gP[0, :] == gP[-2, :]).all() and (gP[-1, :] == gP[1, :]).all()
    assert (gP[:, 0] == gP[:, -2]).all() and (gP[:, -1] == gP[:, 1]).all()
    g1 = g.copy(order="F")
    L.oc_bc_scalar(2, Ng, O._ptr(g1), 0, 1)
    assert (g1[0, 1:-1] == g[2, 1:-1]).all()
    assert (g1[-1, 1:-1] == g[-1, 1:-1]).all()
    assert (g1[:, 0] == g1[:, 1]).all() and (g1[:, -1] == g1[:, -2]).all()
    g1p = g.copy(order="F")
    L.oc_bc_scalar(2, Ng, O._ptr(g1p), 0b01, 1)
    assert (g1p[0, :] == g1p[-2, :]).all() and (g1p[-1, :] == g1p[1, :]).all()

    u = np.asfortranarray(rng.random((6, 6, 2)))
    uN = u.copy(order="F")
    L.oc_bc_facevec(2, Ng, O._ptr(uN), 0)
    assert (uN[0, 1:-1, 0] == u[2, 1:-1, 0]).all()
    assert (uN[-1, 1:-1, 0] == u[-1, 1:-1, 0]).all()
    assert (uN[:, 0, 0] == uN[:, 1, 0]).all() and (uN[:, -1, 0] == uN[:, -2, 0]).all()
    assert (uN[1:-1, 0, 1] == u[1:-1, 2, 1]).all()
    assert (uN[1:-1, -1, 1] == u[1:-1, -1, 1]).all()
    assert (uN[0, :, 1] == uN[1, :, 1]).all() and (uN[-1, :, 1] == uN[-2, :, 1]).all()
    u1 = np.asfortranarray(u[:, :, 0].copy())
    L.oc_bc_scalar(2, Ng, O._ptr(u1), 0, 1)  # BCv1D!(view(u,:,:,1),1) ≡ BCf!(1,·)
    assert (u1 == uN[:, :, 0]).all()

    f = np.asfortranarray(rng.random((6, 6)))
    al = np.asfortranarray(rng.random((6, 6)))
    nh = np.asfortranarray(rng.random((6, 6, 2)))
    al0, n0 = al.copy(), nh.copy()
    L.oc_bc_vof(2, Ng, O._ptr(f), O._ptr(al), O._ptr(nh), 0)
    assert (f[0, :] == f[1, :]).all() and (f[-1, :] == f[-2, :]).all()
    assert (al[0, :] == al0[0, :]).all() and (nh[0, :, :] == n0[0, :, :]).all()
    f = np.asfortranarray(rng.random((6, 6)))
    al = np.asfortranarray(rng.random((6, 6)))
    nh = np.asfortranarray(rng.random((6, 6, 2)))
    L.oc_bc_vof(2, Ng, O._ptr(f), O._ptr(al), O._ptr(nh), 0b11)
    assert (f[0, :] == f[-2, :]).all() and (al[0, :] == al[-2, :]).all()
    assert (nh[0, :, :] == nh[-2, :, :]).all()


def test_applyvof_kat():
    """maintests.jl:131-136 (applyVOF! on a 2x2 plane) and :235-240 (cVOF ctor, ghosts)."""
    sdf = lambda x: (-x[0] - 3 * x[1] + 4.5) / np.sqrt(10)
    fref_inner = np.array([[0, 2 / 3], [1 / 24, 23 / 24]])
    # A 2x2 grid cannot host a MultiLevelPoisson, so the same plane is put on a
    # 8x8 grid: its cells (1..2, 1..2) see identical SDF values.
    sim = O.OracleSim((8, 8), (0.0, 0.0), T=np.float64, InterfaceSDF=sdf)
    f = sim.f
    assert np.allclose(f[1:3, 1:3], fref_inner, atol=1e-12)
    # ghosts are Neumann copies of the interior (cVOF.jl:64 BCf!)
    assert (f[0, :] == f[1, :]).all() and (f[:, 0] == f[:, 1]).all()


def test_normal_kats():
    """maintests.jl:139-156 (WY and MYC)."""
    L = O.lib(np.float64)
    Ng = O._ints([3, 3])
    I = O._ints([2, 2])
    out = np.zeros(2)
    f = _blk([[5 / 12, 1, 2 / 3], [1 / 4, 11 / 12, 1 / 12], [1 / 12, 1 / 3, 0]])
    L.oc_normal_block(2, Ng, O._ptr(f), O.NORMAL["WY"], I, O._ptr(out))
    assert abs(out[0] - 1) < 1e-12 and abs(out[1] + 0.5 - 0.5) < 1e-12
    f = _blk([[0, 1 / 3, 1], [1 / 12, 11 / 12, 1], [1, 1, 1]])
    L.oc_normal_block(2, Ng, O._ptr(f), O.NORMAL["WY"], I, O._ptr(out))
    assert abs(out[0] + 2 / 3) < 1e-12 and abs(out[1] + 1) < 1e-12
    f = _blk([[0, 0, 0], [0, 0.1, 0], [0, 0, 0]])
    L.oc_normal_block(2, Ng, O._ptr(f), O.NORMAL["WY"], I, O._ptr(out))
    assert out[0] == 1 and out[1] == 0
    L.oc_normal_block(2, Ng, O._ptr(f), O.NORMAL["MYC"], I, O._ptr(out))
    assert out[0] == 0.5 and out[1] == 0.5


def test_normal_pcd_youngs_slic_closed_forms():
    """normalEstimation.jl:161-176 (PCD, SLIC) and :232-257 (Youngs) on f linear in x:
    f = 0.2, 0.5, 0.8 along x, constant along y.  PCD = f[I-δ] - f[I+δ] = (-0.6, 0);
    SLIC keeps the sign of the dominant PCD axis = (-1, 0); Youngs sums 2x2 stencils
    either side (0.8 and 3.2 along x, 2.0 and 2.0 along y) -> (-1.2, 0) / 1.2."""
    L = O.lib(np.float64)
    Ng = O._ints([3, 3])
    I = O._ints([2, 2])
    out = np.zeros(2)
    f = _blk([[0.2, 0.2, 0.2], [0.5, 0.5, 0.5], [0.8, 0.8, 0.8]])
    L.oc_normal_block(2, Ng, O._ptr(f), O.NORMAL["PCD"], I, O._ptr(out))
    assert abs(out[0] + 0.6) < 1e-15 and out[1] == 0
    L.oc_normal_block(2, Ng, O._ptr(f), O.NORMAL["SLIC"], I, O._ptr(out))
    assert out[0] == -1 and out[1] == 0
    L.oc_normal_block(2, Ng, O._ptr(f), O.NORMAL["Youngs"], I, O._ptr(out))
    assert abs(out[0] + 1) < 1e-15 and out[1] == 0


def test_vof_flux_kats():
    """maintests.jl:158-180 (getVOFFlux! per face)."""
    L = O.lib(np.float64)
    Ng = O._ints([3, 3])
    f = np.zeros((3, 3), order="F")
    f[1, 1] = 0.32
    al = np.zeros((3, 3), order="F")
    al[1, 1] = -0.2
    nh = np.zeros((3, 3, 2), order="F")
    nh[1, 1, :] = [1, -1]
    ruf = np.zeros((3, 3, 2), order="F")
    ff = np.zeros((3, 3), order="F")
    lr = 0.1
    args = lambda: (O._ptr(ff), O._ptr(f), O._ptr(al), O._ptr(nh))
    L.oc_vof_flux_face(2, Ng, *args(), -0.4, 1, O._ints([2, 2]), O._ptr(ruf), lr)
    L.oc_vof_flux_face(2, Ng, *args(), 0.4, 1, O._ints([3, 2]), O._ptr(ruf), lr)
    assert abs(ff[1, 1] + 0.24) < 1e-12 and abs(ff[2, 1] - 0.02) < 1e-12
    assert abs(ruf[1, 1, 0] + 0.256) < 1e-12 and abs(ruf[2, 1, 0] - 0.058) < 1e-12
    L.oc_vof_flux_face(2, Ng, *args(), -0.4, 2, O._ints([2, 2]), O._ptr(ruf), lr)
    L.oc_vof_flux_face(2, Ng, *args(), 0.4, 2, O._ints([2, 3]), O._ptr(ruf), lr)
    assert abs(ff[1, 1] + 0.02) < 1e-12 and abs(ff[1, 2] - 0.24) < 1e-12
    assert abs(ruf[1, 1, 1] + 0.058) < 1e-12 and abs(ruf[1, 2, 1] - 0.256) < 1e-12


def test_cvof_ctor_kats():
    """maintests.jl:219-242: without an SDF f ≡ 1, α = n̂ = 0; the SDF plane incl. ghosts."""
    sim = O.OracleSim((8, 8), (0.0, 0.0), T=np.float64)
    assert (sim.f == 1).all() and (sim.get(O.ALPHA) == 0).all() and (sim.get(O.NHAT) == 0).all()
    sdf = lambda x: (-x[0] - 3 * x[1] + 4.5) / np.sqrt(10)
    sim = O.OracleSim((8, 8), (0.0, 0.0), T=np.float64, InterfaceSDF=sdf, perdir=())
    fref = np.array([[0, 0, 2 / 3, 2 / 3], [0, 0, 2 / 3, 2 / 3],
                     [1 / 24, 1 / 24, 23 / 24, 23 / 24], [1 / 24, 1 / 24, 23 / 24, 23 / 24]])
    # on the 8x8 grid the low ghost row/column and the first two interior cells
    # coincide with the 2x2 case (indices 0..2 of fRef)
    assert np.allclose(sim.f[0:3, 0:3], fref[0:3, 0:3], atol=1e-12)


def _quiescent(N=8, T=np.float64, **kw):
    """helper.jl:17-18."""
    return O.OracleSim((N, N), (0.0, 0.0), float(N), T=T, perdir=(1, 2),
                       InterfaceSDF=lambda x: np.sqrt((x[0] - N / 2) ** 2 + (x[1] - N / 2) ** 2) - N / 4, **kw)


def _tgv(N=16, T=np.float64, λμ=0.1, λρ=0.1, **kw):
    """helper.jl:6-13."""
    κ = T(2 * np.pi / N)
    R = T(N / 4)

    def u0(i, x):
        return -np.sin(x[0] * κ) * np.cos(x[1] * κ) if i == 1 else np.cos(x[0] * κ) * np.sin(x[1] * κ)

    sdf = lambda x: np.sqrt((x[0] - N / 2) ** 2 + (x[1] - N / 2) ** 2) - R
    return O.OracleSim((N, N), (0.0, 0.0), float(N), T=T, λμ=λμ, λρ=λρ, InterfaceSDF=sdf, u0=u0,
                       perdir=(1, 2), **kw)


def test_quiescent_step_is_noop():
    """maintests.jl:193-202: zero velocity, no forcing -> u stays 0, f bit-identical, p finite."""
    sim = _quiescent()
    f0 = sim.f.copy()
    sim.sim_step()
    assert (sim.u == 0).all()
    assert (sim.f == f0).all()
    assert np.isfinite(sim.p).all()


def test_tgv_droplet_volume():
    """maintests.jl:204-215: TGV droplet, Σf conserved to rtol 1e-4 over 3 steps after the first."""
    sim = _tgv()
    sim.sim_step()
    V0 = O.inside(sim.f).sum()
    for _ in range(3):
        sim.sim_step()
    assert np.isfinite(sim.u).all() and np.isfinite(sim.f).all()
    assert abs(O.inside(sim.f).sum() - V0) <= 1e-4 * abs(V0)


def test_twophase_ctor_kat():
    """maintests.jl:244-261 (construction only): u_BC=(1,0), periodic in y."""
    sim = O.OracleSim((8, 8), (1.0, 0.0), 8.0, T=np.float64, InterfaceSDF=lambda x: x[0] - 4, perdir=(2,))
    assert sim.L == 8.0 and sim.U == 1.0
    f = sim.f
    assert (O.inside(f)[:4, :] == 1).all() and (O.inside(f)[4:, :] == 0).all()


def test_surface_tension_kats():
    """maintests.jl:182-191: getPopinetHeight and the 2D getCurvature (rows are x)."""
    L = O.lib(np.float64)
    f = _blk([[0.3, 0.2, 0.1, 0.1, 0.2, 0.3, 0.0, 0.0],
              [1.0, 1.0, 0.6, 0.5, 0.3, 0.2, 0.0, 0.0],
              [0.0, 0.0, 0.0, 0.1, 0.0, 0.0, 0.0, 0.0]])
    Ng = O._ints([3, 8])
    H = [L.oc_popinet_height(2, Ng, O._ptr(f), O._ints([x, 5]), 2) for x in (1, 2, 3)]
    assert H == [-0.3, -0.9, -1.4]
    assert abs(L.oc_curvature(2, Ng, O._ptr(f), O._ints([2, 5]), 2) - 0.0672718547928328) < 1e-15


def test_metrics_kats():
    """maintests.jl:321-347 (ρkeI, ρuI, ρgh, EnsI in 2D and 3D)."""
    L = O.lib(np.float64)
    u = np.zeros((4, 4, 2), order="F")
    u[1, 1, 0] = 1; u[2, 1, 0] = 2; u[1, 1, 1] = 3; u[1, 2, 1] = 4
    f = np.zeros((4, 4), order="F"); f[1, 1] = 0.5
    Ng, I = O._ints([4, 4]), O._ints([2, 2])
    d = lambda *v: np.ascontiguousarray(v, dtype=np.float64)
    U0, U1 = d(0, 0, 0), d(1, 1, 0)
    assert abs(L.oc_rhoke(2, Ng, O._ptr(u), O._ptr(f), 0.2, I, O._ptr(U0)) - 4.5) < 1e-14
    assert abs(L.oc_rhoke(2, Ng, O._ptr(u), O._ptr(f), 0.2, I, O._ptr(U1)) - 2.1) < 1e-14
    assert abs(L.oc_rhou(2, Ng, 1, O._ptr(u), O._ptr(f), 0.2, I, O._ptr(U0)) - 0.9) < 1e-14
    assert abs(L.oc_rhou(2, Ng, 2, O._ptr(u), O._ptr(f), 0.2, I, O._ptr(U0)) - 2.1) < 1e-14
    assert abs(L.oc_rhou(2, Ng, 1, O._ptr(u), O._ptr(f), 0.2, I, O._ptr(U1)) - 0.3) < 1e-14
    assert abs(L.oc_rhogh(2, Ng, O._ptr(d(0, -1, 0)), O._ptr(f), 0.2, I, O._ptr(d(0, 0, 0))) - 0.3) < 1e-14
    w = np.zeros((4, 4), order="F"); w[1, 1] = 1; w[2, 1] = 2; w[1, 2] = 3; w[2, 2] = 4
    assert abs(L.oc_ens(2, Ng, O._ptr(w), I) - 3.75) < 1e-14
    w3 = np.zeros((4, 4, 4, 3), order="F")
    for c, cells in enumerate([((1, 1, 1), (1, 2, 1), (1, 1, 2), (1, 2, 2)),
                               ((1, 1, 1), (1, 1, 2), (2, 1, 1), (2, 1, 2)),
                               ((1, 1, 1), (2, 1, 1), (1, 2, 1), (2, 2, 1))]):
        for v, cell in enumerate(cells):
            w3[cell + (c,)] = v + 1
    assert abs(L.oc_ens(3, O._ints([4, 4, 4]), O._ptr(w3), O._ints([2, 2, 2])) - 11.25) < 1e-14
