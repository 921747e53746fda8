"""Symmetries of whole steps on the device, FP64: MPFMomStep! (flow.jl:60-109)
through the C ABI commutes with a translation by whole cells on a periodic
grid and with the reflection of a periodic axis.  The device's tiles sit at
other places relative to the data in the moved case, so these check the
periodic images, the tile edges and every upwind choice of the HIP kernels
on configurations beside the oracle comparisons.  The unmoved cases are also
compared with the oracle, at the parity suite's bars.
"""
import numpy as np
import pytest

from __graft_entry__ import load_package
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ia = load_package()
N = 32


def say(*a):
    print("[case]", *a, flush=True)


def _translation_case(shift):
    """A droplet (radius 5.2 cells, density and viscosity ratios 0.1) in a
    Taylor-Green field plus a uniform drift, fully periodic, moved by
    `shift` cells."""
    k = 2 * np.pi / N
    cx, cy = 10.3 + shift[0], 9.6 + shift[1]

    def sdf(x):
        return np.sqrt((x[0] - cx) ** 2 + (x[1] - cy) ** 2) - 5.2

    def u0(i, x):
        a, b = k * (x[0] - shift[0]), k * (x[1] - shift[1])
        return -np.sin(a) * np.cos(b) + 0.3 if i == 1 else np.cos(a) * np.sin(b) + 0.2

    return dict(T=np.float64, λμ=0.1, λρ=0.1, ν=0.02, InterfaceSDF=sdf, u0=u0, perdir=(1, 2))


def _mirror_case(mirror, lim):
    """A bubble below a wavy free surface, x periodic, slip walls in y, gravity
    -y, density ratio 1000; with `mirror` reflected x -> N - x (u_x negated)."""
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

    return dict(T=np.float64, λμ=1e-2, λρ=1e-3, ν=0.02, g=(0.0, -0.01), InterfaceSDF=sdf, u0=u0, perdir=(1,))


def _device(kw, lim="Koren"):
    return ia.TwoPhaseSimulation((N, N), (0.0, 0.0), float(N), λ=lim, poisson_tol=1e-12, poisson_itmx=400, **kw)


def _oracle(kw, lim="Koren"):
    return O.OracleSim((N, N), (0.0, 0.0), float(N), limiter=lim, mg_tol=1e-12, mg_itmx=400, **kw)


def cells(x):
    return np.asarray(x, dtype=np.float64)[1:-1, 1:-1]


def test_device_steps_are_translation_equivariant():
    """Shift (16, 16): every multigrid level of 32² stays aligned, so only the
    order of the sums changes.  Bars: f within 1e-11, u within 1e-11 of
    max|u|, between the moved and the unmoved device runs (measured 4e-15
    and 5e-16); the unmoved run against the oracle within 1e-10 (f) and 1e-4
    of max|u| (u)."""
    shift = (16, 16)
    say("device FP64 32^2 periodic droplet, 3 steps, shift", shift)
    a, b = _device(_translation_case((0, 0))), _device(_translation_case(shift))
    o = _oracle(_translation_case((0, 0)))
    for n in range(3):
        for s in (a, b):
            assert ia.sim_step_(s).status == 0
        o.sim_step()
        fa, fb = np.roll(cells(a.intf.f), shift, axis=(0, 1)), cells(b.intf.f)
        ua, ub = np.roll(cells(a.flow.u), shift, axis=(0, 1)), cells(b.flow.u)
        df, du = np.abs(fa - fb).max(), np.abs(ua - ub).max() / np.abs(ua).max()
        say(f"step {n + 1}: moved vs unmoved max|Δf| {df:.2e}, max|Δu|/max|u| {du:.2e}")
        assert df <= 1e-11 and du <= 1e-11
        assert np.abs(cells(a.intf.f) - cells(o.f)).max() <= 1e-10
        assert np.abs(cells(a.flow.u) - cells(o.u)).max() <= 1e-4 * np.abs(cells(o.u)).max()
    a.close()
    b.close()


def _translation_case3(shift, n=32):
    """The 3D counterpart: a ball (radius 5.2 cells) in a periodic
    Arnold-Beltrami-Childress-like field plus a drift, fully periodic n³."""
    k = 2 * np.pi / n
    c = np.array([8.3, 9.6, 7.2]) + np.array(shift)

    def sdf(x):
        return np.sqrt(sum((x[a] - c[a]) ** 2 for a in range(3))) - 5.2

    def u0(i, x):
        X, Y, Z = (k * (x[a] - shift[a]) for a in range(3))
        return {1: 0.3 * np.sin(Z) + 0.2 * np.cos(Y) + 0.1, 2: 0.2 * np.sin(X) + 0.3 * np.cos(Z) - 0.05,
                3: 0.25 * np.sin(Y) + 0.2 * np.cos(X) + 0.07}[i]

    return dict(T=np.float64, λμ=0.1, λρ=0.1, ν=0.02, InterfaceSDF=sdf, u0=u0, perdir=(1, 2, 3))


def test_device_3d_steps_are_translation_equivariant():
    """3D, 32³ fully periodic, shift (16, 16, 16) (every level aligned): the
    3D kernels (the WH normal from the donor's 3×3×3 block, the column
    tilings of the solver) at other places relative to the data.  Bars as in
    2D: moved against unmoved within 1e-11 (f) and 1e-11 of max|u| (u)
    (measured 7e-15 and 6e-16); the unmoved run against the oracle within
    1e-10 (f) and 1e-4 of max|u|."""
    shift, n = (16, 16, 16), 32
    say("device FP64 32^3 periodic ball, 2 steps, shift", shift)
    mk = lambda kw: ia.TwoPhaseSimulation((n,) * 3, (0.0,) * 3, float(n), poisson_tol=1e-12, poisson_itmx=400, **kw)
    a, b = mk(_translation_case3((0, 0, 0))), mk(_translation_case3(shift))
    o = O.OracleSim((n,) * 3, (0.0,) * 3, float(n), mg_tol=1e-12, mg_itmx=400, **_translation_case3((0, 0, 0)))
    c3 = lambda x: np.asarray(x, dtype=np.float64)[1:-1, 1:-1, 1:-1]
    for step in range(2):
        for s in (a, b):
            assert ia.sim_step_(s).status == 0
        o.sim_step()
        fa, fb = np.roll(c3(a.intf.f), shift, axis=(0, 1, 2)), c3(b.intf.f)
        ua, ub = np.roll(c3(a.flow.u), shift, axis=(0, 1, 2)), c3(b.flow.u)
        df, du = np.abs(fa - fb).max(), np.abs(ua - ub).max() / np.abs(ua).max()
        say(f"step {step + 1}: moved vs unmoved max|Δf| {df:.2e}, max|Δu|/max|u| {du:.2e}")
        assert df <= 1e-11 and du <= 1e-11
        assert np.abs(c3(a.intf.f) - c3(o.f)).max() <= 1e-10
        assert np.abs(c3(a.flow.u) - c3(o.u)).max() <= 1e-4 * np.abs(c3(o.u)).max()
    a.close()
    b.close()


@pytest.mark.parametrize("lim", ["Koren", "superbee"])
def test_device_steps_are_mirror_symmetric_along_a_periodic_axis(lim):
    """x -> N - x on the periodic axis (x-face q maps to N + 2 - q, face 1 to
    itself, u_x negated).  Bars: f within 1e-11, u within 1e-11 of max|u|
    (measured 8e-15 and 2.2e-12: the solves stop at Σr² <= 1e-12, where
    their summation orders differ); the unmoved run against the oracle
    within 1e-10 (f) and 1e-4 of max|u|."""
    say("device FP64 32^2 bubble under a wavy surface, x periodic, 3 steps, mirror,", lim)
    a, b = _device(_mirror_case(False, lim), lim), _device(_mirror_case(True, lim), lim)
    o = _oracle(_mirror_case(False, lim), lim)
    q = np.arange(1, N + 1)
    qm = (N + 2 - q - 1) % N + 1
    for n in range(3):
        for s in (a, b):
            assert ia.sim_step_(s).status == 0
        o.sim_step()
        ua_all, ub_all = np.asarray(a.flow.u, dtype=np.float64), np.asarray(b.flow.u, dtype=np.float64)
        scale = np.abs(ua_all).max()
        df = np.abs(cells(a.intf.f) - cells(b.intf.f)[::-1, :]).max()
        dux = np.abs(ua_all[q, 1:-1, 0] + ub_all[qm, 1:-1, 0]).max() / scale
        duy = np.abs(ua_all[1:-1, 1:-1, 1] - ub_all[1:-1, 1:-1, 1][::-1, :]).max() / scale
        say(f"step {n + 1}: mirrored vs unmirrored max|Δf| {df:.2e}, u_x {dux:.2e}, u_y {duy:.2e}")
        assert df <= 1e-11 and dux <= 1e-11 and duy <= 1e-11
        assert np.abs(cells(a.intf.f) - cells(o.f)).max() <= 1e-10
        assert np.abs(cells(a.flow.u) - cells(o.u)).max() <= 1e-4 * np.abs(cells(o.u)).max()
    a.close()
    b.close()
