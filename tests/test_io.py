"""State I/O: VTK ImageData / PVD round trips (CPU), and checkpoint / VTK
restarts of a running simulation on the GPU path (bit-exact continuation)."""
import importlib
import os
import types

import numpy as np
import pytest

from __graft_entry__ import load_package

ia = load_package()
io = importlib.import_module(ia.__name__ + ".io")


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("D", [2, 3])
def test_vti_round_trip(tmp_path, dt, D):
    rng = np.random.default_rng(D)
    Ng = (6, 5, 4)[:D]
    s = np.asfortranarray(rng.standard_normal(Ng).astype(dt))
    v = np.asfortranarray(rng.standard_normal(Ng + (D,)).astype(dt))
    path = os.path.join(tmp_path, "a.vti")
    io.write_vti(path, {"p": s, "u": v}, Ng)
    got, Ng2 = io.read_vti(path)
    assert tuple(Ng2) == Ng
    assert got["p"].dtype == dt and (got["p"] == s).all()
    assert got["u"].shape == Ng + (D,) and (got["u"] == v).all()
    # components first in the file: the first D values are u at point (0, 0[, 0])
    import base64
    import xml.etree.ElementTree as ET
    da = [d for d in ET.parse(path).getroot().iter("DataArray") if d.get("Name") == "u"][0]
    raw = base64.b64decode(da.text)
    first = np.frombuffer(raw[8:8 + D * np.dtype(dt).itemsize], dtype=dt)
    assert (first == v[(0,) * D]).all()


def test_pvd_collection(tmp_path):
    """VTKWriter appends one file per save with timestep = sim_time."""
    Ng = (6, 6)
    state = dict(u=np.zeros(Ng + (2,)), p=np.zeros(Ng), f=np.zeros(Ng), t=0.0)
    fake = types.SimpleNamespace(Ng=Ng, flow=types.SimpleNamespace(), intf=types.SimpleNamespace())
    fake.sim_time = lambda: state["t"]
    w = io.VTKWriter("run", str(tmp_path))
    for k in range(3):
        state["t"] = 0.5 * k
        fake.flow.u, fake.flow.p, fake.intf.f = state["u"] + k, state["p"] - k, state["f"] + 0.1 * k
        w.save(fake)
    entries = io.read_pvd(os.path.join(tmp_path, "run.pvd"))
    assert [t for t, _ in entries] == [0.0, 0.5, 1.0]
    last, _ = io.read_vti(entries[-1][1])
    assert (last["u"] == 2).all() and (last["p"] == -2).all() and np.allclose(last["f"], 0.2)
    w2 = io.VTKWriter("run", str(tmp_path), append=True)
    assert len(w2.entries) == 3


def test_flow_cfl():
    """WaterLily CFL: 1/(max flux_out + 5ν), capped at 10; flux_out sums the
    outflow through the upper faces (u[I+δi] > 0) and the lower faces (u[I] < 0)."""
    u = np.zeros((6, 6, 2))
    assert io.flow_cfl(u, 0.0) == 10.0
    assert io.flow_cfl(u, 0.5) == 0.4
    u[2, 2, 0] = 1.0    # leaves cell (1,2) through its upper x face
    u[3, 3, 1] = -0.5   # leaves cell (3,3) through its lower y face
    u[4, 3, 0] = 0.75   # and (3,3) through its upper x face
    assert io.flow_cfl(u, 0.0) == 1 / 1.25
    assert io.flow_cfl(u, 0.1) == 1 / 1.75


@pytest.mark.gpu
def test_checkpoint_restart_is_bit_exact(tmp_path):
    from test_gpu_parity import bubble2d, say
    say("checkpoint after 3 steps, restart, 3 more steps")
    case = bubble2d(64, np.float32)
    a = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **case["kw"])
    for _ in range(3):
        ia.sim_step_(a)
    path = os.path.join(tmp_path, "ck.npz")
    io.save_checkpoint(a, path)
    for _ in range(3):
        ia.sim_step_(a)
    b = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **case["kw"])
    io.load_checkpoint(b, path)
    assert len(b.flow.Δt) == 4
    for _ in range(3):
        ia.sim_step_(b)
    assert (a.intf.f == b.intf.f).all() and (a.flow.u == b.flow.u).all() and (a.flow.p == b.flow.p).all()
    assert a.flow.Δt == b.flow.Δt
    a.close()
    b.close()


@pytest.mark.gpu
def test_vtk_save_and_load(tmp_path):
    from test_gpu_parity import bubble2d, say
    say("VTK save, load! from the collection")
    case = bubble2d(32, np.float32)
    case["kw"]["U"] = 1.0
    a = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **case["kw"])
    w = io.VTKWriter("WaterLily", str(tmp_path))
    for _ in range(2):
        ia.sim_step_(a)
        w.save(a)
    b = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **case["kw"])
    w2 = io.load_pvd(b, os.path.join(tmp_path, "WaterLily.pvd"))
    assert (b.flow.u == a.flow.u).all() and (b.flow.p == a.flow.p).all() and (b.intf.f == a.intf.f).all()
    assert b.flow.Δt[-2] == np.float32(a.sim_time() * a.L / a.U)
    assert b.flow.Δt[-1] == np.float32(io.flow_cfl(a.flow.u, a.flow.ν))
    assert len(w2.entries) == 2
    ia.sim_step_(b)
    assert np.isfinite(b.flow.u).all()
    w2.save(b)
    assert [t for t, _ in io.read_pvd(os.path.join(tmp_path, "WaterLily.pvd"))][-1] == b.sim_time()
    a.close()
    b.close()
