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


def _hand_vti(path, arrays, Ng, fmt, zlib_blocks, hdr):
    """A VTK ImageData file written by hand, the way other VTK writers lay it
    out: format "binary" (inline base64) or "appended" (raw bytes after '_'),
    optionally vtkZLibDataCompressor blocks of `zlib_blocks` bytes, with UInt32
    or UInt64 headers."""
    import base64
    import zlib
    hs = np.dtype(hdr).itemsize

    def encode(raw):
        if zlib_blocks is None:
            return [np.array([len(raw)], dtype=hdr).tobytes() + raw]
        chunks = [raw[i:i + zlib_blocks] for i in range(0, len(raw), zlib_blocks)]
        comp = [zlib.compress(c) for c in chunks]
        table = np.array([len(chunks), zlib_blocks, len(chunks[-1])] + [len(c) for c in comp], dtype=hdr)
        return [table.tobytes(), b"".join(comp)]

    ext = " ".join(f"0 {Ng[a] - 1}" if a < len(Ng) else "0 0" for a in range(3))
    comp = ' compressor="vtkZLibDataCompressor"' if zlib_blocks else ""
    head = (f'<?xml version="1.0"?>\n<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" '
            f'header_type="{"UInt64" if hs == 8 else "UInt32"}"{comp}>\n'
            f'<ImageData WholeExtent="{ext}" Origin="0 0 0" Spacing="1 1 1">\n<Piece Extent="{ext}">\n<PointData>\n')
    body, appended = "", b""
    for name, a in arrays.items():
        nc = 1 if a.shape == tuple(Ng) else a.shape[-1]
        flat = np.moveaxis(a, -1, 0) if nc > 1 else a
        raw = np.asfortranarray(flat).tobytes(order="F")
        parts = encode(raw)
        t = {np.float32: "Float32", np.float64: "Float64"}[a.dtype.type]
        if fmt == "binary":
            text = "".join(base64.b64encode(p).decode() for p in parts)
            body += (f'<DataArray type="{t}" Name="{name}" NumberOfComponents="{nc}" format="binary">'
                     f'{text}</DataArray>\n')
        else:
            body += (f'<DataArray type="{t}" Name="{name}" NumberOfComponents="{nc}" format="appended" '
                     f'offset="{len(appended)}"/>\n')
            appended += b"".join(parts)
    tail = "</PointData>\n</Piece>\n</ImageData>\n"
    with open(path, "wb") as fh:
        fh.write((head + body + tail).encode())
        if fmt == "appended":
            fh.write(b'<AppendedData encoding="raw">\n_' + appended + b"\n</AppendedData>\n")
        fh.write(b"</VTKFile>\n")


@pytest.mark.parametrize("fmt,blocks,hdr", [("appended", None, np.uint64), ("appended", 96, np.uint32),
                                            ("binary", 96, np.uint64), ("binary", 1 << 15, np.uint32),
                                            ("binary", None, np.uint32)])
def test_read_vti_other_layouts(tmp_path, fmt, blocks, hdr):
    """Appended raw data (with '_' and bytes that look like XML) and zlib
    blocks, several blocks with a short last one, both header widths."""
    rng = np.random.default_rng(7)
    Ng = (7, 5, 3)
    s = np.asfortranarray(rng.standard_normal(Ng))
    s.flat[3] = np.frombuffer(b"<_>\n</Ap", dtype=np.float64)[0]
    v = np.asfortranarray(rng.standard_normal(Ng + (3,)).astype(np.float32))
    path = os.path.join(tmp_path, "o.vti")
    _hand_vti(path, {"p": s, "u": v}, Ng, fmt, blocks, hdr)
    got, Ng2 = io.read_vti(path)
    assert tuple(Ng2) == Ng
    assert got["p"].dtype == np.float64 and np.array_equal(got["p"], s)
    assert got["u"].dtype == np.float32 and np.array_equal(got["u"], v)


def test_read_vti_rejects_other_compressors(tmp_path):
    path = os.path.join(tmp_path, "c.vti")
    _hand_vti(path, {"p": np.zeros((3, 3))}, (3, 3), "binary", 64, np.uint64)
    txt = open(path, "rb").read().replace(b"vtkZLibDataCompressor", b"vtkLZ4DataCompressor")
    open(path, "wb").write(txt)
    with pytest.raises(ValueError):
        io.read_vti(path)


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
