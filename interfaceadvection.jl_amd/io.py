"""State I/O of a TwoPhaseSimulation: binary checkpoints, and the VTK
ImageData / PVD collection files that WaterLily's `vtkWriter` writes and the
reference's `load!(sim, Val(:pvd); fname)` reads (ext/IntfAdvReadVTKExt.jl:29-50).

Checkpoints are `.npz` archives of plain arrays (no pickled objects): the
fields f, f⁰, u, u⁰, p, the Δt history and the grid and element type they
belong to.  Restoring one into a simulation built with the same parameters
reproduces the run bit for bit.

VTK files hold every array entry, ghosts included, as point data on a grid of
Ng points per axis (extent 0..Ng-1), vectors with their components first, in
base64-encoded little-endian binary with a UInt64 byte-count header -- the
layout ParaView and ReadVTK.jl read.  The reader also takes appended raw data
and zlib-compressed blocks, the layouts other VTK writers use.
"""
import base64
import json
import os
import xml.etree.ElementTree as ET
import zlib

import numpy as np

_VTK_TYPE = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64", np.dtype(np.int8): "Int8"}
_NP_TYPE = {v: k for k, v in _VTK_TYPE.items()}


# ----------------------------------------------------------------------------- checkpoints
def save_checkpoint(sim, path):
    """Write f, f⁰, u, u⁰, p and the Δt history of `sim` to `path` (.npz)."""
    meta = dict(dims=list(sim.dims), T=np.dtype(sim.T).name, Δt=[float(x) for x in sim.flow.Δt])
    np.savez(path, f=sim.intf.f, f0=sim.intf.f0, u=sim.flow.u, u0=sim.flow.u0, p=sim.flow.p,
             meta=np.array(json.dumps(meta)))


def load_checkpoint(sim, path):
    """Restore a checkpoint written by `save_checkpoint` into `sim`, which must
    have the same grid and element type."""
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        if list(sim.dims) != meta["dims"] or np.dtype(sim.T).name != meta["T"]:
            raise ValueError(f"checkpoint is for dims={meta['dims']} T={meta['T']}, "
                             f"the simulation has dims={list(sim.dims)} T={np.dtype(sim.T).name}")
        sim.intf.f = z["f"]
        sim.intf.f0 = z["f0"]
        sim.flow.u = z["u"]
        sim.flow.u0 = z["u0"]
        sim.flow.p = z["p"]
    sim.flow.Δt[:] = [sim.T.type(x) for x in meta["Δt"]]


# ----------------------------------------------------------------------------- VTK ImageData
def _encode(a):
    raw = np.ascontiguousarray(a).tobytes()
    return base64.b64encode(np.uint64(len(raw)).tobytes() + raw).decode("ascii")


def write_vti(path, fields, Ng):
    """Write `fields` ({name: array}) on a grid of Ng points per axis.  A field of
    shape Ng is a scalar; one of shape Ng + (D,) a vector (components first in
    the file)."""
    D = len(Ng)
    ext = " ".join(f"0 {Ng[a] - 1}" if a < D else "0 0" for a in range(3))
    root = ET.Element("VTKFile", type="ImageData", version="1.0", byte_order="LittleEndian", header_type="UInt64")
    img = ET.SubElement(root, "ImageData", WholeExtent=ext, Origin="0 0 0", Spacing="1 1 1")
    piece = ET.SubElement(img, "Piece", Extent=ext)
    pd = ET.SubElement(piece, "PointData")
    for name, a in fields.items():
        a = np.asarray(a)
        if a.shape == tuple(Ng):
            nc, flat = 1, a.ravel(order="F")
        elif a.shape == tuple(Ng) + (D,):
            nc, flat = D, np.moveaxis(a, -1, 0).ravel(order="F")
        else:
            raise ValueError(f"{name}: shape {a.shape} is neither {tuple(Ng)} nor {tuple(Ng) + (D,)}")
        da = ET.SubElement(pd, "DataArray", type=_VTK_TYPE[a.dtype], Name=name, NumberOfComponents=str(nc),
                           format="binary")
        da.text = _encode(flat.astype(a.dtype.newbyteorder("<"), copy=False))
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")


def _blocks(buf, pos, hdr, compressed):
    """One DataArray's bytes starting at `pos` of `buf`: a byte-count header
    then the raw bytes, or with zlib compression the block table
    [nblocks, block size, last block size, compressed sizes...] then the
    compressed blocks (the VTK XML binary layout)."""
    hs = np.dtype(hdr).itemsize
    if not compressed:
        n = int(np.frombuffer(buf[pos:pos + hs], dtype=hdr)[0])
        return buf[pos + hs:pos + hs + n]
    nb = int(np.frombuffer(buf[pos:pos + hs], dtype=hdr)[0])
    table = np.frombuffer(buf[pos:pos + (3 + nb) * hs], dtype=hdr).astype(np.int64)
    q = pos + (3 + nb) * hs
    out = []
    for k in range(nb):
        c = int(table[3 + k])
        out.append(zlib.decompress(buf[q:q + c]))
        q += c
    return b"".join(out)


def _inline_blocks(text, hdr, compressed):
    """An inline base64 DataArray: uncompressed, one base64 stream of header +
    bytes; zlib-compressed, the block table and the blocks as two streams."""
    if not compressed:
        return _blocks(base64.b64decode(text), 0, hdr, False)
    hs = np.dtype(hdr).itemsize
    first = base64.b64decode(text[:4 * ((hs + 2) // 3)])
    nb = int(np.frombuffer(first[:hs], dtype=hdr)[0])
    hl = 4 * (((3 + nb) * hs + 2) // 3)
    table = base64.b64decode(text[:hl])
    return _blocks(table + base64.b64decode(text[hl:]), 0, hdr, True)


def read_vti(path):
    """Read a VTK ImageData file: binary DataArrays inline (base64) or appended
    (raw, after the AppendedData marker), uncompressed or zlib-compressed
    (vtkZLibDataCompressor), UInt32 or UInt64 headers -- what `write_vti`
    writes and what WriteVTK-based writers write by default.  Returns
    ({name: array}, Ng)."""
    data = open(path, "rb").read()
    raw = None
    i = data.find(b"<AppendedData")
    if i >= 0:
        j = data.index(b"_", data.index(b">", i)) + 1
        raw = data[j:]
        data = data[:i] + b"</VTKFile>"
    root = ET.fromstring(data)
    comp = root.get("compressor")
    if comp not in (None, "", "vtkZLibDataCompressor"):
        raise ValueError(f"unsupported VTK compressor {comp}")
    compressed = bool(comp)
    hdr = np.uint64 if root.get("header_type", "UInt32") == "UInt64" else np.uint32
    img = root.find("ImageData")
    e = [int(v) for v in img.get("WholeExtent").split()]
    Ng = tuple(e[2 * a + 1] - e[2 * a] + 1 for a in range(3))
    D = 3 if Ng[2] > 1 else 2
    Ng = Ng[:D]
    out = {}
    for da in img.find("Piece").find("PointData").findall("DataArray"):
        fmt = da.get("format")
        if fmt == "binary":
            buf = _inline_blocks(da.text.strip(), hdr, compressed)
        elif fmt == "appended":
            if raw is None:
                raise ValueError("appended DataArray without AppendedData")
            buf = _blocks(raw, int(da.get("offset")), hdr, compressed)
        else:
            raise ValueError(f"DataArray format {fmt!r} is not supported (binary or appended)")
        dt = np.dtype(_NP_TYPE[da.get("type")]).newbyteorder("<")
        flat = np.frombuffer(buf, dtype=dt)
        nc = int(da.get("NumberOfComponents", "1"))
        if nc == 1:
            out[da.get("Name")] = flat.reshape(Ng, order="F").astype(dt.newbyteorder("="))
        else:
            a = flat.reshape((nc,) + Ng, order="F")
            out[da.get("Name")] = np.asfortranarray(np.moveaxis(a, 0, -1)).astype(dt.newbyteorder("="))
    return out, Ng


# ----------------------------------------------------------------------------- PVD collections
def _write_pvd(fname, entries):
    root = ET.Element("VTKFile", type="Collection", version="1.0", byte_order="LittleEndian")
    col = ET.SubElement(root, "Collection")
    for t, f in entries:
        ET.SubElement(col, "DataSet", timestep=repr(float(t)), part="0", file=f)
    ET.ElementTree(root).write(fname, xml_declaration=True, encoding="utf-8")


def read_pvd(fname):
    """[(timestep, vti path)] of a collection, paths relative to the .pvd's folder resolved."""
    base = os.path.dirname(os.path.abspath(fname))
    col = ET.parse(fname).getroot().find("Collection")
    return [(float(d.get("timestep")), os.path.join(base, d.get("file"))) for d in col.findall("DataSet")]


def default_attrib():
    """The fields WaterLily's vtkWriter saves for a TwoPhaseSimulation: u, p, f."""
    return {"u": lambda s: s.flow.u, "p": lambda s: s.flow.p, "f": lambda s: s.intf.f}


class VTKWriter:
    """vtkWriter(fname; dir, attrib): each `save(sim)` writes `dir/fname_NNNNNN.vti`
    and appends it to `fname.pvd` with timestep = sim_time(sim), the
    dimensionless time t·U/L."""

    def __init__(self, fname="WaterLily", dir=".", attrib=None, append=False):
        self.fname, self.dir = fname, dir
        self.attrib = attrib or default_attrib()
        os.makedirs(dir, exist_ok=True)
        pvd = os.path.join(dir, fname + ".pvd")
        self.entries = []
        if append and os.path.exists(pvd):
            self.entries = [(t, os.path.relpath(p, dir)) for t, p in read_pvd(pvd)]

    def save(self, sim):
        k = len(self.entries) + 1
        name = f"{self.fname}_{k:06d}.vti"
        write_vti(os.path.join(self.dir, name), {n: get(sim) for n, get in self.attrib.items()}, sim.Ng)
        self.entries.append((sim.sim_time(), name))
        _write_pvd(os.path.join(self.dir, self.fname + ".pvd"), self.entries)
        return name


def flow_cfl(u, ν, Δt_max=10):
    """WaterLily's single-phase `CFL(flow)`: min(Δt_max, inv(max_I σ[I] + 5ν)) with
    σ[I] = flux_out(I) = Σ_i max(0, u[I+δi, i]) + max(0, -u[I, i]) over the interior
    cells I, the terms summed in Float64 and stored in the element type T of u, as
    Julia promotes `max(0., x)`; the rest in T (WaterLily keeps σ in its Poisson
    source array, whose ghosts it leaves as they were; they take no part here)."""
    D = u.shape[-1]
    T = u.dtype.type
    inner = tuple(slice(1, n - 1) for n in u.shape[:D])
    s = np.zeros(tuple(n - 2 for n in u.shape[:D]))
    for i in range(D):
        up = tuple(slice(2, n) if a == i else slice(1, n - 1) for a, n in enumerate(u.shape[:D]))
        s += np.maximum(0.0, u[up + (i,)].astype(np.float64)) + np.maximum(0.0, -u[inner + (i,)].astype(np.float64))
    rate = T(s.astype(T).max(initial=T(0))) + T(5) * T(ν)
    return T(Δt_max) if rate == 0 else min(T(Δt_max), T(1) / rate)


def load_pvd(sim, fname="WaterLily.pvd"):
    """load!(sim, Val(:pvd); fname) (ext/IntfAdvReadVTKExt.jl:29-50): the last
    file of the collection overwrites p, f and u; the last Δt becomes the saved
    dimensionless time times L/U and the next one is WaterLily's single-phase
    CFL of the loaded velocity, as the reference pushes it.  Returns a writer
    that appends to the same collection."""
    entries = read_pvd(fname)
    t, path = entries[-1]
    fields, Ng = read_vti(path)
    if tuple(Ng) != tuple(sim.Ng):
        raise ValueError("The dimensions of the simulation do not match the dimensions of the vtk file.")
    if "p" in fields:
        sim.flow.p = fields["p"].astype(sim.T)
    if "f" in fields:
        sim.intf.f = fields["f"].astype(sim.T)
    sim.flow.u = fields["u"].astype(sim.T)
    sim.flow.Δt[-1] = sim.T.type(t * sim.L / sim.U)
    sim.flow.Δt.append(flow_cfl(sim.flow.u, sim.flow.ν))
    d = os.path.dirname(os.path.abspath(fname))
    return VTKWriter(os.path.splitext(os.path.basename(fname))[0], d, append=True)
