"""ia_attach_field: the library adopts a caller's device buffers as cVOF.f,
cVOF.f⁰, Flow.u, Flow.u⁰ and Flow.p (zero-copy, the way a Julia host whose
arrays are ROCArrays would hand them over).  A simulation running on attached
buffers must be bit-identical to one on the library's own, and after every
call the caller's buffers must hold the fields' current values (the sweeps
rotate f and f⁰ through a pool of four buffers; the library moves the values
back).  Device memory for the caller is allocated here through the HIP
runtime directly."""
import ctypes as C

import numpy as np
import pytest

from __graft_entry__ import load_package
from test_gpu_parity import bubble2d, bubbles3d, say

pytestmark = pytest.mark.gpu

ia = load_package()
_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipFree.argtypes = [C.c_void_p]
    return _hip


class DevBuf:
    """A caller-owned device buffer holding one field."""

    def __init__(self, host):
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), self.nbytes) == 0
        self.upload(host)

    def upload(self, host):
        a = np.asfortranarray(host)
        assert hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, 1) == 0

    def read(self):
        a = np.zeros(self.shape, dtype=self.dtype, order="F")
        assert hip().hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 2) == 0
        return a

    def free(self):
        hip().hipFree(self.ptr)


NAMES = ("f", "f0", "u", "u0", "p")


def attached_pair(case, **extra):
    kw = dict(case["kw"])
    kw.update(extra)
    a = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **kw)
    b = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **kw)
    bufs = {}
    for n in NAMES:
        bufs[n] = DevBuf(b._get(n))
        b._call("ia_attach_field", ia.FIELD[n], bufs[n].ptr, bufs[n].nbytes)
    return a, b, bufs


def check_same(a, b, bufs):
    for n in NAMES:
        want = a._get(n)
        assert np.array_equal(bufs[n].read(), want), n   # the caller's buffer holds the field
        assert np.array_equal(b._get(n), want), n        # and the library reads the same


@pytest.mark.parametrize("D", [2, 3])
def test_attached_buffers_are_bit_identical(D):
    say("attach", f"{D}D")
    if D == 2:
        case, extra = bubble2d(32, np.float64), {}
    else:
        case, extra = bubbles3d(32, np.float32), dict(poisson="refined")
    a, b, bufs = attached_pair(case, **extra)
    check_same(a, b, bufs)
    for n in range(3):
        ia.sim_step_(a)
        ia.sim_step_(b)
        assert a.flow.Δt[-1] == b.flow.Δt[-1]
        check_same(a, b, bufs)
    # advect! of f⁰ (the other pool slot), then of f, with the step's velocities
    for fld in ("f0", "f"):
        ia.advect_(a, f=fld, dt=0.2)
        ia.advect_(b, f=fld, dt=0.2)
        check_same(a, b, bufs)
    # the caller writes its buffer directly; the next step uses it
    u = bufs["u"].read()
    u *= 0.5
    bufs["u"].upload(u)
    a.flow.u = u
    ia.sim_step_(a)
    ia.sim_step_(b)
    check_same(a, b, bufs)
    a.close()
    b.close()
    for v in bufs.values():
        v.free()


def test_attach_rejects_bad_buffers():
    case = bubble2d(16, np.float64)
    s = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **case["kw"])
    f = s._get("f")
    host = np.zeros_like(f)
    with pytest.raises(ia.IAError):  # host memory
        s._call("ia_attach_field", ia.FIELD["f"], host.ctypes.data_as(C.c_void_p), host.nbytes)
    buf = DevBuf(f)
    with pytest.raises(ia.IAError):  # wrong size
        s._call("ia_attach_field", ia.FIELD["f"], buf.ptr, buf.nbytes - 8)
    with pytest.raises(ia.IAError):  # not an attachable field
        s._call("ia_attach_field", ia.FIELD["sigma"], buf.ptr, buf.nbytes)
    s.close()
    buf.free()
