"""Host-side mirror of InterfaceAdvection.jl's API over the gfx950 C ABI.

The reference's Julia host (TwoPhaseSimulation, sim_step!, MPFMomStep!,
advect!; InterfaceAdvection.jl:64-103, flow.jl:60, advection.jl:17) is
mirrored name for name, argument for argument; every call goes through
libia.so (include/ia.h).  There is no CPU code path: if libia.so is missing
or no GPU is visible, construction raises.

Julia's `!` cannot appear in a Python name, so `sim_step!` is `sim_step_`,
`MPFMomStep!` is `MPFMomStep_` and `advect!` is `advect_`.  Fields come back
as Fortran-ordered numpy arrays shaped like the Julia arrays: (Ng...,) for a
scalar and (Ng..., D) for a vector, Ng = N + 2.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libia.so")

IA_F32, IA_F64 = 0, 1
FIELD = dict(f=0, f0=1, alpha=2, nhat=3, cbar=5, rhou=6, rhouf=7, u=9, u0=10, force=11, p=12, sigma=13, mu0=14)
_VECTOR = {"nhat", "rhou", "rhouf", "u", "u0", "force", "mu0"}
POISSON = {"reference": 0, "refined": 1, "learned": 2, "plain": 3, "pcg": 4}
PROFILE = {"vof": 0, "momentum": 1, "visc_update": 2, "pcg_Ae": 3, "pcg_update": 4, "increment": 5,
           "load_inc": 6, "prolong_inc": 7, "fcg_Ap": 8, "fcg_update": 9, "pcg_last": 10, "cheb2_dots": 10, "redist": 11,
           "cg1_l1": 12, "cg1_l2": 13, "cg1_l3": 14}
STATUS = {0: "OK", 1: "NaN", 2: "divergence explosion", 3: "fill warning", 4: "HIP error", 5: "bad argument"}


class IAParams(C.Structure):
    _fields_ = [("D", C.c_int), ("N", C.c_int * 3), ("dtype", C.c_int), ("perdir", C.c_int),
                ("u_bc", C.c_double * 3), ("g", C.c_double * 3), ("has_g", C.c_int), ("nu", C.c_double),
                ("lambda_rho", C.c_double), ("lambda_mu", C.c_double), ("eta", C.c_double),
                ("normal_scheme", C.c_int), ("limiter", C.c_int), ("poisson_mode", C.c_int),
                ("poisson_tol", C.c_double), ("poisson_itmx", C.c_int), ("poisson_rtol", C.c_double),
                ("poisson_max_outer", C.c_int), ("device", C.c_int), ("exit_bc", C.c_int)]


class IASolveStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("r2", C.c_double), ("rel", C.c_double), ("ms", C.c_double)]


class IAStepStats(C.Structure):
    _fields_ = [("dt_next", C.c_double), ("status", C.c_int), ("fill_warnings", C.c_int),
                ("solve", IASolveStats * 2), ("vol", C.c_double), ("fmin", C.c_double), ("fmax", C.c_double)]


class IALearnedInfo(C.Structure):
    _fields_ = [("projections", C.c_int), ("accepted", C.c_int), ("train_steps", C.c_int),
                ("last_ratio", C.c_double)]


# parameters of the learned correction: 5 layers x (27 taps x 16 x 16) weights, then 5 x 16 biases
LEARNED_NPAR = 5 * (27 * 16 * 16 + 16)

_lib = None


def lib():
    """Load libia.so (built by `build()`); raises when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build the HIP extension first (__graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, sz = C.c_void_p, C.c_size_t
    sig = {
        "ia_create": (C.c_int, [C.POINTER(IAParams), C.POINTER(vp)]),
        "ia_destroy": (None, [vp]),
        "ia_set_field": (C.c_int, [vp, C.c_int, vp, sz]),
        "ia_get_field": (C.c_int, [vp, C.c_int, vp, sz]),
        "ia_field_device_ptr": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.POINTER(sz)]),
        "ia_attach_field": (C.c_int, [vp, C.c_int, vp, sz]),
        "ia_mg_levels": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int)]),
        "ia_apply_vof": (C.c_int, [vp, vp, sz]),
        "ia_set_initial_velocity": (C.c_int, [vp, vp, sz]),
        "ia_cfl": (C.c_int, [vp, C.POINTER(C.c_double)]),
        "ia_mom_step": (C.c_int, [vp, C.c_double, C.c_int, C.POINTER(IAStepStats)]),
        "ia_advect": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(IAStepStats)]),
        "ia_reconstruct": (C.c_int, [vp]),
        "ia_vof_sweep": (C.c_int, [vp, C.c_int, C.c_double, C.POINTER(IAStepStats)]),
        "ia_visc": (C.c_int, [vp, C.c_int]),
        "ia_update_L": (C.c_int, [vp, C.c_int]),
        "ia_mg_update": (C.c_int, [vp]),
        "ia_project": (C.c_int, [vp, C.c_double, C.POINTER(IASolveStats)]),
        "ia_synchronize": (C.c_int, [vp]),
        "ia_profile_enable": (C.c_int, [vp, C.c_int, C.c_int]),
        "ia_profile_read": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "ia_last_error": (C.c_char_p, [vp]),
        "ia_metrics": (C.c_int, [vp, vp, vp, vp]),
        "ia_metrics_noalloc": (C.c_int, [vp, vp, vp, vp]),
        "ia_levelset": (C.c_int, [vp]),
        "ia_redistance": (C.c_int, [vp, C.c_double, C.c_double, C.c_int]),
        "ia_learned_init": (C.c_int, [vp, C.c_uint]),
        "ia_learned_set_weights": (C.c_int, [vp, vp, sz]),
        "ia_learned_get_weights": (C.c_int, [vp, vp, sz]),
        "ia_learned_train": (C.c_int, [vp, C.c_int, C.c_double]),
        "ia_learned_stats": (C.c_int, [vp, C.POINTER(IALearnedInfo)]),
        "ia_learned_forward": (C.c_int, [vp, vp, vp]),
        "ia_learned_grad": (C.c_int, [vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class IAError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


# plug-ins selectable by name or by these function objects (InterfaceAdvection.jl:28, :13)
def Koren(u, c, d):
    """Koren limiter, flow.jl:7 (a selector here: the kernels implement it)."""
    return np.median([(7 * c + d - 2 * u) / 6, c, np.median([2 * c - u, c, d])])


def minmod(u, c, d):
    """minmod limiter, flow.jl:6."""
    return np.median([(3 * c - u) / 2, c, (c + d) / 2])


def upwind(u, c, d):
    """upwind, flow.jl:5."""
    return c


def vanAlbada1(u, c, d):
    """vanAlbada1 limiter, flow.jl:8-11."""
    a, b = c - u, d - c
    return c + max(a * b, 0) * (0 if a == b == 0 else (a + b) / (a * a + b * b)) / 2


def Sweby(u, c, d, γ=1.5):
    """Sweby limiter, flow.jl:12 (γ = 1.5; the kernels implement γ = 1.5 and γ = 2)."""
    s = np.sign(d - u)
    if c <= min(u, d) or c >= max(u, d):
        return c
    return c + s * max(0, min(s * γ * (c - u), s * (d - c)), min(s * (c - u), s * γ * (d - c))) / 2


def superbee(u, c, d):
    """superbee = Sweby(γ = 2), flow.jl:13."""
    return Sweby(u, c, d, 2)


def TVDcen(u, c, d):
    """TVDcen, flow.jl:14."""
    s = np.sign(d - u)
    return c if (c <= min(u, d) or c >= max(u, d)) else c + s * min(s * (c - u), s * (d - c) / 2)


def TVDdown(u, c, d):
    """TVDdown, flow.jl:15."""
    s = np.sign(d - u)
    return c if (c <= min(u, d) or c >= max(u, d)) else c + s * min(s * (c - u), s * (d - c))


def getInterfaceNormal_WH(f, nhat, I):
    """Selector for the WH normal scheme (normalEstimation.jl:105)."""
    raise TypeError("normal schemes run on the device; pass the function as normalScheme=")


def getInterfaceNormal_Column(f, nhat, I):
    """Selector for the Column normal scheme (normalEstimation.jl:75)."""
    raise TypeError("normal schemes run on the device; pass the function as normalScheme=")


def getInterfaceNormal_WY(f, nhat, I):
    """Selector for the WY normal scheme (normalEstimation.jl:35)."""
    raise TypeError("normal schemes run on the device; pass the function as normalScheme=")


def getInterfaceNormal_MYC(f, nhat, I):
    """Selector for the MYC normal scheme (normalEstimation.jl:185)."""
    raise TypeError("normal schemes run on the device; pass the function as normalScheme=")


def getInterfaceNormal_PCD(f, nhat, I):
    """Selector for the PCD normal scheme (normalEstimation.jl:161)."""
    raise TypeError("normal schemes run on the device; pass the function as normalScheme=")


def getInterfaceNormal_Y(f, nhat, I):
    """Selector for the Youngs normal scheme (normalEstimation.jl:232)."""
    raise TypeError("normal schemes run on the device; pass the function as normalScheme=")


def getInterfaceNormal_SLIC(f, nhat, I):
    """Selector for the SLIC normal scheme (normalEstimation.jl:172)."""
    raise TypeError("normal schemes run on the device; pass the function as normalScheme=")


_LIMITERS = {Koren: 0, minmod: 1, upwind: 2, vanAlbada1: 3, superbee: 4, TVDcen: 5, TVDdown: 6, Sweby: 7,
             "Koren": 0, "minmod": 1, "upwind": 2, "vanAlbada1": 3, "superbee": 4, "TVDcen": 5, "TVDdown": 6,
             "Sweby": 7}
_SCHEMES = {getInterfaceNormal_WH: 0, getInterfaceNormal_Column: 1, getInterfaceNormal_WY: 2,
            getInterfaceNormal_MYC: 3, getInterfaceNormal_PCD: 4, getInterfaceNormal_Y: 5, getInterfaceNormal_SLIC: 6,
            "WH": 0, "Column": 1, "WY": 2, "MYC": 3, "PCD": 4, "Youngs": 5, "SLIC": 6}


def sdf_samples(dims, sdf, T):
    """Per interior cell (x fastest): sdf(x), then sdf(x ± Δx e_i) (VOFutil.jl:203-225,
    x = loc(0,I) = I .- 1.5, Δx = T(0.01)); the SDF is called on arrays of dtype T."""
    T = np.dtype(T).type
    D = len(dims)
    axes = [np.arange(n, dtype=T) + T(0.5) for n in dims]
    X = [x.ravel(order="F") for x in np.meshgrid(*axes, indexing="ij")]
    dx = T(0.01)
    cols = [sdf(X)]
    for i in range(D):
        for sgn in (1, -1):
            cols.append(sdf([(x + (sgn * dx if j == i else T(0))).astype(T) for j, x in enumerate(X)]))
    return np.ascontiguousarray(np.stack([np.broadcast_to(np.asarray(c, dtype=T), X[0].shape) for c in cols], axis=1))


def velocity_samples(dims, uλ, T):
    """apply!(uλ, u): every index incl. ghosts, loc(i,I) = I .- 1.5 .- 0.5δ(i)."""
    T = np.dtype(T).type
    D = len(dims)
    Ng = [n + 2 for n in dims]
    out = np.zeros(tuple(Ng) + (D,), dtype=T, order="F")
    base = [np.arange(n, dtype=T) - T(0.5) for n in Ng]
    for i in range(D):
        X = np.meshgrid(*[(b - (T(0.5) if j == i else T(0))).astype(T) for j, b in enumerate(base)], indexing="ij")
        out[..., i] = np.asarray(uλ(i + 1, X), dtype=T)
    return out


class _Fields:
    """Attribute access to a context's fields (Flow / cVOF field names)."""

    def __init__(self, sim, names):
        object.__setattr__(self, "_sim", sim)
        object.__setattr__(self, "_names", names)

    def __getattr__(self, k):
        if k in self._names:
            return self._sim._get(self._names[k])
        raise AttributeError(k)

    def __setattr__(self, k, v):
        if k in self._names:
            self._sim._set(self._names[k], v)
        else:
            object.__setattr__(self, k, v)


class TwoPhaseSimulation:
    """TwoPhaseSimulation(dims, u_BC, L; T, mem, λμ, λρ, η, InterfaceSDF, λ, normalScheme,
    Δt, ν, g, U, ϵ, perdir, exitBC, body, u0)  (InterfaceAdvection.jl:64-85).

    Extra keyword (this build): `poisson` = "reference" (WaterLily solver!, tol Σr²,
    flow.jl:346), "refined" (FP32 V-cycles in FP64 iterative refinement to
    `poisson_rtol`), or "learned".  `g` is a constant vector or a callable
    g(i, x, t) that is constant in x and t; `udf` hooks are not supported.
    """

    def __init__(self, dims, u_BC, L=1.0, *, T=np.float32, mem="hip", λμ=1e-2, λρ=1e-3, η=None,
                 InterfaceSDF=None, λ=Koren, normalScheme=getInterfaceNormal_WH, Δt=0.25, ν=0.0, g=None,
                 U=None, ϵ=1, perdir=(), exitBC=False, body=None, u0=None, uλ=None, poisson="reference",
                 poisson_tol=1e-4, poisson_itmx=200, poisson_rtol=1e-6, poisson_max_outer=100, device=0):
        self.T = np.dtype(T)
        if self.T not in (np.float32, np.float64):
            raise ValueError("T must be Float32 or Float64")
        if body is not None:
            raise ValueError("immersed bodies are not supported by MPFMomStep! (InterfaceAdvection.jl:97)")
        if η is not None and not (float(η) > 0):
            raise ValueError("surface tension coefficient η must be positive (or None)")
        self.dims = tuple(int(n) for n in dims)
        self.D = len(self.dims)
        self.Ng = tuple(n + 2 for n in self.dims)
        self.L = L
        self.U = U if U is not None else float(np.sqrt(sum(float(x) ** 2 for x in u_BC)))
        self.ϵ = ϵ
        self.mem = mem
        gv = None
        if g is not None:
            gv = [float(g(i + 1, np.zeros(self.D), 0.0)) if callable(g) else float(g[i]) for i in range(self.D)]
        p = IAParams()
        p.D = self.D
        for d in range(3):
            p.N[d] = self.dims[d] if d < self.D else 1
            p.u_bc[d] = float(u_BC[d]) if d < self.D else 0.0
            p.g[d] = gv[d] if (gv is not None and d < self.D) else 0.0
        p.dtype = IA_F32 if self.T == np.float32 else IA_F64
        p.perdir = sum(1 << (d - 1) for d in perdir)
        p.has_g = 0 if gv is None else 1
        p.nu = float(ν)
        p.lambda_rho = float(λρ)
        p.lambda_mu = float(λμ)
        p.eta = 0.0 if η is None else float(η)
        p.normal_scheme = _SCHEMES[normalScheme]
        p.limiter = _LIMITERS[λ]
        p.poisson_mode = POISSON[poisson]
        p.poisson_tol = float(poisson_tol)
        p.poisson_itmx = int(poisson_itmx)
        p.poisson_rtol = float(poisson_rtol)
        p.poisson_max_outer = int(poisson_max_outer)
        p.device = int(device)
        p.exit_bc = 1 if exitBC else 0
        if exitBC and gv is not None and gv[0] != 0:
            raise ValueError("exitBC with gravity along x is not supported (the exit face's uStar would be g_x)")
        self.params = p
        self._lib = lib()
        ctx = C.c_void_p()
        rc = self._lib.ia_create(C.byref(p), C.byref(ctx))
        if rc != 0:
            raise IAError(rc, "ia_create failed (a GPU is required; grid sides must be a*2^n, n>2)")
        self._ctx = ctx
        self.perdir = tuple(perdir)
        self.flow = _Fields(self, {"u": "u", "u0": "u0", "p": "p", "σ": "sigma", "sigma": "sigma", "μ₀": "mu0",
                                   "mu0": "mu0", "f": "force", "force": "force"})
        self.intf = _Fields(self, {"f": "f", "f0": "f0", "α": "alpha", "alpha": "alpha", "n̂": "nhat", "nhat": "nhat",
                                   "c̄": "cbar", "cbar": "cbar", "ρu": "rhou", "rhou": "rhou", "ρuf": "rhouf",
                                   "rhouf": "rhouf"})
        self.flow.Δt = [self.T.type(Δt)]
        self.flow.ν = ν
        self.flow.g = g
        self.flow.perdir = self.perdir
        self.intf.λρ, self.intf.λμ = λρ, λμ
        self.intf.μ = None if ν == 0 else ν
        self.intf.η = None
        self.intf.perdir = self.perdir
        self.pois = self
        self.body = None
        uλ = u0 if u0 is not None else uλ
        if uλ is not None:
            us = velocity_samples(self.dims, uλ, self.T)
            self._call("ia_set_initial_velocity", _ptr(us), us.nbytes)
        if InterfaceSDF is not None:
            s = sdf_samples(self.dims, InterfaceSDF, self.T)
            self._call("ia_apply_vof", _ptr(s), s.nbytes)
        # flow.Δt[end] = min(last(flow.Δt), MPCFL(flow,intf))  (InterfaceAdvection.jl:81)
        self.flow.Δt[-1] = min(self.flow.Δt[-1], self.T.type(self.MPCFL()))
        self.last_stats = None

    # ------------------------------------------------------------------ plumbing
    def _call(self, name, *args):
        rc = getattr(self._lib, name)(self._ctx, *args)
        if rc in (1, 2, 4, 5):
            raise IAError(rc, self._lib.ia_last_error(self._ctx).decode())
        if rc == 3:
            print("ERROR: " + self._lib.ia_last_error(self._ctx).decode())
        return rc

    def _shape(self, name):
        return self.Ng + ((self.D,) if name in _VECTOR else ())

    def _get(self, name):
        dt = np.int8 if name == "cbar" else self.T
        a = np.zeros(self._shape(name), dtype=dt, order="F")
        self._call("ia_get_field", FIELD[name], _ptr(a), a.nbytes)
        return a

    def _set(self, name, v):
        dt = np.int8 if name == "cbar" else self.T
        a = np.asfortranarray(v, dtype=dt)
        if a.shape != self._shape(name):
            raise ValueError(f"{name}: shape {a.shape}, expected {self._shape(name)}")
        self._call("ia_set_field", FIELD[name], _ptr(a), a.nbytes)

    def device_ptr(self, name):
        p, n = C.c_void_p(), C.c_size_t()
        self._call("ia_field_device_ptr", FIELD[name], C.byref(p), C.byref(n))
        return p.value, n.value

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.ia_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ reference API
    def MPCFL(self):
        """MPCFL(flow, intf) (flow.jl:262-281)."""
        v = C.c_double()
        self._call("ia_cfl", C.byref(v))
        return v.value

    def sim_time(self):
        """WaterLily sim_time: the dimensionless time Σ Δt[1:end-1] · U/L."""
        return float(sum(float(x) for x in self.flow.Δt[:-1])) * self.U / self.L

    def metrics(self, U=None, StatWL=None):
        """Domain integrals of the diagnostics of metrics.jl:15-50 (FP64): kinetic
        energy Σρke, momentum Σρu, potential energy Σρgh, enstrophy and volume Σf."""
        u = np.zeros(3)
        w = np.zeros(3)
        if U is not None:
            u[:len(U)] = U
        if StatWL is not None:
            w[:len(StatWL)] = StatWL
        out = np.zeros(7)
        self._call("ia_metrics", _ptr(u), _ptr(w), _ptr(out))
        return dict(kinetic_energy=out[0], momentum=tuple(out[1:1 + self.D]), potential_energy=out[4],
                    enstrophy=out[5], volume=out[6])

    # ---- learned Poisson correction (no reference counterpart)
    def learned_init(self, seed=0):
        """He-normal initialisation of the correction network."""
        self._call("ia_learned_init", int(seed))

    def learned_train(self, on=True, lr=1e-3):
        """One Adam step per LEARNED projection while on."""
        self._call("ia_learned_train", int(bool(on)), float(lr))

    @property
    def learned_weights(self):
        w = np.empty(LEARNED_NPAR, dtype=np.float32)
        self._call("ia_learned_get_weights", _ptr(w), w.size)
        return w

    @learned_weights.setter
    def learned_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._call("ia_learned_set_weights", _ptr(w), w.size)

    def learned_stats(self):
        s = IALearnedInfo()
        self._call("ia_learned_stats", C.byref(s))
        return dict(projections=s.projections, accepted=s.accepted, train_steps=s.train_steps,
                    last_ratio=s.last_ratio)

    def learned_forward(self, feat):
        """The network on features of shape (3, N0, N1, N2) -> (2, N0, N1, N2)."""
        N = tuple(n - 2 for n in self.Ng)
        src = np.stack([np.asarray(feat[c], dtype=np.float32).reshape(-1, order="F") for c in range(3)])
        out = np.empty((2, int(np.prod(N))), dtype=np.float32)
        self._call("ia_learned_forward", _ptr(np.ascontiguousarray(src)), _ptr(out))
        return np.stack([out[c].reshape(N, order="F") for c in range(2)])

    def learned_grad(self, feat, dout):
        """Parameter gradient of Σ dout·net(feat) (LEARNED_NPAR values)."""
        src = np.ascontiguousarray(np.stack([np.asarray(feat[c], dtype=np.float32).reshape(-1, order="F")
                                             for c in range(3)]))
        d = np.ascontiguousarray(np.stack([np.asarray(dout[c], dtype=np.float32).reshape(-1, order="F")
                                           for c in range(2)]))
        gw = np.empty(LEARNED_NPAR, dtype=np.float32)
        self._call("ia_learned_grad", _ptr(src), _ptr(d), _ptr(gw))
        return gw

    def mg_levels(self):
        n = self._lib.ia_mg_levels(self._ctx, 0, None)
        out = []
        for l in range(n):
            Ng = (C.c_int * 3)()
            self._lib.ia_mg_levels(self._ctx, l, Ng)
            out.append(tuple(Ng[:self.D]))
        return out


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def MPFMomStep_(flow, pois, intf, body=None, *, δt=None, udf=None):
    """MPFMomStep!(flow, pois, intf, body; δt = last(flow.Δt))  (flow.jl:60-109)."""
    sim = flow._sim if isinstance(flow, _Fields) else flow
    if udf is not None:
        raise ValueError("udf hooks cannot cross the C ABI")
    dt = sim.flow.Δt[-1] if δt is None else δt
    st = IAStepStats()
    sim._call("ia_mom_step", float(dt), len(sim.flow.Δt), C.byref(st))
    sim.flow.Δt.append(sim.T.type(st.dt_next))
    sim.last_stats = st
    return st


def sim_step_(sim, t_end=None, *, remeasure=False, udf=None, verbose=False, max_steps=None):
    """sim_step!(sim; remeasure, udf) (InterfaceAdvection.jl:100-103); with t_end,
    WaterLily's generic loop `sim_step!(sim, t_end)` over that method: steps while
    the dimensionless time sim_time(sim) is below t_end (and fewer than max_steps
    were taken)."""
    if t_end is None:
        return MPFMomStep_(sim, sim, sim, None, udf=udf)
    if not sim.U > 0:
        raise ValueError("sim_step!(sim, t_end) needs a velocity scale U > 0 (sim_time is t·U/L)")
    n0 = len(sim.flow.Δt)
    while sim.sim_time() < t_end and (max_steps is None or len(sim.flow.Δt) - n0 < max_steps):
        MPFMomStep_(sim, sim, sim, None, udf=udf)
        if verbose:
            sim_info(sim)


def sim_info(sim):
    """sim_info (InterfaceAdvection.jl:104-107)."""
    print(f"    tU/L={sim.sim_time():10.6f}, ΔtU/L={sim.flow.Δt[-1] * sim.U / sim.L:.10f}")


_ADVECT_F = {"f": FIELD["f"], "f0": FIELD["f0"]}
_ADVECT_U = {"u": FIELD["u"], "u0": FIELD["u0"], "force": FIELD["force"], "rhou": FIELD["rhou"]}


def advect_(flow, intf=None, f="f", u1="u0", u2="u", dt=None):
    """advect!(a, c, f=c.f, u¹=a.u⁰, u²=a.u, dt=last(a.Δt))  (advection.jl:17-23).

    f names the advected field ("f" = intf.f or "f0" = intf.f⁰, advected in
    place); u1 and u2 name the velocities ("u", "u0", "force", "rhou").  An
    array instead of a name is first uploaded into the default field of that
    argument (intf.f, flow.u⁰, flow.u)."""
    sim = flow._sim if isinstance(flow, _Fields) else flow

    def pick(v, table, default, setter):
        if isinstance(v, str):
            if v not in table:
                raise ValueError(f"advect!: {v!r} is not one of {sorted(table)}")
            return table[v]
        setter(v)
        return table[default]

    fid = pick(f, _ADVECT_F, "f", lambda a: setattr(sim.intf, "f", a))
    u1id = pick(u1, _ADVECT_U, "u0", lambda a: setattr(sim.flow, "u0", a))
    u2id = pick(u2, _ADVECT_U, "u", lambda a: setattr(sim.flow, "u", a))
    dt = sim.flow.Δt[-1] if dt is None else dt
    st = IAStepStats()
    sim._call("ia_advect", fid, u1id, u2id, float(dt), len(sim.flow.Δt), C.byref(st))
    sim.last_stats = st
    return st


class LevelSet:
    """LevelSet(sim) (redistaning.jl:8-26): the signed level set ϕ = 2f - 1, kept
    in the storage of sim.intf.f⁰ (so `ls.ϕ` is `sim.intf.f0`), with ϕini = ϕ in a
    scratch buffer of the simulation and ϕ⁰ (`ls.ϕ0`) / L in sim.intf.α / sim.flow.σ.  It
    stays valid until f is next advected; build a new one after a step."""

    def __init__(self, sim):
        sim._call("ia_levelset")
        self.sim = sim
        self.perdir = sim.perdir

    @property
    def ϕ(self):
        return self.sim.intf.f0

    @ϕ.setter
    def ϕ(self, a):
        self.sim.intf.f0 = a

    phi = ϕ

    @property
    def ϕ0(self):
        return self.sim.intf.α

    phi0 = ϕ0

    @property
    def L(self):
        return self.sim.flow.σ


def redistaning_(ls, d=5, dτ=0.5, perdir=()):
    """redistaning!(ls; d=5, dτ=0.5, perdir=()) (redistaning.jl:44-56): march
    ∂τϕ = ϕini (1 - |∇ϕ|) with SSP-RK3 steps of dτ for a pseudo-time d.  As in the
    reference, perdir is its own argument (default non-periodic), not ls.perdir."""
    mask = 0
    for a in perdir:
        mask |= 1 << (int(a) - 1)
    ls.sim._call("ia_redistance", float(d), float(dτ), mask)
    return ls


def inside(a, D=None):
    """WaterLily inside(a): drop one ghost layer per spatial axis."""
    nd = a.ndim if D is None else D
    return a[tuple(slice(1, -1) for _ in range(nd))]
