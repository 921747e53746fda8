"""ctypes wrapper around the CPU oracle (oracle/ia_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

`OracleSim` mirrors the reference's `TwoPhaseSimulation` construction
(InterfaceAdvection.jl:67-85) so that a test can build the same case on the
oracle and on the HIP library and compare fields.  Arrays are returned as
Fortran-ordered numpy arrays with the reference's shapes: (Ng...,) for
scalars and (Ng..., D) for vectors.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

F, F0, ALPHA, NHAT, FFLUX, CBAR, RHOU, RHOUF, DRHO, U, U0, FORCE, P, SIGMA, MU0, MG_D, MG_ID, MG_R, MG_EPS = range(19)
_VECTOR = {NHAT, RHOU, RHOUF, DRHO, U, U0, FORCE, MU0}
NORMAL = {"WH": 0, "Column": 1, "WY": 2, "MYC": 3, "PCD": 4, "Youngs": 5, "SLIC": 6}
LIMITER = {"Koren": 0, "minmod": 1, "upwind": 2, "vanAlbada1": 3, "superbee": 4, "TVDcen": 5, "TVDdown": 6,
           "Sweby": 7}


class OcParams(C.Structure):
    _fields_ = [("D", C.c_int), ("N", C.c_int * 3), ("perdir", C.c_int),
                ("u_bc", C.c_double * 3), ("g", C.c_double * 3), ("has_g", C.c_int), ("eta", C.c_double),
                ("nu", C.c_double), ("lambda_rho", C.c_double), ("lambda_mu", C.c_double),
                ("normal_scheme", C.c_int), ("limiter", C.c_int), ("dt0", C.c_double),
                ("mg_itmx", C.c_int), ("mg_tol", C.c_double), ("poisson_mode", C.c_int),
                ("poisson_rtol", C.c_double), ("smooth_it", C.c_int), ("exit_bc", C.c_int)]


_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(T):
    T = np.dtype(T)
    key = "f32" if T == np.float32 else "f64"
    if key in _libs:
        return _libs[key]
    path = os.path.join(_HERE, f"libia_oracle_{key}.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    R = C.c_float if key == "f32" else C.c_double
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
    sig = {
        "oc_create": (vp, [C.POINTER(OcParams)]),
        "oc_destroy": (None, [vp]),
        "oc_field_len": (C.c_int, [vp, C.c_int]),
        "oc_set_field": (C.c_int, [vp, C.c_int, vp]),
        "oc_get_field": (C.c_int, [vp, C.c_int, vp]),
        "oc_mg_nlevels": (C.c_int, [vp]),
        "oc_mg_level_dims": (C.c_int, [vp, C.c_int, ip]),
        "oc_mg_level_get": (C.c_int, [vp, C.c_int, C.c_int, vp]),
        "oc_mg_transfer": (C.c_int, [vp, C.c_int, C.c_int, vp]),
        "oc_mg_level_set": (C.c_int, [vp, C.c_int, C.c_int, vp]),
        "oc_mg_pcg": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "oc_apply_vof": (C.c_int, [vp, vp]),
        "oc_set_initial_velocity": (C.c_int, [vp, vp]),
        "oc_cfl": (C.c_double, [vp]),
        "oc_mom_step": (C.c_int, [vp, C.c_double, C.c_int, dp]),
        "oc_advect": (C.c_int, [vp, C.c_double, C.c_int]),
        "oc_reconstruct": (C.c_int, [vp]),
        "oc_vof_sweep": (C.c_int, [vp, C.c_int, C.c_double]),
        "oc_visc": (C.c_int, [vp, C.c_int]),
        "oc_update_L": (C.c_int, [vp, C.c_int]),
        "oc_mg_update": (C.c_int, [vp]),
        "oc_project": (C.c_int, [vp, C.c_double, ip, dp]),
        "oc_mg_solve": (C.c_int, [vp, C.c_double, C.c_int, ip, dp]),
        "oc_project_refined": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, ip, dp]),
        "oc_project_fcg": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int, ip, dp]),
        "oc_set_fine_cheb": (None, [vp, C.c_int]),
        "oc_getIntercept2": (R, [R, R, R]),
        "oc_getIntercept3": (R, [R, R, R, R]),
        "oc_getVolumeFraction2": (R, [R, R, R]),
        "oc_getVolumeFraction3": (R, [R, R, R, R]),
        "oc_f2alpha3": (R, [R, R, R, R]),
        "oc_alpha2f3": (R, [R, R, R, R]),
        "oc_koren": (R, [R, R, R]),
        "oc_limiter": (R, [C.c_int, R, R, R]),
        "oc_phiq_value": (R, [R, R, R, R, C.c_int, R, R, R]),
        "oc_popinet_height": (R, [C.c_int, ip, vp, ip, C.c_int]),
        "oc_rhoke": (C.c_double, [C.c_int, ip, vp, vp, R, ip, vp]),
        "oc_rhou": (C.c_double, [C.c_int, ip, C.c_int, vp, vp, R, ip, vp]),
        "oc_rhogh": (C.c_double, [C.c_int, ip, vp, vp, R, ip, vp]),
        "oc_ens": (C.c_double, [C.c_int, ip, vp, ip]),
        "oc_metrics": (C.c_int, [vp, vp, vp, vp]),
        "oc_curvature": (R, [C.c_int, ip, vp, ip, C.c_int]),
        "oc_normal_block": (C.c_int, [C.c_int, ip, vp, C.c_int, ip, vp]),
        "oc_getmu": (R, [C.c_int, ip, vp, C.c_int, C.c_int, ip, R, R, R]),
        "oc_bc_scalar": (C.c_int, [C.c_int, ip, vp, C.c_int, C.c_int]),
        "oc_bc_facevec": (C.c_int, [C.c_int, ip, vp, C.c_int]),
        "oc_bc_vel": (C.c_int, [C.c_int, ip, vp, vp, C.c_int, C.c_int]),
        "oc_bc_vof": (C.c_int, [C.c_int, ip, vp, vp, vp, C.c_int]),
        "oc_f2face": (C.c_int, [C.c_int, ip, vp, vp, C.c_int]),
        "oc_vof_flux_face": (C.c_int, [C.c_int, ip, vp, vp, vp, vp, R, C.c_int, ip, vp, R]),
        "oc_ls_grad2": (R, [R, R, R, R, R, R]),
        "oc_ls_computeL": (C.c_int, [C.c_int, ip, vp, vp, vp, C.c_int]),
        "oc_ls_stage": (C.c_int, [C.c_int, ip, vp, vp, vp, vp, R, R, C.c_int]),
        "oc_redistance": (C.c_int, [C.c_int, ip, vp, vp, vp, vp, C.c_double, C.c_double, C.c_int]),
        "oc_set_threads": (None, [C.c_int]),
        "oc_last_iters": (C.c_int, [vp, C.c_int]),
        "oc_last_error": (C.c_char_p, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _libs[key] = L
    return L


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _ints(v):
    return (C.c_int * len(v))(*v)


def perdir_mask(perdir):
    m = 0
    for d in perdir:
        m |= 1 << (d - 1)
    return m


def _ls_arrays(*arrs):
    T = arrs[0].dtype
    for a in arrs:
        if a.dtype != T or not a.flags.f_contiguous:
            raise ValueError("level-set arrays must share one float dtype and be Fortran-ordered")
    return lib(T), T


def ls_grad2(a, b, c, d, e, s, T=np.float64):
    """𝛁ϕᵢ²(a,b,c,d,e,s) (redistaning.jl:122-143)."""
    return lib(T).oc_ls_grad2(a, b, c, d, e, s)


def computeL(L, phi, phiini, perdir=()):
    """computeL!(L, ϕ, ϕini; perdir) (redistaning.jl:67-85), in place on L."""
    lb, _ = _ls_arrays(L, phi, phiini)
    st = lb.oc_ls_computeL(phi.ndim, _ints(phi.shape), _ptr(L), _ptr(phi), _ptr(phiini), perdir_mask(perdir))
    if st:
        raise ValueError("computeL!: every axis needs at least 2 interior cells")


def redistaning_stage(phi, phi0, phiini, L, dτ, α, perdir=()):
    """_redistaningStage!(ϕ, ϕ⁰, ϕini, L, dτ, α; perdir) (redistaning.jl:28-31), in place."""
    lb, _ = _ls_arrays(phi, phi0, phiini, L)
    st = lb.oc_ls_stage(phi.ndim, _ints(phi.shape), _ptr(phi), _ptr(phi0), _ptr(phiini), _ptr(L), dτ, α,
                        perdir_mask(perdir))
    if st:
        raise ValueError("_redistaningStage!: every axis needs at least 2 interior cells")


def redistaning(phi, phi0, phiini, L, d=5, dτ=0.5, perdir=()):
    """redistaning!(ls; d, dτ, perdir) (redistaning.jl:44-56) on the arrays of a
    LevelSet (ϕ updated in place; ϕ⁰ and L are scratch)."""
    lb, _ = _ls_arrays(phi, phi0, phiini, L)
    st = lb.oc_redistance(phi.ndim, _ints(phi.shape), _ptr(phi), _ptr(phi0), _ptr(phiini), _ptr(L), float(d),
                          float(dτ), perdir_mask(perdir))
    if st:
        raise ValueError("redistaning!: every axis needs at least 2 interior cells")


def sdf_samples(dims, sdf, T):
    """SDF samples per interior cell, x fastest: centre, then ±Δx per axis
    (VOFutil.jl:203-225 with loc(0,I) = I .- 1.5, Δx = T(0.01))."""
    T = np.dtype(T).type
    D = len(dims)
    axes = [np.arange(n, dtype=T) + T(0.5) for n in dims]
    X = np.meshgrid(*axes, indexing="ij")
    X = [x.ravel(order="F") for x in X]
    dx = T(0.01)
    cols = [sdf(X)]
    for i in range(D):
        for sgn in (1, -1):
            Y = [x + (sgn * dx if j == i else T(0)) for j, x in enumerate(X)]
            Y = [y.astype(T) for y in Y]
            cols.append(sdf(Y))
    return np.ascontiguousarray(np.stack([np.asarray(c, dtype=T) for c in cols], axis=1))


def velocity_samples(dims, u0, T):
    """apply!(uλ,u): u[Ii] = uλ(i, loc(Ii)) over every index incl. ghosts,
    loc(i,I) = I .- 1.5 .- 0.5δ(i) (1-based), in Fortran order (Ng..., D)."""
    T = np.dtype(T).type
    D = len(dims)
    Ng = [n + 2 for n in dims]
    out = np.zeros(tuple(Ng) + (D,), dtype=T, order="F")
    base = [np.arange(n, dtype=T) - T(0.5) for n in Ng]  # I-1.5 with I = q+1
    for i in range(D):
        axes = [b - (T(0.5) if j == i else T(0)) for j, b in enumerate(base)]
        X = np.meshgrid(*[a.astype(T) for a in axes], indexing="ij")
        out[..., i] = np.asarray(u0(i + 1, X), dtype=T)
    return out


class OracleSim:
    """Oracle counterpart of TwoPhaseSimulation (InterfaceAdvection.jl:67-85)."""

    def __init__(self, dims, u_BC, L=1.0, *, T=np.float32, λμ=1e-2, λρ=1e-3, ν=0.0, g=None,
                 perdir=(), Δt=0.25, InterfaceSDF=None, u0=None, normal="WH", limiter="Koren",
                 mg_tol=1e-4, mg_itmx=200, U=None, poisson="reference", poisson_rtol=1e-6, smooth_it=2, η=None,
                 exitBC=False):
        self.T = np.dtype(T)
        self.lib = lib(self.T)
        self.dims = tuple(int(n) for n in dims)
        self.D = len(dims)
        self.Ng = tuple(n + 2 for n in self.dims)
        self.L = L
        self.U = U if U is not None else float(np.sqrt(sum(x * x for x in u_BC)))
        p = OcParams()
        p.D = self.D
        for d in range(3):
            p.N[d] = self.dims[d] if d < self.D else 1
            p.u_bc[d] = float(u_BC[d]) if d < self.D else 0.0
            p.g[d] = float(g[d]) if (g is not None and d < self.D) else 0.0
        p.has_g = 0 if g is None else 1
        p.perdir = perdir_mask(perdir)
        p.nu = float(ν)
        p.lambda_rho = float(λρ)
        p.lambda_mu = float(λμ)
        p.eta = 0.0 if η is None else float(η)
        p.normal_scheme = NORMAL[normal]
        p.limiter = LIMITER[limiter]
        p.dt0 = float(Δt)
        p.mg_itmx = int(mg_itmx)
        p.mg_tol = float(mg_tol)
        p.poisson_mode = {"reference": 0, "refined": 1, "pcg": 2}[poisson]
        p.poisson_rtol = float(poisson_rtol)
        p.smooth_it = int(smooth_it)
        p.exit_bc = 1 if exitBC else 0
        self.ctx = self.lib.oc_create(C.byref(p))
        if not self.ctx:
            raise ValueError("oracle: bad parameters (MultiLevelPoisson needs >2 levels)")
        if u0 is not None:
            us = velocity_samples(self.dims, u0, self.T)
            self.lib.oc_set_initial_velocity(self.ctx, _ptr(us))
        if InterfaceSDF is not None:
            s = sdf_samples(self.dims, InterfaceSDF, self.T)
            self.lib.oc_apply_vof(self.ctx, _ptr(s))
        self.Δt = [self.T.type(Δt)]
        self.Δt[-1] = min(self.Δt[-1], self.T.type(self.lib.oc_cfl(self.ctx)))

    def __del__(self):
        if getattr(self, "ctx", None):
            self.lib.oc_destroy(self.ctx)
            self.ctx = None

    def shape(self, fid):
        return self.Ng + ((self.D,) if fid in _VECTOR else ())

    def get(self, fid):
        if fid == CBAR:
            a = np.zeros(self.Ng, dtype=np.int8, order="F")
        else:
            a = np.zeros(self.shape(fid), dtype=self.T, order="F")
        self.lib.oc_get_field(self.ctx, fid, _ptr(a))
        return a

    def set(self, fid, a):
        dt = np.int8 if fid == CBAR else self.T
        a = np.asfortranarray(a, dtype=dt)
        assert a.shape == (self.Ng if fid == CBAR else self.shape(fid))
        self.lib.oc_set_field(self.ctx, fid, _ptr(a))

    def sim_step(self):
        """sim_step!(sim) -> MPFMomStep! (flow.jl:60-109)."""
        nxt = C.c_double()
        st = self.lib.oc_mom_step(self.ctx, float(self.Δt[-1]), len(self.Δt), C.byref(nxt))
        if st in (1, 2):
            raise RuntimeError(self.lib.oc_last_error(self.ctx).decode())
        self.Δt.append(self.T.type(nxt.value))
        return st

    def advect(self, dt, step_index):
        """advect!(flow, intf, f, u⁰, u, dt) (advection.jl:17-23)."""
        return self.lib.oc_advect(self.ctx, float(dt), int(step_index))

    def project(self, dt_w):
        it, r2 = C.c_int(), C.c_double()
        self.lib.oc_project(self.ctx, float(dt_w), C.byref(it), C.byref(r2))
        return it.value, r2.value

    def project_refined(self, dt_w, rtol=1e-6, max_outer=100):
        it, rel = C.c_int(), C.c_double()
        self.lib.oc_project_refined(self.ctx, float(dt_w), float(rtol), int(max_outer), C.byref(it), C.byref(rel))
        return it.value, rel.value

    def project_fcg(self, dt_w, rtol=1e-6, max_it=200, smooth_it=2):
        """Flexible CG to rtol.  The fine smoother after each V-cycle is the two
        Chebyshev steps of mg_cheb2 unless set_fine_cheb(False) selects
        `smooth_it` pcg! iterations instead."""
        it, rel = C.c_int(), C.c_double()
        self.lib.oc_project_fcg(self.ctx, float(dt_w), float(rtol), int(max_it), int(smooth_it), C.byref(it),
                                C.byref(rel))
        return it.value, rel.value

    def set_fine_cheb(self, on=True):
        self.lib.oc_set_fine_cheb(self.ctx, 1 if on else 0)

    def metrics(self, U=None, StatWL=None):
        u = np.zeros(3)
        w = np.zeros(3)
        if U is not None:
            u[:len(U)] = U
        if StatWL is not None:
            w[:len(StatWL)] = StatWL
        out = np.zeros(7)
        self.lib.oc_metrics(self.ctx, _ptr(u), _ptr(w), _ptr(out))
        return out

    def cfl(self):
        return self.lib.oc_cfl(self.ctx)

    def sim_time(self):
        return float(sum(self.Δt[:-1]))

    @property
    def f(self):
        return self.get(F)

    @property
    def u(self):
        return self.get(U)

    @property
    def p(self):
        return self.get(P)


def inside(a, D=None):
    """Interior view (WaterLily inside): drop one ghost layer per spatial axis."""
    nd = a.ndim if D is None else D
    sl = tuple(slice(1, -1) for _ in range(nd))
    return a[sl]
