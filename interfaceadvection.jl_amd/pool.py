"""The 3D two-phase pool workloads: a liquid pool holding non-overlapping
spherical bubbles, slip walls, gravity -z, density ratio 1000, viscosity ratio
100, L = N cells, U = 1, g0 = U^2/L, ν = UL/1e4.

- "pool": 256^3, liquid below z = 0.7 L, 32 bubbles of radius U(0.03, 0.08) L,
  seed 3 (the benchmark case);
- "slab": 512^3, liquid below z = 0.6 L, 256 bubbles of radius U(0.02, 0.05) L,
  seed 5.

Centres are rejection-sampled uniformly in the liquid with at least 2 cells
of clearance to the walls, the free surface and each other (numpy PCG64 with
the given seed).  The volume fractions come from applyVOF! (VOFutil.jl:8-37)
on the SDF φ = max(z - zmax L, max_b(r_b - |x - c_b|)).
"""
CASES = {
    "pool": dict(N=256, nb=32, seed=3, rmin=0.03, rmax=0.08, zmax=0.7),
    "slab": dict(N=512, nb=256, seed=5, rmin=0.02, rmax=0.05, zmax=0.6),
}
import numpy as np


def bubbles(L, nb, seed, rmin, rmax, zmax, clearance=2.0):
    """Rejection-sampled non-overlapping bubbles inside the pool z < zmax."""
    rng = np.random.default_rng(seed)
    out = []
    tries = 0
    while len(out) < nb:
        tries += 1
        if tries > 200000:
            raise RuntimeError("could not place bubbles")
        r = rng.uniform(rmin, rmax) * L
        c = rng.uniform(0, L, 3)
        lo, hi = r + clearance, L - r - clearance
        if not (lo <= c[0] <= hi and lo <= c[1] <= hi and lo <= c[2] <= zmax * L - r - clearance):
            continue
        if any(np.linalg.norm(c - np.array(b[:3])) < r + b[3] + clearance for b in out):
            continue
        out.append((c[0], c[1], c[2], r))
    return out


def pool_sdf_samples(N, bl, zmax, T=np.float32):
    """applyVOF! samples (centre, ±Δx per axis) of the SDF
    φ = max(z - zmax·L, max_b(r_b - |x - c_b|)) for every interior cell.

    The SDF is evaluated exactly within 3 cells of every bubble; elsewhere the
    plane term alone is used.  Where they differ both are below -3, and the
    PLIC formula then gives f = 1 exactly either way, so f is identical to the
    full evaluation."""
    L = float(N)
    dx = T(0.01)
    D = 3
    z = (np.arange(N, dtype=T) + T(0.5))
    S = np.empty((N, N, N, 1 + 2 * D), dtype=T, order="F")
    zc = z[None, None, :] - T(zmax * L)
    S[..., 0] = zc
    S[..., 1] = zc
    S[..., 2] = zc
    S[..., 3] = zc
    S[..., 4] = zc
    S[..., 5] = (z + dx).astype(T)[None, None, :] - T(zmax * L)
    S[..., 6] = (z - dx).astype(T)[None, None, :] - T(zmax * L)
    for (cx, cy, cz, r) in bl:
        lo = [max(0, int(np.floor(c - r - 4))) for c in (cx, cy, cz)]
        hi = [min(N, int(np.ceil(c + r + 4))) for c in (cx, cy, cz)]
        ax = [np.arange(lo[a], hi[a], dtype=T) + T(0.5) for a in range(3)]
        X = np.meshgrid(*ax, indexing="ij")
        cc = (T(cx), T(cy), T(cz))
        offs = [(0, 0, 0), (dx, 0, 0), (-dx, 0, 0), (0, dx, 0), (0, -dx, 0), (0, 0, dx), (0, 0, -dx)]
        blk = S[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2], :]
        for q, o in enumerate(offs):
            Y = [(X[a] + T(o[a])).astype(T) for a in range(3)]
            d = np.sqrt((Y[0] - cc[0]) ** 2 + (Y[1] - cc[1]) ** 2 + (Y[2] - cc[2]) ** 2)
            blk[..., q] = np.maximum(blk[..., q], (T(r) - d).astype(T))
    return np.ascontiguousarray(S.reshape(N * N * N, 1 + 2 * D, order="F"))


def make_case(N, poisson, T=np.float32, nb=32, seed=3, rmin=0.03, rmax=0.08, zmax=0.7):
    L = float(N)
    bl = bubbles(N, nb, seed, rmin, rmax, zmax)
    kw = dict(T=T, λρ=1e-3, λμ=1e-2, ν=1.0 * L / 1e4, g=(0.0, 0.0, -1.0 / L), perdir=())
    return bl, kw


def build_sim(ia, N, poisson, rtol, seed=3, nb=32, rmin=0.03, rmax=0.08, zmax=0.7):
    """The pool workload at N^3 with the given Poisson mode, f initialised by
    applyVOF! from the SDF samples, Δt from MPCFL."""
    bl, kw = make_case(N, poisson, seed=seed, nb=nb, rmin=rmin, rmax=rmax, zmax=zmax)
    sim = ia.TwoPhaseSimulation((N, N, N), (0.0, 0.0, 0.0), float(N), poisson=poisson, poisson_rtol=rtol,
                                poisson_max_outer=200, **kw)
    s = pool_sdf_samples(N, bl, zmax)
    sim._call("ia_apply_vof", ia._ptr(s), s.nbytes)
    del s
    sim.flow.Δt[-1] = min(sim.flow.Δt[-1], sim.T.type(sim.MPCFL()))
    return sim, bl, kw
