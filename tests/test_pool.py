"""The pool workload generator (interfaceadvection.jl_amd/pool.py): bubble
placement constraints, the committed bubble lists (tests/golden/bubbles_*.json,
written by tests/golden/make_bubble_lists.py), and the truncated SDF sampling,
whose applyVOF! fractions must equal those of the full SDF
max(z - zmax L, max_b(r_b - |x - c_b|))."""
import importlib
import json
import os

import numpy as np
import pytest

from __graft_entry__ import load_package
from oracle import oracle as O

ia = load_package()
pool = importlib.import_module(ia.__name__ + ".pool")


@pytest.mark.parametrize("name", sorted(pool.CASES))
def test_bubbles_respect_clearances(name):
    c = pool.CASES[name]
    N = c["N"]
    bl = pool.bubbles(N, c["nb"], c["seed"], c["rmin"], c["rmax"], c["zmax"])
    assert len(bl) == c["nb"]
    a = np.array(bl)
    r = a[:, 3]
    assert (r >= c["rmin"] * N).all() and (r <= c["rmax"] * N).all()
    assert (a[:, :3] - r[:, None] >= 2).all() and (a[:, :2] + r[:, None] <= N - 2).all()
    assert (a[:, 2] + r <= c["zmax"] * N - 2).all()
    d = np.linalg.norm(a[:, None, :3] - a[None, :, :3], axis=-1)
    np.fill_diagonal(d, np.inf)
    assert (d >= r[:, None] + r[None, :] + 2).all()
    assert pool.bubbles(N, c["nb"], c["seed"], c["rmin"], c["rmax"], c["zmax"]) == bl


@pytest.mark.parametrize("name", sorted(pool.CASES))
def test_bubbles_match_the_committed_lists(name):
    """The workload is its bubble list: the generator (seeded PCG64) must
    reproduce the committed one exactly, so that benchmark and parity runs on
    any numpy build advance the same state."""
    c = pool.CASES[name]
    with open(os.path.join(os.path.dirname(__file__), "golden", f"bubbles_{name}.json")) as fh:
        rec = json.load(fh)
    assert (rec["N"], rec["seed"]) == (c["N"], c["seed"])
    bl = pool.bubbles(c["N"], c["nb"], c["seed"], c["rmin"], c["rmax"], c["zmax"])
    assert [list(map(float, b)) for b in bl] == rec["bubbles"]


def full_sdf_samples(N, bl, zmax, T=np.float32):
    dx = T(0.01)
    ax = np.arange(N, dtype=T) + T(0.5)
    X = np.meshgrid(ax, ax, ax, indexing="ij")
    offs = [(0, 0, 0), (dx, 0, 0), (-dx, 0, 0), (0, dx, 0), (0, -dx, 0), (0, 0, dx), (0, 0, -dx)]
    S = np.empty((N, N, N, 7), dtype=T, order="F")
    for q, o in enumerate(offs):
        Y = [(X[a] + T(o[a])).astype(T) for a in range(3)]
        s = (Y[2] - T(zmax * N)).astype(T)
        for (cx, cy, cz, r) in bl:
            d = np.sqrt((Y[0] - T(cx)) ** 2 + (Y[1] - T(cy)) ** 2 + (Y[2] - T(cz)) ** 2)
            s = np.maximum(s, (T(r) - d).astype(T))
        S[..., q] = s
    return np.ascontiguousarray(S.reshape(N ** 3, 7, order="F"))


def test_truncated_sdf_gives_the_same_fractions():
    N, zmax = 40, 0.7
    bl = pool.bubbles(N, 6, 11, 0.05, 0.12, zmax)
    fr = []
    for S in (pool.pool_sdf_samples(N, bl, zmax), full_sdf_samples(N, bl, zmax)):
        o = O.OracleSim((N, N, N), (0.0, 0.0, 0.0), float(N), T=np.float32)
        o.lib.oc_apply_vof(o.ctx, O._ptr(S))
        fr.append(o.f)
    assert 0 < O.inside(fr[0]).mean() < 1
    assert (fr[0] == fr[1]).all()
