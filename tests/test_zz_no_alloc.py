"""ia_metrics_noalloc (the allocation-free variant of ia_metrics) through the
C ABI: its seven FP64 integrals equal ia_metrics's to round-off (the two sum
the same per-cell terms in different fixed orders: relative 1e-12, or 1e-12
of the largest magnitude for a total that cancels), on two calls in a row on
one GPU, the second on a smaller grid, so that the shared scratch array holds
the first call's partials when the second one runs."""
import ctypes as C

import numpy as np
import pytest

from __graft_entry__ import load_package
from test_gpu_parity import _swirl, bubbles3d, bubble2d

pytestmark = pytest.mark.gpu

ia = load_package()


def _metrics(sim, name, U, wl):
    out = np.zeros(7)
    u = np.zeros(3)
    w = np.zeros(3)
    u[:len(U)] = U
    w[:len(wl)] = wl
    rc = getattr(ia.lib(), name)(sim._ctx, ia._ptr(u), ia._ptr(w), ia._ptr(out))
    assert rc == 0, ia.lib().ia_last_error(sim._ctx)
    return out


def _close(a, b):
    scale = max(np.abs(a).max(), np.abs(b).max(), 1e-300)
    for x, y in zip(a, b):
        assert abs(x - y) <= 1e-12 * max(abs(y), 1e-3 * scale), (a, b)


def _case(case, T, amp):
    sim = ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], **case["kw"])
    sim.flow.u = _swirl(case, T, amp=amp)
    return sim


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_metrics_noalloc_matches_metrics(T):
    print("[case] ia_metrics_noalloc vs ia_metrics, 96^3 then 16^2,", T.__name__, flush=True)
    big = _case(bubbles3d(96, T), T, 0.8)       # 2 x 24 x 96 = 4608 tiles -> 1024 blocks of 4 or 5 tiles
    small = _case(bubble2d(16, T), T, 0.5)      # 1 x 4 tiles -> 4 blocks
    U, wl = (0.1, -0.2, 0.05), (1.0, 2.0, 3.0)
    a0 = _metrics(big, "ia_metrics", U, wl)
    a1 = _metrics(big, "ia_metrics_noalloc", U, wl)
    _close(a1, a0)
    b0 = _metrics(small, "ia_metrics", U[:2], wl[:2])
    b1 = _metrics(small, "ia_metrics_noalloc", U[:2], wl[:2])
    _close(b1, b0)
    # the first context again after the smaller call: nothing of the small
    # call's partials leaks into it
    a2 = _metrics(big, "ia_metrics_noalloc", U, wl)
    _close(a2, a0)
    assert np.array_equal(a2, a1)  # same grid, same order: bit-identical
    big.close()
    small.close()
