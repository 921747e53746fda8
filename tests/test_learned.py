"""The learned Poisson correction (no reference counterpart): the BF16 MFMA
network against a float64 PyTorch CPU restatement of the same network, its
parameter gradient against PyTorch autograd, and the LEARNED solve mode.

Bars: the forward pass within 1e-2 (relative L2) of the reference, which
rounds weights, inputs and every layer's output to bfloat16 exactly where the
device stores them; parameter gradients within 2e-2 (relative L2) of autograd
through that same rounded network (autograd then also carries the layer
gradients in bfloat16, as the device does); every LEARNED solve
reaches the 1e-6 relative FP64 residual, and a correction is kept only when it
lowered the residual.
"""
import numpy as np
import pytest

from __graft_entry__ import load_package
from test_gpu_parity import bubbles3d, say

pytestmark = pytest.mark.gpu

ia = load_package()
torch = pytest.importorskip("torch")

NW = 27 * 16 * 16


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float64)


def torch_net(w, feat, bf16=True):
    """float64 restatement: 5 x (3x3x3 conv, zero padding), ReLU between layers.
    feat: (3, N0, N1, N2).  Weights W[l][tap][ci][co], tap = (dz+1)*9+(dy+1)*3+dx+1."""
    r = (lambda t: _bf(t)) if bf16 else (lambda t: t)
    x = torch.zeros((1, 16) + feat.shape[1:], dtype=torch.float64)
    x[0, :3] = torch.as_tensor(feat, dtype=torch.float64)
    x = r(x)
    for l in range(5):
        W = torch.as_tensor(w[l * NW:(l + 1) * NW].reshape(3, 3, 3, 16, 16), dtype=torch.float64)  # dz,dy,dx,ci,co
        K = r(W.permute(4, 3, 2, 1, 0).contiguous())  # co, ci, dx, dy, dz
        b = torch.as_tensor(w[5 * NW + 16 * l:5 * NW + 16 * (l + 1)], dtype=torch.float64)
        x = torch.nn.functional.conv3d(x, K, b, padding=1)
        if l < 4:
            x = torch.relu(x)
        x = r(x)
    return x[0, :2]


def make_sim(N=32):
    case = bubbles3d(N, np.float32)
    return ia.TwoPhaseSimulation(case["dims"], case["u_BC"], case["L"], poisson="learned", poisson_rtol=1e-6,
                                 poisson_max_outer=100, **case["kw"])


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_learned_forward_matches_torch():
    say("learned network forward vs torch, 32^3")
    g = make_sim(32)
    g.learned_init(seed=7)
    w = g.learned_weights
    rng = np.random.default_rng(1)
    feat = rng.standard_normal((3, 32, 32, 32)).astype(np.float32)
    # the last layer is initialised small; scale it up so the output is not tiny
    w2 = w.copy()
    w2[4 * NW:5 * NW] *= 10
    g.learned_weights = w2
    out = g.learned_forward(feat)
    ref = torch_net(w2, feat).numpy()
    assert out.shape == ref.shape == (2, 32, 32, 32)
    assert np.isfinite(out).all()
    assert rel_l2(out, ref) <= 1e-2
    g.close()


def test_learned_gradient_matches_autograd():
    say("learned network parameter gradient vs torch autograd, 32^3")
    g = make_sim(32)
    g.learned_init(seed=11)
    w = g.learned_weights
    w[4 * NW:5 * NW] *= 10
    g.learned_weights = w
    rng = np.random.default_rng(2)
    feat = rng.standard_normal((3, 32, 32, 32)).astype(np.float32)
    dout = rng.standard_normal((2, 32, 32, 32)).astype(np.float32)
    gw = g.learned_grad(feat, dout)
    wt = torch.tensor(w.astype(np.float64), requires_grad=True)
    out = torch_net(wt, feat, bf16=True)
    (out * torch.as_tensor(dout, dtype=torch.float64)).sum().backward()
    ref = wt.grad.numpy()
    for l in range(5):
        sl = slice(l * NW, (l + 1) * NW)
        assert rel_l2(gw[sl], ref[sl]) <= 2e-2, l
        bs = slice(5 * NW + 16 * l, 5 * NW + 16 * (l + 1))
        assert rel_l2(gw[bs], ref[bs]) <= 2e-2, l
    # unused parameters have no gradient: layer-0 input channels >= 3, layer-4 outputs >= 2
    W0 = gw[:NW].reshape(27, 16, 16)
    W4 = gw[4 * NW:5 * NW].reshape(27, 16, 16)
    assert (W0[:, 3:, :] == 0).all() and (W4[:, :, 2:] == 0).all()
    g.close()


def _project_fresh(g, rng, u0):
    g._call("ia_set_initial_velocity", ia._ptr(u0), u0.nbytes)
    g.flow.p = np.zeros(g.Ng, dtype=np.float32, order="F")
    g._call("ia_update_L", 0)
    g._call("ia_mg_update")
    st = ia.IASolveStats()
    g._call("ia_project", 1.0, ia.C.byref(st))
    return st


def test_learned_solves_reach_tolerance_and_training_lowers_the_loss():
    say("LEARNED projection, 32^3: 40 training solves on one right-hand side")
    g = make_sim(32)
    g.learned_init(seed=3)
    g.learned_train(True, lr=2e-3)
    rng = np.random.default_rng(5)
    u = np.zeros(g.Ng + (3,), dtype=np.float32, order="F")
    u[1:-1, 1:-1, 1:-1] = rng.standard_normal((32, 32, 32, 3))
    ratios = []
    for k in range(40):
        st = _project_fresh(g, rng, u)
        assert st.rel <= 1e-6
        s = g.learned_stats()
        ratios.append(s["last_ratio"])
    s = g.learned_stats()
    assert s["projections"] == 40 and s["train_steps"] == 40
    # the network starts as a damped Jacobi step and learns to undo more of the residual
    assert np.mean(ratios[-5:]) < np.mean(ratios[:5])
    assert np.mean(ratios[-5:]) < 1
    g.learned_train(False)
    st = _project_fresh(g, rng, u)
    assert st.rel <= 1e-6
    assert g.learned_stats()["train_steps"] == 40
    g.close()
