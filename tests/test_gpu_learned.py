"""GPU tests of the learned coarse correction: forward and gradient parity of
the hand-written HIP network against the CPU checker, training on pairs from
the GPU's FP64 multigrid, and the learned flexible CG against the plain one
(accuracy judged by the FP64 true residual; plain GMG-FCG as the fallback)."""
import sys

import numpy as np
import pytest

from gnu_femlisp_amd import flgpu as F
from gnu_femlisp_amd import iteration as it
from gnu_femlisp_amd import learned
from gnu_femlisp_amd.matlisp import GpuVector
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CIN = [2, 16, 16, 32, 32, 64, 96, 48, 16]
COUT = [16, 16, 32, 32, 64, 64, 32, 16, 1]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def he_theta(seed):
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in zip(CIN, COUT):
        parts.append(rng.normal(scale=np.sqrt(2.0 / (27 * ci)), size=27 * ci * co))
        parts.append(rng.normal(scale=0.1, size=co))
    return np.concatenate(parts)


def layer_slices():
    out, o = [], 0
    for ci, co in zip(CIN, COUT):
        out.append(slice(o, o + 27 * ci * co + co))
        o += 27 * ci * co + co
    return out


def test_unet_forward_parity():
    say("unet forward")
    n, batch = 8, 2
    theta = he_theta(0)
    net = learned.UNet(2, 0)
    net.set_params(theta)
    x = np.random.default_rng(1).normal(size=batch * n ** 3 * 2)
    xv = GpuVector.from_array(x, F.F32)
    y = net.forward(xv, batch, n).numpy()
    # the checker rounds the same operands to BF16; FP32 vs FP64 sums remain
    theta32 = theta.astype(np.float32).astype(np.float64)
    yo = po.unet_forward(theta32, x.astype(np.float32).astype(np.float64), batch, n, bf16=True)
    assert np.abs(y - yo).max() <= 1e-2 * np.abs(yo).max()


def test_unet_gradient_parity():
    say("unet gradient")
    n, batch = 8, 2
    theta = he_theta(2)
    net = learned.UNet(2, 0)
    net.set_params(theta)
    rng = np.random.default_rng(3)
    x = rng.normal(size=batch * n ** 3 * 2).astype(np.float32).astype(np.float64)
    t = rng.normal(size=batch * n ** 3).astype(np.float32).astype(np.float64)
    loss, g = net.loss_grad(GpuVector.from_array(x, F.F32), GpuVector.from_array(t, F.F32), batch, n)
    theta32 = theta.astype(np.float32).astype(np.float64)
    lo, go = po.unet_loss_grad(theta32, x, t, batch, n, bf16=True)
    assert abs(loss - lo) <= 1e-3 * abs(lo)
    for l, sl in enumerate(layer_slices()):
        d = np.linalg.norm(g[sl] - go[sl])
        assert d <= 3e-2 * np.linalg.norm(go[sl]) + 1e-12, (l, d, np.linalg.norm(go[sl]))


def test_unet_training_reduces_loss():
    say("unet training")
    n = 16
    X, T, cnt = learned.generate_pairs(n=n, problems=16, snapshots=(1, 2), seed=5, base_level=1)
    assert cnt == 32
    net = learned.UNet(2, 7)
    hist, vloss = learned.train(net, X, T, cnt, n, epochs=12, batch=8, lr=2e-3, seed=1)
    assert hist[-1] < 0.8 * hist[0], hist


def _problem(n, seed, kind):
    kap = learned.coefficient(kind, n, seed)
    return kap


def test_learned_fcg_and_fallback():
    """A briefly trained network inside FCG: FP64-verified convergence, the
    iteration count against plain FCG, and the fallback forced from the host
    (an untrained network with a one-step stall window)."""
    say("learned fcg")
    n = 32
    X, T, cnt = learned.generate_pairs(n=n, problems=48, snapshots=(1, 2, 3), seed=11)
    net = learned.UNet(2, 3)
    hist, vloss = learned.train(net, X, T, cnt, n, epochs=6, batch=16, lr=1e-3, seed=2)
    say(f"train losses {hist} holdout {vloss}")
    kap = _problem(n, 987654, "inclusions")
    kv = GpuVector.from_array(kap.ravel(), F.F64)
    cs = it.GeometricCS(it.SOR(1.0), 2, 2, 2, symmetric=True)
    H = it.Hierarchy.lattice(cs, 3, 1, 1, 6, kappa=(kv, n))
    A, b = H.top_matrix(), H.rhs
    bn = np.linalg.norm(b.numpy())
    plain = it.solve(it.LinearSolver(it.FCG(cs), true_residual=True, failure_if=(">", "step", 300)),
                     {"matrix": A, "rhs": b, "hierarchy": H})
    learned.attach(H, net, kv, n)
    lrn = it.solve(it.LinearSolver(it.FCG(cs), true_residual=True, learned=True,
                                   failure_if=(">", "step", 300)),
                   {"matrix": A, "rhs": b, "hierarchy": H})
    say(f"plain FCG {plain['steps']} steps, learned FCG {lrn['steps']} steps")
    for bb in (plain, lrn):
        assert bb["status"] == "success"
        assert bb["true_residual"] / bn < 1e-10
    diff = np.linalg.norm(lrn["solution"].numpy() - plain["solution"].numpy())
    assert diff / np.linalg.norm(plain["solution"].numpy()) < 1e-9
    # forced fallback: untrained network, stall window 1
    bad = learned.UNet(2, 99)
    learned.attach(H, bad, kv, n)
    fb = it.solve(it.LinearSolver(it.FCG(cs), true_residual=True, learned=True, stall_window=1,
                                  failure_if=(">", "step", 300)),
                  {"matrix": A, "rhs": b, "hierarchy": H})
    assert fb["status"] == "success" and fb["true_residual"] / bn < 1e-10
    assert fb["fell_back"]
