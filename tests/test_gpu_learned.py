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


def bf16(a):
    """round to BF16 (nearest, ties to even) as the kernels do when operands enter MFMA"""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def test_unet_forward_closed_forms():
    """the parameter sets that reduce the checker to closed forms (tests/test_oracle_unet.py)
    on the GPU network: a single tap is the shift with zero padding, bit for bit on BF16
    inputs; the pools, upsamplings and [up, skip] concatenations give the block means"""
    say("unet closed forms")
    from tests.test_oracle_unet import CENTER, block_mean, positive_input, tap, theta_from
    n, batch = 8, 2
    net = learned.UNet(2, 0)
    x = positive_input(n, batch, 5)
    xv = GpuVector.from_array(x.ravel(), F.F32)
    for d in ((1, 0, -1), (-1, -1, -1)):
        net.set_params(theta_from([(0, tap(*d), 0, 0, 1.0), (1, CENTER, 0, 0, 1.0),
                                   (7, CENTER, 32, 0, 1.0), (8, CENTER, 0, 0, 1.0)]))
        y = net.forward(xv, batch, n).numpy().reshape(batch, n, n, n)
        ref = np.zeros((batch, n, n, n))
        zs, ys, xs = (slice(max(0, -k), n - max(0, k)) for k in d)
        zd, yd, xd = (slice(max(0, k), n - max(0, -k)) for k in d)
        ref[:, zs, ys, xs] = bf16(x[..., 0])[:, zd, yd, xd]
        assert np.array_equal(y, ref), d
    ent = [(l, CENTER, 0, 0, 1.0) for l in (0, 1, 2, 3, 4, 5)]
    ent += [(6, CENTER, 0, 0, 1.0), (6, CENTER, 64, 0, 2.0), (7, CENTER, 0, 0, 1.0),
            (7, CENTER, 32, 0, 3.0), (8, CENTER, 0, 0, 1.0)]
    net.set_params(theta_from(ent, biases=[(8, 0, 0.25)]))
    y = net.forward(xv, batch, n).numpy().reshape(batch, n, n, n)
    x0 = x[..., 0]
    ref = block_mean(x0, 4) + 2 * block_mean(x0, 2) + 3 * x0 + 0.25
    # every layer rounds its input to BF16 (relative 2^-9 each, a few on every path)
    assert np.abs(y - ref).max() <= 2e-2 * np.abs(ref).max()


def test_unet_adam_step_matches_checker():
    """one training step on the GPU: the parameters and moments Adam leaves equal the
    checker's Adam (po.adam, FP64) applied to the GPU's own gradient of the same batch"""
    say("unet adam step")
    n, batch, lr, step = 8, 2, 1e-3, 5
    rng = np.random.default_rng(8)
    theta = he_theta(4).astype(np.float32).astype(np.float64)
    m = (rng.normal(size=theta.size) * 1e-3).astype(np.float32).astype(np.float64)
    v = (rng.uniform(size=theta.size) * 1e-6).astype(np.float32).astype(np.float64)
    net = learned.UNet(2, 0)
    net.set_params(theta, m, v, step)
    x = GpuVector.from_array(rng.normal(size=batch * n ** 3 * 2), F.F32)
    t = GpuVector.from_array(rng.normal(size=batch * n ** 3), F.F32)
    loss0, g = net.loss_grad(x, t, batch, n)
    loss1 = net.train_step(x, t, batch, n, lr)
    assert abs(loss1 - loss0) <= 1e-6 * abs(loss0)  # the loss before the update
    th_gpu, m_gpu, v_gpu, st = net.params()
    assert st == step + 1
    # the GPU's Adam runs in FP32 with the FP32 roundings of b1, b2 and eps (0.999f differs
    # from 0.999 by 1.3e-5 relative): the checker gets the same constants
    b1, b2, eps = (float(np.float32(c)) for c in (0.9, 0.999, 1e-8))
    th_ref, m_ref, v_ref = po.adam(theta, g, m, v, step + 1, lr=lr, b1=b1, b2=b2, eps=eps)
    # FP32 moments: m = b1 m + (1 - b1) g may cancel, so its error is bounded by the rounding
    # of its terms, not of the result; v adds positive terms
    u = 2.0 ** -23
    assert np.all(np.abs(m_gpu - m_ref) <= 3 * u * (0.9 * np.abs(m) + 0.1 * np.abs(g)) + 1e-30)
    assert np.allclose(v_gpu, v_ref, rtol=4 * u, atol=1e-30)
    # the update from the GPU's own moments, to FP32 rounding of theta and of the update
    t = step + 1
    d_exp = -lr * (m_gpu / (1 - b1 ** t)) / (np.sqrt(v_gpu / (1 - b2 ** t)) + eps)
    d_gpu = th_gpu - theta
    assert np.all(np.abs(d_gpu - d_exp) <= 1e-5 * np.abs(d_exp) + 2 * u * np.abs(theta) + 1e-12)
    # and the checker's update agrees wherever the moments do not cancel
    d_ref = th_ref - theta
    sel = np.abs(m_ref) > 1e-3 * np.abs(m_ref).max()
    assert np.allclose(d_gpu[sel], d_ref[sel], rtol=1e-4, atol=2 * u * np.abs(theta).max())


def test_unet_training_batch_32_on_32_cubed():
    """the training configuration's shapes: random-coefficient Q1 problems on 32^3 cells
    (kappa = exp(g), g a smoothed Gaussian field, correlation length 4 cells, sigma 2),
    pairs from the FP64 multigrid, Adam with lr 1e-3 and batch 32: the loss falls"""
    say("unet training, batch 32, 32^3")
    n = 32
    X, T, cnt = learned.generate_pairs(n=n, problems=256, snapshots=(1,), kinds=("gaussian",),
                                       seed=21)
    assert cnt == 256
    net = learned.UNet(2, 5)
    hist, vloss = learned.train(net, X, T, cnt, n, epochs=8, batch=32, lr=1e-3, seed=3)
    say(f"losses {hist} holdout {vloss}")
    assert hist[-1] < 0.9 * hist[0], hist


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


def test_learned_correction_not_added_leaves_fcg_unchanged():
    """A network whose step carries less than 1% of the cycle step's energy (here: all
    parameters zero, so its correction is zero) is never added to the cycle's output and is
    dropped in the first step without restarting FCG: the learned solve then equals plain FCG
    with the same hierarchy bit for bit (steps, defect norms, solution)."""
    say("learned correction not added")
    n = 16
    kap = _problem(n, 31337, "inclusions")
    kv = GpuVector.from_array(kap.ravel(), F.F64)
    cs = it.GeometricCS(it.Jacobi(0.9), 2, 2, 1, transfer="blackbox")
    H = it.Hierarchy.lattice(cs, 3, 1, 1, 5, kappa=(kv, n))
    A, b = H.top_matrix(), H.rhs
    plain = it.solve(it.LinearSolver(it.FCG(cs), true_residual=True, failure_if=(">", "step", 200)),
                     {"matrix": A, "rhs": b, "hierarchy": H})
    net = learned.UNet(2, 7)
    net.set_params(np.zeros(net.nparams))
    learned.attach(H, net, kv, n)
    lrn = it.solve(it.LinearSolver(it.FCG(cs), true_residual=True, learned=True,
                                   stall_window=1000, failure_if=(">", "step", 200)),
                   {"matrix": A, "rhs": b, "hierarchy": H})
    assert plain["status"] == "success" and lrn["status"] == "success"
    assert lrn["fell_back"]
    assert lrn["steps"] == plain["steps"]
    assert np.array_equal(np.asarray(lrn["defnorms"]), np.asarray(plain["defnorms"]))
    assert np.array_equal(lrn["solution"].numpy(), plain["solution"].numpy())


def test_learned_solves_reuse_the_hook():
    """The hierarchy keeps the learned correction's buffers between solves: two learned solves
    in a row give the same result bit for bit, a solve after another network is attached
    equals one on a fresh hierarchy with that network, and the first network again gives the
    first result."""
    say("learned hook reuse")
    n = 16
    kap = _problem(n, 4242, "inclusions")
    kv = GpuVector.from_array(kap.ravel(), F.F64)
    nets = [learned.UNet(2, 5), learned.UNet(2, 6)]

    def solve(H, net):
        learned.attach(H, net, kv, n)
        r = it.solve(it.LinearSolver(it.FCG(cs), true_residual=True, learned=True,
                                     stall_window=1000, failure_if=(">", "step", 200)),
                     {"matrix": H.top_matrix(), "rhs": H.rhs, "hierarchy": H})
        assert r["status"] == "success"
        return r["steps"], r["solution"].numpy()

    cs = it.GeometricCS(it.SOR(1.0), 2, 2, 1, symmetric=True)
    H = it.Hierarchy.lattice(cs, 3, 1, 1, 5, kappa=(kv, n))
    a1, a2 = solve(H, nets[0]), solve(H, nets[0])
    b = solve(H, nets[1])
    a3 = solve(H, nets[0])
    Hf = it.Hierarchy.lattice(cs, 3, 1, 1, 5, kappa=(kv, n))
    bf = solve(Hf, nets[1])
    for x, y in ((a1, a2), (a1, a3), (b, bf)):
        assert x[0] == y[0] and np.array_equal(x[1], y[1])
