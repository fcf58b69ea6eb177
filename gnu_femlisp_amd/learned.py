"""Learned coarse correction: network handle, training data from the GPU's
own FP64 multigrid, training loop and attachment to a hierarchy.

Training pairs: for random-coefficient problems, the FP64 GMG is run as a
stationary iteration from x = 0; after k = 1..K V-cycles the residual
r_k = b - A x_k and the error e_k = x* - x_k (x* solved to 1e-13) form one
pair.  Inputs [r/rms(r), log(kappa)/4], targets e/rms(e).  In the flexible
CG the network's output is used along the energy-optimal step, so only its
shape matters.
"""
import ctypes
import time

import numpy as np

from . import flgpu as F
from . import iteration as it
from .fields import gaussian_log_field, inclusions
from .flgpu import check, lib, ptr
from .matlisp import GpuVector, axpy, compute_residual, copy, default_context


class UNet:
    def __init__(self, cin=2, seed=0, ctx=None):
        self.ctx = ctx or default_context()
        h = ctypes.c_void_p()
        check(lib().flgpu_unet_create(self.ctx.h, cin, seed, ctypes.byref(h)))
        self.h = h
        self.cin = cin
        n = ctypes.c_long()
        check(lib().flgpu_unet_nparams(h, ctypes.byref(n)))
        self.nparams = n.value

    def params(self):
        th, m, v = (np.zeros(self.nparams) for _ in range(3))
        step = ctypes.c_int()
        check(lib().flgpu_unet_get_params(self.h, ptr(th), ptr(m), ptr(v), ctypes.byref(step)))
        return th, m, v, step.value

    def set_params(self, theta, m=None, v=None, step=0):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
        theta, m, v = conv(theta), conv(m), conv(v)
        check(lib().flgpu_unet_set_params(self.h, ptr(theta), ptr(m), ptr(v), step))

    def save(self, path):
        th, m, v, step = self.params()
        np.savez(path, theta=th, m=m, v=v, step=step)

    def load(self, path):
        d = np.load(path)
        self.set_params(d["theta"], d["m"], d["v"], int(d["step"]))

    def forward(self, x, batch, n, y=None):
        y = y or GpuVector(batch * n ** 3, F.F32, self.ctx)
        check(lib().flgpu_unet_forward(self.h, batch, n, x.h, y.h))
        return y

    def loss_grad(self, x, t, batch, n):
        loss = ctypes.c_double()
        g = np.zeros(self.nparams)
        check(lib().flgpu_unet_loss_grad(self.h, batch, n, x.h, t.h, ctypes.byref(loss), ptr(g)))
        return loss.value, g

    def train_step(self, x, t, batch, n, lr):
        loss = ctypes.c_double()
        check(lib().flgpu_unet_train_step(self.h, batch, n, x.h, t.h, lr, ctypes.byref(loss)))
        return loss.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().flgpu_unet_destroy(self.h)
                self.h = None
        except Exception:
            pass


def coefficient(kind, n, seed):
    if kind == "gaussian":
        return gaussian_log_field(n, seed, corr=4.0, sigma=2.0)
    if kind == "inclusions":
        return inclusions(n, seed=seed, frac=0.30, rmin=3.0, rmax=8.0, contrast=1e4)
    raise ValueError(kind)


def generate_pairs(n=32, problems=64, snapshots=(1, 2, 3), kinds=("gaussian", "inclusions"),
                   seed=1, base_level=2, log=None, ctx=None):
    """Residual/error pairs from the FP64 GMG on random-coefficient problems.
    Returns (X, T, count) device tensors."""
    ctx = ctx or default_context()
    levels = int(round(np.log2(n))) + 1
    S = problems * len(snapshots)
    X = GpuVector(S * n ** 3 * 2, F.F32, ctx)
    T = GpuVector(S * n ** 3, F.F32, ctx)
    cs = it.GeometricCS(it.SOR(1.0), 2, 2, base_level, symmetric=True, prec=F.F64)
    k0 = coefficient(kinds[0], n, seed)
    kv = GpuVector.from_array(k0.ravel(), F.F64, ctx)
    H = it.Hierarchy.lattice(cs, 3, 1, 1, levels, kappa=(kv, n))
    A = H.top_matrix()
    N = A.nrows
    rng = np.random.default_rng(seed)
    interior = np.zeros((n + 1,) * 3, bool)
    interior[1:n, 1:n, 1:n] = True
    interior = interior.ravel()
    x, xs, r, e = (GpuVector(N, F.F64, ctx) for _ in range(4))
    bv = GpuVector(N, F.F64, ctx)
    s = 0
    t0 = time.time()
    for p in range(problems):
        kind = kinds[p % len(kinds)]
        kap = coefficient(kind, n, seed * 100003 + p)
        kv.upload(kap.ravel())
        check(lib().flgpu_mg_reassemble(H.h, kv.h, n))
        b = np.where(interior, rng.uniform(-1, 1, N), 0.0)
        bv.upload(b)
        bb = it.solve(it.LinearSolver(it.CG(cs), success_if=("<", "reduction", 1e-13),
                                      failure_if=(">", "step", 400)),
                      {"matrix": A, "rhs": bv, "hierarchy": H, "solution": xs})
        x.upload(np.zeros(N))
        copy(bv, r)
        done = 0
        for k in range(1, max(snapshots) + 1):
            H.apply(x, bv, r)
            compute_residual(A, x, bv, r)
            if k in snapshots:
                copy(xs, e)
                axpy(-1.0, x, e)
                check(lib().flgpu_unet_pack(r.h, kv.h, n, X.h, s, 0))
                check(lib().flgpu_unet_pack(e.h, None, n, T.h, s, 1))
                s += 1
                done += 1
        if log and (p + 1) % max(1, problems // 8) == 0:
            log(f"[learned] pairs: {p + 1}/{problems} problems ({time.time() - t0:.1f}s)")
    ctx.synchronize()
    return X, T, s


def train(net, X, T, count, n, epochs=10, batch=32, lr=1e-3, seed=0, holdout=0.05, log=None):
    """Adam on the pairs; returns (train_losses per epoch, holdout loss)."""
    rng = np.random.default_rng(seed)
    order = rng.permutation(count)
    nval = max(1, int(round(holdout * count))) if count > batch else 0
    val, tr = order[:nval], order[nval:]
    xb = GpuVector(batch * n ** 3 * 2, F.F32, net.ctx)
    tb = GpuVector(batch * n ** 3, F.F32, net.ctx)
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        perm = rng.permutation(tr)
        losses = []
        for i in range(0, len(perm) - batch + 1, batch):
            idx = np.ascontiguousarray(perm[i:i + batch], np.int64)
            check(lib().flgpu_vec_gather_rows(X.h, n ** 3 * 2, ptr(idx), batch, xb.h))
            check(lib().flgpu_vec_gather_rows(T.h, n ** 3, ptr(idx), batch, tb.h))
            losses.append(net.train_step(xb, tb, batch, n, lr))
        hist.append(float(np.mean(losses)) if losses else float("nan"))
        if log:
            log(f"[learned] epoch {ep + 1}/{epochs}: loss {hist[-1]:.4f} ({time.time() - t0:.1f}s)")
    vloss = None
    if nval:
        vl = []
        xv = GpuVector(n ** 3 * 2, F.F32, net.ctx)
        tv = GpuVector(n ** 3, F.F32, net.ctx)
        for j in val:
            idx = np.array([j], np.int64)
            check(lib().flgpu_vec_gather_rows(X.h, n ** 3 * 2, ptr(idx), 1, xv.h))
            check(lib().flgpu_vec_gather_rows(T.h, n ** 3, ptr(idx), 1, tv.h))
            vl.append(net.loss_grad(xv, tv, 1, n)[0])
        vloss = float(np.mean(vl))
    return hist, vloss


def coarse_kappa(kappa, n):
    """the coefficient on n cells per axis from a finer per-cell field: the geometric mean
    over each coarse cell's fine cells (the network reads log(kappa))"""
    nf = kappa.shape[0]
    f = nf // n
    assert f * n == nf and kappa.shape == (nf, nf, nf)
    lk = np.log(kappa).reshape(n, f, n, f, n, f).mean(axis=(1, 3, 5))
    return np.exp(lk)


def attach(hierarchy, net, kappa_vec, n):
    """attach the network to a lattice hierarchy: it runs on the level with n cells per axis
    (the top level or a coarser one) and reads kappa_vec, n^3 cell values on that level"""
    check(lib().flgpu_mg_attach_unet(hierarchy.h, net.h, kappa_vec.h if kappa_vec else None, n))
    hierarchy._unet = (net, kappa_vec)


def attach_coarse(hierarchy, net, kappa, n, ctx=None):
    """attach the network on the level with n cells per axis of a hierarchy whose top-level
    coefficient is the per-cell field kappa (finer): the level's coefficient is its
    geometric mean"""
    kc = coarse_kappa(np.asarray(kappa), n)
    kv = GpuVector.from_array(kc.ravel(), F.F64, ctx or hierarchy.ctx)
    attach(hierarchy, net, kv, n)
    return kv
