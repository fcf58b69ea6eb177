"""GPU tests of the solver loop's <cg> restart-cycle and <safe-linear-solver> fallback against
the oracle, of per-cell factors after re-assembly, of context checks, of captured V-cycles
after a level matrix changes its storage, and of calls from several threads.  Every call goes
through the C-ABI."""
import ctypes
import sys
import threading

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from gnu_femlisp_amd import flgpu as F
from gnu_femlisp_amd import iteration as it
from gnu_femlisp_amd import lattice as lat
from gnu_femlisp_amd.flgpu import check, lib
from gnu_femlisp_amd.matlisp import Context, GpuVector, gemm
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(x, y):
    d = np.linalg.norm(y)
    return np.linalg.norm(x - y) / (d if d > 0 else 1.0)


def _solve(H, cs, itc, b, maxsteps, tol=1e-10, safe=False, restart=None, A=None):
    itn = itc(cs, restart_cycle=restart)
    cls = it.SafeLinearSolver if safe else it.LinearSolver
    return it.solve(cls(itn, success_if=("<", "reduction", tol), failure_if=(">", "step", maxsteps)),
                    {"matrix": A or H.top_matrix(), "rhs": b, "hierarchy": H})


@pytest.mark.parametrize("outer", [po.OUTER_CG, po.OUTER_FCG])
@pytest.mark.parametrize("restart", [1, 3, 7])
def test_restart_cycle_parity(outer, restart):
    """<cg> with restart-cycle (krylow.lisp:101-105,128-131): same steps, defect history and
    solution as the oracle."""
    say(f"restart cycle {restart} outer {outer}")
    rng = np.random.default_rng(5)
    kap = np.exp(rng.normal(scale=1.0, size=(8, 8, 8)))
    oh = po.Hierarchy(3, 1, 1, 4, kappa=kap)
    b = rng.uniform(-1, 1, oh.ndofs())
    b[po.boundary(3, 1, 8)] = 0
    prm = po.params(outer=outer, smoother=po.SMOOTH_SOR, omega=1.0, ordering=po.ORDER_COLOUR,
                    pre_steps=2, post_steps=2, base_level=1, symmetric=1, tol=1e-10,
                    maxsteps=300, restart_cycle=restart)
    st, xo, so, dno = oh.solve(prm, b)
    assert st == 0
    cs = it.GeometricCS(it.SOR(1.0), 2, 2, 1, symmetric=True)
    H = it.Hierarchy.lattice(cs, 3, 1, 1, 4, kappa=kap)
    itc = it.CG if outer == po.OUTER_CG else it.FCG
    bb = _solve(H, cs, itc, GpuVector.from_array(b), 300, restart=restart)
    assert bb["status"] == "success"
    assert bb["steps"] == so
    assert np.allclose(bb["defnorms"], dno, rtol=1e-8)
    assert relerr(bb["solution"].numpy(), xo) < 1e-10


def test_restart_cycle_plain_cg_parity():
    say("restart cycle, plain CG")
    oh = po.Hierarchy(2, 1, 1, 5)
    b = oh.rhs()
    prm = po.params(outer=po.OUTER_CG_PLAIN, tol=1e-10, maxsteps=400, restart_cycle=10)
    st, xo, so, dno = oh.solve(prm, b)
    A = lat.assemble(2, 1, 16)[0]
    bb = it.solve(it.LinearSolver(it.CG(restart_cycle=10), failure_if=(">", "step", 400)),
                  {"matrix": A, "rhs": GpuVector.from_array(b)})
    assert st == 0 and bb["status"] == "success" and bb["steps"] == so
    assert np.allclose(bb["defnorms"], dno, rtol=1e-8)
    assert relerr(bb["solution"].numpy(), xo) < 1e-10


def test_safe_linear_solver_lu_fallback_parity():
    """<safe-linear-solver> (linsolve.lisp:45-56,75-89) on the 2D 32^2 configuration: CG
    fails at its step limit, <lu> takes one step and the loop succeeds, as in the oracle."""
    say("safe linear solver, 2D 32^2")
    oh = po.Hierarchy(2, 1, 1, 6)
    b = oh.rhs()
    prm = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_JACOBI, omega=0.6, pre_steps=2,
                    post_steps=2, base_level=0, tol=1e-10, maxsteps=2, fallback=po.FALLBACK_LU)
    st, xo, so, dno = oh.solve(prm, b)
    assert st == 0 and so == 4
    cs = it.GeometricCS(it.Jacobi(0.6), 2, 2, 0)
    H = it.Hierarchy.lattice(cs, 2, 1, 1, 6)
    bb = _solve(H, cs, it.CG, H.rhs, 2, safe=True)
    assert bb["status"] == "success" and bb["fallback_used"]
    assert bb["steps"] == so
    assert np.allclose(bb["defnorms"][:so], dno[:so], rtol=1e-8)
    assert bb["defnorms"][so] < 1e-10 * bb["defnorms"][0]
    assert relerr(bb["solution"].numpy(), xo) < 1e-10
    # the same solver without the fallback fails at the step limit
    plain = _solve(H, cs, it.CG, H.rhs, 2)
    assert plain["status"] == "failure" and not plain["fallback_used"]


@pytest.mark.parametrize("prec", [F.F64, F.F32])
def test_safe_linear_solver_fallback_failure(prec):
    """a second failure returns status 2: femlisp's "The fallback iteration has failed." """
    say(f"safe linear solver fails twice, prec {prec}")
    A, b = lat.assemble(2, 1, 8, prec=prec)
    bb = it.solve(it.SafeLinearSolver(it.CG(), success_if=("<", "reduction", 1e-300),
                                      failure_if=(">", "step", 1)),
                  {"matrix": A, "rhs": b})
    assert bb["status"] == "fallback-failed" and bb["fallback_used"] and bb["steps"] == 3


def test_fallback_rejects_large_systems():
    A, b = lat.assemble(3, 1, 32)
    with pytest.raises(F.FlgpuError):
        it.solve(it.SafeLinearSolver(it.CG(), failure_if=(">", "step", 1)), {"matrix": A, "rhs": b})


def _kap(seed, n, nk):
    return np.exp(np.random.default_rng(seed).normal(scale=1.0, size=(nk,) * 3))


@pytest.mark.parametrize("prec", [F.F64, F.F32])
@pytest.mark.parametrize("nk_factor", [1, 2])
def test_lattice_reassemble_refreshes_cell_factors(prec, nk_factor):
    """flgpu_lattice_reassemble on a matrix in the per-cell form: products, <jacobi> steps and
    <sor> sweeps equal those of a fresh assembly with the new coefficient (nk = n keeps the
    per-cell form with new factors; nk = 2n drops it)."""
    n = 12
    say(f"reassemble prec {prec} nk {nk_factor * n}")
    k1, k2 = _kap(1, n, n), _kap(2, n, nk_factor * n)
    A, b = lat.assemble(3, 1, n, k1, prec=prec)
    assert A.storage() == F.STORAGE_CELL
    lat.reassemble(A, k2)
    R, _ = lat.assemble(3, 1, n, k2, prec=prec)
    assert A.storage() == R.storage()
    assert A.storage() == (F.STORAGE_CELL if nk_factor == 1 else F.STORAGE_STENCIL)
    rp, ci, v = A.csr()
    orp, oci, ov, _ = po.assemble(3, 1, n, k2, 1.0)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    tol = 1e-13 if prec == F.F64 else 1e-6
    assert np.abs(v - ov).max() <= tol * np.abs(ov).max()
    N = A.nrows
    y = np.random.default_rng(3).uniform(-1, 1, N)
    outs = []
    for M in (A, R):
        yv = GpuVector.from_array(y, prec)
        z = GpuVector(N, prec)
        gemm(1.0, M, yv, 0.0, z)
        x, t = GpuVector(N, prec), GpuVector(N, prec)
        bb = GpuVector.from_array(b.numpy(), prec)
        check(lib().flgpu_jacobi_steps(M.h, 0.8, 3, x.h, bb.h, t.h))
        xs, rs = GpuVector(N, prec), GpuVector(N, prec)
        check(lib().flgpu_sor_iterate(M.h, 1.0, 0, xs.h, bb.h, rs.h))
        outs.append((z.numpy(), x.numpy(), xs.numpy()))
    for a, r in zip(*outs):
        assert np.array_equal(a, r)
    Ao = sp.csr_matrix((ov, oci, orp))
    assert relerr(outs[0][0], Ao @ y) < (1e-13 if prec == F.F64 else 1e-6)


def test_context_checks():
    """vectors and hierarchies of another context are rejected before any work is queued"""
    say("context checks")
    c2 = Context(0)
    cs = it.GeometricCS(it.Jacobi(0.6), 2, 2, 0)
    H = it.Hierarchy.lattice(cs, 2, 1, 1, 4)
    A = H.top_matrix()
    N = A.nrows
    b_other = GpuVector.from_array(np.ones(N), ctx=c2)
    x_other = GpuVector(N, ctx=c2)
    x, r = GpuVector(N), GpuVector(N)
    prm = it.LinearSolver(it.CG(cs)).params()
    rep = F.SolveReport()
    for bv, xv in ((b_other, x), (H.rhs, x_other)):
        st = lib().flgpu_solve(H.h, A.h, ctypes.byref(prm), bv.h, xv.h, ctypes.byref(rep), None, 0)
        assert st == -1
    H2 = it.Hierarchy.lattice(cs, 2, 1, 1, 4, ctx=c2)
    st = lib().flgpu_solve(H2.h, A.h, ctypes.byref(prm), H.rhs.h, x.h, ctypes.byref(rep), None, 0)
    assert st == -1
    assert lib().flgpu_mg_apply(H.h, x_other.h, H.rhs.h, r.h) == -1
    assert lib().flgpu_mg_apply(H.h, x.h, b_other.h, r.h) == -1
    # the same calls with the right context work
    bb = it.solve(it.LinearSolver(it.CG(cs)), {"matrix": A, "rhs": H.rhs, "hierarchy": H})
    assert bb["status"] == "success"


@pytest.mark.parametrize("smoother", ["jacobi", "sor"])
def test_graphs_follow_level_matrix_changes(smoother):
    """A captured V-cycle is dropped when a level matrix that the hierarchy hands out changes
    its storage form or colouring: the next solve captures again and equals the eager one."""
    say(f"graph cache after a storage change, {smoother}")
    kap = _kap(4, 16, 16)
    sm = it.Jacobi(0.8) if smoother == "jacobi" else it.SOR(1.0)
    res = {}
    for graphs in (True, False):
        cs = it.GeometricCS(sm, 2, 2, 1, symmetric=True)
        H = it.Hierarchy.lattice(cs, 3, 1, 1, 5, kappa=kap)
        H.use_graphs(graphs)
        A = H.top_matrix()
        first = _solve(H, cs, it.CG, H.rhs, 100)
        assert first["status"] == "success"
        top = H.top_matrix()
        assert top.storage(F.STORAGE_SELL) == F.STORAGE_SELL  # frees the per-cell factors
        mid = H.level_matrix(2)
        assert mid.storage(F.STORAGE_SELL) == F.STORAGE_SELL  # frees the stencil planes
        second = _solve(H, cs, it.CG, H.rhs, 100, A=A)
        res[graphs] = (first, second)
    for k in (0, 1):
        g, e = res[True][k], res[False][k]
        assert g["steps"] == e["steps"]
        assert np.array_equal(g["solution"].numpy(), e["solution"].numpy())


def test_solve_work_vectors_reused_across_sizes():
    """A context keeps flgpu_solve's work vectors between solves and grows them when a larger
    system comes: solves of different sizes and precisions in turn on one context equal the
    same solves on fresh contexts bit for bit."""
    say("solve work reuse")
    shared = Context(0)

    def run(ctx, n, prec):
        kap = _kap(n, n, n)
        if prec == F.F64:
            cs = it.GeometricCS(it.SOR(1.0), 2, 2, 1, symmetric=True)
            H = it.Hierarchy.lattice(cs, 3, 1, 1, int(np.log2(n)) + 1, kappa=kap, ctx=ctx)
            r = it.solve(it.LinearSolver(it.CG(cs), failure_if=(">", "step", 100)),
                         {"matrix": H.top_matrix(), "rhs": H.rhs, "hierarchy": H})
        else:
            A, b = lat.assemble(3, 1, n, kap, prec=prec, ctx=ctx)
            r = it.solve(it.LinearSolver(it.CG(), success_if=("<", "reduction", 1e-5),
                                         failure_if=(">", "step", 400)),
                         {"matrix": A, "rhs": b})
        assert r["status"] == "success"
        return r["steps"], r["solution"].numpy()

    for n, prec in ((16, F.F64), (8, F.F64), (16, F.F32), (32, F.F64), (8, F.F32)):
        a, b = run(shared, n, prec), run(Context(0), n, prec)
        assert a[0] == b[0] and np.array_equal(a[1], b[1]), (n, prec)


def test_concurrent_contexts_and_shared_context():
    """Calls from several threads (femlisp's lparallel workers): solves on two contexts run
    concurrently, and products on one shared context are serialised by its lock; both give
    the results of the same calls made one after another."""
    say("threads")
    kap = _kap(6, 16, 16)

    def solve_on(ctx):
        cs = it.GeometricCS(it.SOR(1.0), 2, 2, 1, symmetric=True)
        H = it.Hierarchy.lattice(cs, 3, 1, 1, 5, kappa=kap, ctx=ctx)
        r = _solve(H, cs, it.CG, H.rhs, 100)
        return r["steps"], r["solution"].numpy()

    ref = solve_on(Context(0))
    ctxs = [Context(0), Context(0)]
    out = [None, None]
    errs = []

    def worker(i):
        try:
            out[i] = solve_on(ctxs[i])
        except Exception as e:  # pragma: no cover - reported below
            errs.append(e)

    ts = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not errs, errs
    for o in out:
        assert o[0] == ref[0] and np.array_equal(o[1], ref[1])

    # one context shared by four threads: products of one matrix into separate vectors
    c = Context(0)
    A, _ = lat.assemble(3, 1, 24, _kap(7, 24, 24), ctx=c)
    N = A.nrows
    ys = [GpuVector.from_array(np.random.default_rng(10 + i).uniform(-1, 1, N), ctx=c)
          for i in range(4)]
    seq = []
    for y in ys:
        z = GpuVector(N, ctx=c)
        for _ in range(5):
            gemm(1.0, A, y, 0.5, z)
        seq.append(z.numpy())
    zs = [GpuVector(N, ctx=c) for _ in range(4)]

    def prod(i):
        try:
            for _ in range(5):
                gemm(1.0, A, ys[i], 0.5, zs[i])
        except Exception as e:  # pragma: no cover
            errs.append(e)

    ts = [threading.Thread(target=prod, args=(i,)) for i in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not errs, errs
    for z, s in zip(zs, seq):
        assert np.array_equal(z.numpy(), s)


@pytest.mark.parametrize("form", [F.STORAGE_CELL, F.STORAGE_STENCIL])
@pytest.mark.parametrize("n,prec", [(40, F.F64), (40, F.F32), (70, F.F64), (70, F.F32),
                                    (128, F.F32)])
def test_cell_sor_sweeps_over_many_tiles(n, prec, form):
    """the per-cell <sor> sweep in plane strips on larger lattices (strips per plane: n = 40,
    6 of 7 lines; n = 70, 8 of 9; n = 128, 4 of 33; the last strip one to three lines shorter)
    and the stencil-form colour launches equal the SELL colour sweep, forwards and backwards,
    and with rhs and residual one vector (femlisp's (iterate q w w))"""
    say(f"sor strips n={n} prec={prec} form={form}")
    rng = np.random.default_rng(n)
    kap = np.exp(rng.normal(scale=1.5, size=(n, n, n)))
    A, b = lat.assemble(3, 1, n, kap, prec=prec)
    B, _ = lat.assemble(3, 1, n, kap, prec=prec)
    assert A.storage(form) == form and B.storage(F.STORAGE_SELL) == F.STORAGE_SELL
    tol = 1e-13 if prec == F.F64 else 2e-6
    y = rng.uniform(-1, 1, A.nrows)
    bh = b.numpy() + rng.uniform(-1, 1, A.nrows)
    for bw in (0, 1):
        for alias in (False, True):
            xs = [GpuVector.from_array(y, prec) for _ in range(2)]
            rs = [GpuVector.from_array(bh, prec) for _ in range(2)]
            bs = [GpuVector.from_array(bh, prec) for _ in range(2)]
            for M, xv, rv, bv in ((A, xs[0], rs[0], bs[0]), (B, xs[1], rs[1], bs[1])):
                for _ in range(2):
                    check(lib().flgpu_sor_iterate(M.h, 1.1, bw, xv.h, (rv if alias else bv).h,
                                                  rv.h))
            assert relerr(xs[0].numpy(), xs[1].numpy()) < tol, (bw, alias)
            assert relerr(rs[0].numpy(), rs[1].numpy()) < 10 * tol, (bw, alias)


@pytest.mark.parametrize("n", [24, 40])
@pytest.mark.parametrize("prec", [F.F64, F.F32])
def test_stencil_sor_strips_nonsymmetric(prec, n):
    """the stencil-form <sor> sweep in strips on a lattice matrix whose values are not
    symmetric (all 27 planes streamed) equals the SELL colour sweep, forwards and backwards"""
    say(f"stencil sor strips, nonsymmetric, n={n} prec={prec}")
    rng = np.random.default_rng(77 + n)
    A, b = lat.assemble(3, 1, n, prec=prec)
    rp, ci, v = A.csr()
    v = v * (1.0 + 0.1 * rng.uniform(-1, 1, v.size))  # unequal A(i, j) and A(j, i)
    B, _ = lat.assemble(3, 1, n, prec=prec)
    A.update_values(v)
    B.update_values(v)
    assert A.storage(F.STORAGE_STENCIL) == F.STORAGE_STENCIL and A.stencil_planes() == 27
    assert B.storage(F.STORAGE_SELL) == F.STORAGE_SELL
    tol = 1e-13 if prec == F.F64 else 2e-6
    y = rng.uniform(-1, 1, A.nrows)
    for bw in (0, 1):
        xs = [GpuVector.from_array(y, prec) for _ in range(2)]
        rs = [GpuVector(A.nrows, prec) for _ in range(2)]
        bs = GpuVector.from_array(b.numpy(), prec)
        for M, xv, rv in ((A, xs[0], rs[0]), (B, xs[1], rs[1])):
            check(lib().flgpu_sor_iterate(M.h, 1.1, bw, xv.h, bs.h, rv.h))
        assert relerr(xs[0].numpy(), xs[1].numpy()) < tol, bw
        assert relerr(rs[0].numpy(), rs[1].numpy()) < 10 * tol, bw


@pytest.mark.parametrize("n,prec", [(16, F.F64), (16, F.F32), (7, F.F64)])
def test_base_level_dense_solve(n, prec):
    """a one-level hierarchy is its base-level solve: x = A^-1 b from the compressed dense
    inverse (the Dirichlet DOFs kept as unit rows; at n = 16 the other 3375 DOFs are inverted in
    53 blocks of 64 pivots, the last of 47; at n = 7, 216 in 4, the last of 24), against scipy's
    sparse LU on the same matrix values"""
    say(f"base-level dense solve n={n} prec={prec}")
    rng = np.random.default_rng(n)
    kap = np.exp(rng.normal(scale=1.5, size=(n,) * 3))
    cs = it.GeometricCS(it.Jacobi(0.9), 2, 2, 0, prec=prec)
    H = it.Hierarchy.lattice(cs, 3, 1, n, 1, kappa=kap)
    rp, ci, v = H.top_matrix().csr()
    M = sp.csr_matrix((v, ci, rp)).tocsc()
    bh = rng.uniform(-1, 1, M.shape[0])
    ref = spla.spsolve(M, bh.astype(np.float32).astype(np.float64) if prec == F.F32 else bh)
    x = GpuVector(M.shape[0], prec)
    b = GpuVector.from_array(bh, prec)
    r = GpuVector.from_array(bh, prec)
    H.apply(x, b, r)
    assert relerr(x.numpy(), ref) < (1e-11 if prec == F.F64 else 1e-5)
    # from x0 with r = b - A x0: x0 + A^-1 r is the solution again
    x0 = rng.uniform(-1, 1, M.shape[0])
    x = GpuVector.from_array(x0, prec)
    r = GpuVector.from_array(bh - M @ x0, prec)
    H.apply(x, b, r)
    assert relerr(x.numpy(), ref) < (1e-11 if prec == F.F64 else 1e-5)


def _gather_host(E, n, M):
    """femlisp's scatter of element matrices (global-local-matrix-operation) restated: for every
    row P and lattice neighbour Q, the sum in ascending cell order, in FP64, of E over the cells
    holding both; Dirichlet rows and columns eliminated (identity diagonal)"""
    vals = []
    for P in np.ndindex(M, M, M):
        rb = any(p in (0, M - 1) for p in P)
        for d in np.ndindex(3, 3, 3):
            Q = tuple(p + dd - 1 for p, dd in zip(P, d))
            if any(q < 0 or q >= M for q in Q):
                continue
            cb = any(q in (0, M - 1) for q in Q)
            if rb or cb:
                vals.append(1.0 if (rb and Q == P) else 0.0)
                continue
            acc = 0.0
            rng = [range(max(max(p, q) - 1, 0), min(min(p, q), n - 1) + 1) for p, q in zip(P, Q)]
            for c in np.ndindex(*[len(r) for r in rng]):
                cc = tuple(r[i] for r, i in zip(rng, c))
                cell = (cc[0] * n + cc[1]) * n + cc[2]
                la = ((P[0] - cc[0]) * 2 + (P[1] - cc[1])) * 2 + (P[2] - cc[2])
                lb = ((Q[0] - cc[0]) * 2 + (Q[1] - cc[1])) * 2 + (Q[2] - cc[2])
                acc += E[cell, la, lb]
            vals.append(acc)
    return np.array(vals)


@pytest.mark.parametrize("prec,mode", [(F.F64, F.ASM_F64), (F.F64, F.ASM_MFMA),
                                       (F.F32, F.ASM_F64), (F.F32, F.ASM_MFMA)])
@pytest.mark.parametrize("nk_factor", [1, 2])
def test_one_pass_q1_assembly_equals_element_gather(prec, mode, nk_factor):
    """3D Q1 values from the one-pass kernel (element matrices in LDS only) equal the element
    kernel's matrices gathered on the host in ascending cell order, bit for bit"""
    n = 9
    say(f"one-pass assembly prec {prec} mode {mode} nk {nk_factor * n}")
    kap = _kap(11, n, nk_factor * n)
    A, b = lat.assemble(3, 1, n, kap, prec=prec, mode=mode)
    E = lat.element_matrices(3, 1, n, kap, prec=prec, mode=mode).reshape(n ** 3, 8, 8)
    ref = _gather_host(E, n, n + 1)
    if prec == F.F32:
        ref = ref.astype(np.float32).astype(np.float64)
    assert np.array_equal(A.csr()[2], ref)
    # the load vector of f = 1 against the oracle
    orp, oci, ov, ob = po.assemble(3, 1, n, kap, 1.0)
    assert relerr(b.numpy(), ob) < (1e-14 if prec == F.F64 else 1e-7)
