"""Vertex-centred block smoothers on the GPU (<geometric-ssc>, <geometric-psc>,
src/iteration/blockit.lisp:113-203, src/iteration/geoblock.lisp:65-157) against the oracle:
single iterates through flgpu_block_iterate, V-cycles and preconditioned solves."""
import ctypes
import sys

import numpy as np
import pytest

from gnu_femlisp_amd import flgpu as F
from gnu_femlisp_amd import iteration as it
from gnu_femlisp_amd import lattice as lat
from gnu_femlisp_amd.flgpu import check, lib
from gnu_femlisp_amd.matlisp import GpuSparseMatrix, GpuVector, default_context
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(x, y):
    d = np.linalg.norm(y)
    return np.linalg.norm(x - y) / (d if d > 0 else 1.0)


# smallest lattices first: one cell (every vertex Dirichlet, each vertex its own colour class)
ITER_CASES = [(1, 2, 1), (2, 2, 1), (3, 2, 1), (3, 2, 2), (3, 1, 1), (1, 2, 8), (2, 1, 9),
              (2, 2, 6), (3, 1, 6), (3, 2, 3), (3, 2, 5)]


@pytest.mark.parametrize("dim,order,n", ITER_CASES)
@pytest.mark.parametrize("kind", [F.SMOOTH_SSC, F.SMOOTH_PSC])
def test_block_iterate_parity(dim, order, n, kind):
    """One iterate forwards and backwards against the oracle (x random at Dirichlet nodes
    too); <ssc> also with b aliased to r, <psc> leaves r untouched."""
    say(f"block iterate dim={dim} order={order} n={n} kind={kind}")
    rng = np.random.default_rng(100 * dim + 10 * order + n)
    kap = np.exp(rng.normal(size=(n,) * dim))
    A, _ = lat.assemble(dim, order, n, kappa=kap)
    rp, ci, v = A.csr()
    N = A.nrows
    x0, b = rng.normal(size=N), rng.normal(size=N)
    ssc = kind == F.SMOOTH_SSC
    for backwards in (0, 1):
        r0 = np.zeros(N) if ssc else b - po.spmv(rp, ci, v, x0)
        xo, ro = po.block_iterate(rp, ci, v, dim, order, n, ssc, 0.8, x0, b, r0,
                                  backwards=bool(backwards))
        gx, gb, gr = GpuVector.from_array(x0), GpuVector.from_array(b), GpuVector.from_array(r0)
        check(lib().flgpu_block_iterate(A.h, kind, 0.8, backwards, gx.h, gb.h, gr.h))
        assert relerr(gx.numpy(), xo) < 1e-13
        if ssc:
            assert relerr(gr.numpy(), ro) < 1e-13
        else:
            assert np.array_equal(gr.numpy(), r0)
    if ssc:
        xo, ro = po.block_iterate(rp, ci, v, dim, order, n, True, 1.0, x0, None, b)
        gx, gw = GpuVector.from_array(x0), GpuVector.from_array(b)
        check(lib().flgpu_block_iterate(A.h, kind, 1.0, 0, gx.h, gw.h, gw.h))
        assert relerr(gx.numpy(), xo) < 1e-13 and relerr(gw.numpy(), ro) < 1e-13


@pytest.mark.parametrize("kind", [F.SMOOTH_SSC, F.SMOOTH_PSC])
def test_block_iterate_fp32(kind):
    """FP32 lattice matrices: inverses formed in FP64 and stored in FP32; the iterate agrees
    with the FP64 oracle to FP32 accuracy."""
    say(f"block iterate fp32 kind={kind}")
    dim, order, n = 3, 2, 4
    rng = np.random.default_rng(3)
    kap = np.exp(rng.normal(size=(n,) * dim))
    A, _ = lat.assemble(dim, order, n, kappa=kap, prec=F.F32)
    A64, _ = lat.assemble(dim, order, n, kappa=kap)
    rp, ci, v = A64.csr()
    N = A.nrows
    x0, b = rng.normal(size=N), rng.normal(size=N)
    ssc = kind == F.SMOOTH_SSC
    r0 = np.zeros(N) if ssc else b - po.spmv(rp, ci, v, x0)
    xo, _ = po.block_iterate(rp, ci, v, dim, order, n, ssc, 0.9, x0, b, r0)
    gx = GpuVector.from_array(x0, F.F32)
    gb, gr = GpuVector.from_array(b, F.F32), GpuVector.from_array(r0, F.F32)
    check(lib().flgpu_block_iterate(A.h, kind, 0.9, 0, gx.h, gb.h, gr.h))
    assert relerr(gx.numpy(), xo) < 2e-5


def test_block_iterate_rejects_non_lattice_matrices():
    say("block iterate argument checks")
    ctx = default_context()
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([0, 1], np.int32)
    v = np.array([2.0, 3.0])
    A = GpuSparseMatrix.from_csr(rp, ci, v, ctx=ctx)
    x, b = GpuVector(2), GpuVector(2)
    assert lib().flgpu_block_iterate(A.h, F.SMOOTH_SSC, 1.0, 0, x.h, b.h, b.h) < 0
    buf = ctypes.create_string_buffer(256)
    lib().flgpu_last_error(buf, 256)
    assert b"lattice" in buf.value
    L, _ = lat.assemble(2, 2, 4)
    y, c = GpuVector(L.nrows), GpuVector(L.nrows)
    assert lib().flgpu_block_iterate(L.h, 7, 1.0, 0, y.h, c.h, c.h) < 0
    assert lib().flgpu_block_iterate(L.h, F.SMOOTH_SSC, 1.0, 0, y.h, y.h, c.h) < 0


def _hier(dim, order, n0, levels, sm, pre, post, base, symmetric=False, kappa=None,
          prec=F.F64, galerkin=False, mode=F.ASM_F64):
    cs = it.GeometricCS(sm, pre, post, base, symmetric=symmetric, prec=prec, galerkin=galerkin)
    return cs, it.Hierarchy.lattice(cs, dim, order, n0, levels, kappa=kappa, mode=mode)


@pytest.mark.parametrize("dim,order,n0,levels", [(2, 2, 1, 4), (3, 2, 1, 3), (3, 1, 1, 4)])
@pytest.mark.parametrize("kind", [po.SMOOTH_SSC, po.SMOOTH_PSC])
@pytest.mark.parametrize("alias,steps", [(0, 1), (0, 2), (1, 2)])
def test_block_vcycle_parity(dim, order, n0, levels, kind, alias, steps):
    """One V-cycle from zero with block smoothers on every level, as <cg> calls its
    preconditioner (aliased or not), against the oracle."""
    say(f"block vcycle dim={dim} order={order} kind={kind} alias={alias} steps={steps}")
    omega = 1.0 if kind == po.SMOOTH_SSC else 0.2
    sm = it.GeometricSSC(omega) if kind == po.SMOOTH_SSC else it.GeometricPSC(omega)
    rng = np.random.default_rng(7)
    ntop = n0 << (levels - 1)
    kap = np.exp(rng.normal(size=(ntop,) * dim))
    cs, H = _hier(dim, order, n0, levels, sm, steps, steps, 0, symmetric=True, kappa=kap)
    oh = po.Hierarchy(dim, order, n0, levels, kappa=kap)
    prm = po.params(smoother=kind, omega=omega, pre_steps=steps, post_steps=steps,
                    base_level=0, symmetric=1, alias=alias)
    r = rng.uniform(-1, 1, oh.ndofs())
    r[po.boundary(dim, order, ntop)] = 0
    ref = oh.apply(prm, r)
    q, w = GpuVector(len(r)), GpuVector.from_array(r)
    H.apply(q, w, w if alias else GpuVector.from_array(r))
    assert relerr(q.numpy(), ref) < 1e-12


@pytest.mark.parametrize("galerkin", [False, True])
def test_ssc_preconditioned_cg_parity(galerkin):
    """<cg> with the V(1,1) <geometric-ssc> cycle femlisp selects for Q2 (geomg.lisp:58-71):
    the oracle's step count, defects and solution."""
    say(f"ssc cg galerkin={galerkin}")
    dim, order, n0, levels = 3, 2, 1, 3
    rng = np.random.default_rng(17)
    kap = np.exp(rng.normal(scale=1.5, size=(4, 4, 4)))
    cs, H = _hier(dim, order, n0, levels, None, 1, 1, 0, symmetric=True, kappa=kap,
                  galerkin=galerkin)
    oh = po.Hierarchy(dim, order, n0, levels, kappa=kap, galerkin=galerkin)
    b = rng.uniform(-1, 1, oh.ndofs())
    b[po.boundary(dim, order, 4)] = 0
    prm = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_SSC, omega=1.0, pre_steps=1,
                    post_steps=1, base_level=0, symmetric=1, tol=1e-10, maxsteps=100)
    st, xo, so, dno = oh.solve(prm, b)
    assert st == 0
    bb = it.solve(it.LinearSolver(it.CG(cs), success_if=("<", "reduction", 1e-10),
                                  failure_if=(">", "step", 100)),
                  {"matrix": H.top_matrix(), "rhs": GpuVector.from_array(b), "hierarchy": H})
    assert bb["status"] == "success" and bb["steps"] == so
    assert relerr(bb["solution"].numpy(), xo) < 1e-10
    assert np.allclose(bb["defnorms"], dno, rtol=1e-8)


def test_ssc_graph_replays_and_reassembly():
    """The captured SSC cycle replays bit for bit; a new coefficient refreshes the block
    inverses (the cycle then equals a hierarchy built for that coefficient)."""
    say("ssc graphs + reassemble")
    rng = np.random.default_rng(23)
    dim, order, n0, levels = 3, 2, 1, 3
    k1 = np.exp(rng.normal(size=(4, 4, 4)))
    k2 = np.exp(rng.normal(size=(4, 4, 4)))
    cs, H = _hier(dim, order, n0, levels, it.GeometricSSC(), 1, 1, 0, symmetric=True, kappa=k1)
    N = H.top_matrix().nrows
    r = rng.uniform(-1, 1, N)
    r[po.boundary(dim, order, 4)] = 0
    H.use_graphs(False)
    qe = GpuVector(N)
    H.apply(qe, GpuVector.from_array(r), GpuVector.from_array(r))
    ref = qe.numpy()
    H.use_graphs(True)
    for rep in range(3):
        q = GpuVector(N)
        H.apply(q, GpuVector.from_array(r), GpuVector.from_array(r))
        assert np.array_equal(q.numpy(), ref), rep
    H.reassemble(k2)
    _, H2 = _hier(dim, order, n0, levels, it.GeometricSSC(), 1, 1, 0, symmetric=True, kappa=k2)
    q, q2 = GpuVector(N), GpuVector(N)
    H.apply(q, GpuVector.from_array(r), GpuVector.from_array(r))
    H2.apply(q2, GpuVector.from_array(r), GpuVector.from_array(r))
    assert np.array_equal(q.numpy(), q2.numpy())


def test_block_iterate_on_uploaded_lattice_matrix():
    """A caller's CSR in lattice numbering (as femlisp's sparse-matrix->ccs in
    sort-lexicographically order hands it over) takes block smoothing once
    flgpu_lattice_colouring has recorded its lattice; the iterate equals the one on the
    matrix assembled on the GPU."""
    say("block iterate on an uploaded matrix")
    dim, order, n = 3, 2, 3
    rng = np.random.default_rng(41)
    A, _ = lat.assemble(dim, order, n, kappa=np.exp(rng.normal(size=(n,) * dim)))
    rp, ci, v = A.csr()
    U = GpuSparseMatrix.from_csr(rp, ci, v)
    N = A.nrows
    x0, b = rng.normal(size=N), rng.normal(size=N)
    x, bb, r = GpuVector.from_array(x0), GpuVector.from_array(b), GpuVector(N)
    assert lib().flgpu_block_iterate(U.h, F.SMOOTH_SSC, 1.0, 0, x.h, bb.h, r.h) < 0
    assert lib().flgpu_lattice_colouring(U.h, dim, order, n + 1) < 0  # wrong size
    check(lib().flgpu_lattice_colouring(U.h, dim, order, n))
    check(lib().flgpu_block_iterate(U.h, F.SMOOTH_SSC, 1.0, 0, x.h, bb.h, r.h))
    x2, r2 = GpuVector.from_array(x0), GpuVector(N)
    check(lib().flgpu_block_iterate(A.h, F.SMOOTH_SSC, 1.0, 0, x2.h, bb.h, r2.h))
    assert np.array_equal(x.numpy(), x2.numpy())


def test_ssc_mixed_fcg_eager_matches_graphs():
    """FP32 <geometric-ssc> cycles inside FP64 FCG, launched eagerly and replayed from HIP
    graphs: the same steps and the same solution, to 1e-10 in the FP64 true residual."""
    say("ssc mixed fcg eager vs graphs")
    dim, order, n0, levels = 3, 2, 1, 4
    rng = np.random.default_rng(43)
    kap = np.exp(rng.normal(scale=1.5, size=(8, 8, 8)))
    A, b = lat.assemble(dim, order, 8, kappa=kap)
    out = []
    for graphs in (False, True):
        cs, H = _hier(dim, order, n0, levels, it.GeometricSSC(), 1, 1, 0, symmetric=True,
                      kappa=kap, prec=F.F32, mode=F.ASM_MFMA)
        H.use_graphs(graphs)
        bb = it.solve(it.LinearSolver(it.FCG(cs), success_if=("<", "reduction", 1e-10),
                                      failure_if=(">", "step", 100), true_residual=True),
                      {"matrix": A, "rhs": b, "hierarchy": H})
        assert bb["status"] == "success"
        assert bb["true_residual"] / np.linalg.norm(b.numpy()) < 1e-10
        out.append((bb["steps"], bb["solution"].numpy()))
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])
