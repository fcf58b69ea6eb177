"""GPU parity against the oracle: pattern/indexing bit-exact, assembly and
SpMV within 1e-13 relative, V-cycle operator within 1e-12, converged
solutions within 1e-10 and identical iteration counts where the ordering is
defined (Jacobi; <sor> in colour order).  All calls go through the C-ABI."""
import json
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from gnu_femlisp_amd import flgpu as F
from gnu_femlisp_amd import iteration as it
from gnu_femlisp_amd import lattice as lat
from gnu_femlisp_amd.matlisp import (GpuSparseMatrix, GpuVector, compute_residual,
                                     default_context, dot, gemm, norm)
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rowrel(v_gpu, v_ref, rp):
    """max over rows of |row error|_inf / |reference row|_inf"""
    worst = 0.0
    for r in range(len(rp) - 1):
        a, b = rp[r], rp[r + 1]
        den = np.abs(v_ref[a:b]).max()
        if den > 0:
            worst = max(worst, np.abs(v_gpu[a:b] - v_ref[a:b]).max() / den)
    return worst


def relerr(x, y):
    d = np.linalg.norm(y)
    return np.linalg.norm(x - y) / (d if d > 0 else 1.0)


CASES = [(1, 1, 8), (2, 1, 16), (3, 1, 8), (1, 2, 5), (2, 2, 6), (3, 2, 4), (3, 1, 16)]


@pytest.mark.parametrize("dim,order,n", CASES)
@pytest.mark.parametrize("mode", [F.ASM_F64, F.ASM_MFMA])
def test_assembly_parity(dim, order, n, mode):
    say(f"assembly dim={dim} order={order} n={n} mode={mode}")
    rng = np.random.default_rng(7)
    nk = 2 * n
    kap = np.exp(rng.normal(scale=1.5, size=(nk,) * dim))
    A, b = lat.assemble(dim, order, n, kap, f=1.0, prec=F.F64, mode=mode)
    rp, ci, v = A.csr()
    orp, oci, ov, ob = po.assemble(dim, order, n, kap, 1.0)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)  # bit-exact pattern
    assert rowrel(v, ov, orp) < 1e-13
    assert relerr(b.numpy(), ob) < 1e-13


@pytest.mark.parametrize("dim,order,n", [(2, 1, 8), (3, 1, 4), (3, 2, 3)])
def test_element_kernel_parity(dim, order, n):
    say(f"element dim={dim} order={order} n={n}")
    rng = np.random.default_rng(11)
    kap = np.exp(rng.normal(size=(n,) * dim))
    for mode in (F.ASM_F64, F.ASM_MFMA):
        E = lat.element_matrices(dim, order, n, kap, prec=F.F64, mode=mode)
        for c in (0, n ** dim // 2, n ** dim - 1):
            kq = np.full(((order + 1) ** dim), kap.ravel()[c])
            K, _ = po.element_matrix(dim, order, 1.0 / n, kq)
            assert np.abs(E[c] - K).max() <= 1e-13 * np.abs(K).max()
    # BF16 MFMA on operands split in three BF16 parts (the six products down to 2^-16 of the
    # leading one) with FP32 accumulation: about 24 significant bits per product, as FP32;
    # stated bound 2e-6 of the largest entry (FP32 rounding of the entries alone is 6e-8)
    E32 = lat.element_matrices(dim, order, n, kap, prec=F.F32, mode=F.ASM_MFMA)
    E64 = lat.element_matrices(dim, order, n, kap, prec=F.F64, mode=F.ASM_F64)
    err = np.abs(E32 - E64).max() / np.abs(E64).max()
    say(f"  BF16x3 element error {err:.2e} of the largest entry")
    assert err <= 2e-6
    # one coefficient value per cell: the contraction is S_0 K (quadrature point 0 against the
    # reference stiffness K, the others against the zero variation), so every entry is within a
    # few FP32 roundings of its own size, and the entries where K vanishes keep the FP64 sum's
    # residue (about 1e-17 of the largest) instead of the FP32 accumulator's cancellation error
    emax = np.abs(E64).max(axis=(1, 2), keepdims=True)
    big = np.abs(E64) > 1e-12 * emax
    zero = np.abs(E32) / emax
    assert np.all(zero[~big] <= 1e-15), zero[~big].max()
    rel = (np.abs(E32 - E64)[big] / np.abs(E64)[big]).max()
    say(f"  per-entry relative error {rel:.2e}; {np.count_nonzero(~big)} vanishing entries, "
        f"at most {zero[~big].max() if (~big).any() else 0.0:.1e} of the largest")
    assert rel <= 1e-6
    # coefficient varying inside the cells (4 samples per cell and axis): the bound above
    kv = np.exp(rng.normal(size=(4 * n,) * dim))
    E32 = lat.element_matrices(dim, order, n, kv, prec=F.F32, mode=F.ASM_MFMA)
    E64 = lat.element_matrices(dim, order, n, kv, prec=F.F64, mode=F.ASM_F64)
    err = np.abs(E32 - E64).max() / np.abs(E64).max()
    say(f"  BF16x3 element error, varying coefficient: {err:.2e} of the largest entry")
    assert err <= 2e-6


@pytest.mark.parametrize("dim,order,nc", [(1, 1, 4), (2, 1, 4), (3, 1, 4), (2, 2, 3), (3, 2, 2)])
def test_interpolation_parity(dim, order, nc):
    say(f"interp dim={dim} order={order} nc={nc}")
    P = lat.interpolation(dim, order, nc)
    rp, ci, v = P.csr()
    orp, oci, ov = po.interpolation(dim, order, nc)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    assert np.array_equal(v, ov)  # dyadic weights: exact


def test_spmv_parity_random_csr():
    say("spmv random")
    rng = np.random.default_rng(5)
    for n, dens in ((1, 1.0), (7, 0.5), (300, 0.03), (5000, 0.002)):
        M = sp.random(n, n, density=dens, random_state=rng, format="csr") + sp.eye(n, format="csr")
        M.sort_indices()
        A = GpuSparseMatrix.from_scipy(M)
        y = rng.normal(size=n)
        x0 = rng.normal(size=n)
        gy, gx = GpuVector.from_array(y), GpuVector.from_array(x0)
        gemm(0.7, A, gy, -1.3, gx)
        ref = po.spmv(M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data, y, 0.7, -1.3, x0)
        assert relerr(gx.numpy(), ref) < 1e-13
        # gemm-tn!
        gz = GpuVector.from_array(x0)
        gemm(1.0, A, gy, 0.0, gz, job="tn")
        assert relerr(gz.numpy(), M.T @ y) < 1e-13


def test_empty_and_ragged_rows():
    say("spmv ragged")
    rp = np.array([0, 0, 3, 3, 4, 4], np.int32)
    ci = np.array([0, 2, 4, 1], np.int32)
    v = np.array([1.0, 2.0, 3.0, 4.0])
    A = GpuSparseMatrix.from_csr(rp, ci, v, ncols=5)
    y = GpuVector.from_array(np.arange(1.0, 6.0))
    x = GpuVector.from_array(np.full(5, 9.0))
    gemm(1.0, A, y, 0.0, x)
    assert np.array_equal(x.numpy(), [0.0, 1 + 6 + 15, 0.0, 8.0, 0.0])


def test_csr_upload_rejects_bad_input():
    rp = np.array([0, 2, 1], np.int32)
    with pytest.raises(F.FlgpuError):
        GpuSparseMatrix.from_csr(rp, np.array([0, 1], np.int32), np.ones(2))
    with pytest.raises(F.FlgpuError):
        GpuSparseMatrix.from_csr(np.array([0, 2], np.int32), np.array([1, 0], np.int32), np.ones(2))


@pytest.mark.parametrize("dim,order,n", [(2, 1, 32), (3, 1, 16), (3, 2, 6)])
def test_lattice_spmv_parity(dim, order, n):
    say(f"lattice spmv dim={dim} order={order} n={n}")
    A, b = lat.assemble(dim, order, n)
    rp, ci, v = A.csr()
    y = np.random.default_rng(2).uniform(-1, 1, A.ncols)
    gy = GpuVector.from_array(y)
    gx = A.make_image_vector()
    gemm(1.0, A, gy, 0.0, gx)
    assert relerr(gx.numpy(), po.spmv(rp, ci, v, y)) < 1e-13
    # compute-residual with b and r the same vector: r <- r - A x
    gr = GpuVector.from_array(b.numpy())
    compute_residual(A, gy, gr, gr)
    assert relerr(gr.numpy(), b.numpy() - po.spmv(rp, ci, v, y)) < 1e-13


@pytest.mark.parametrize("form", [F.STORAGE_CELL, F.STORAGE_STENCIL])
@pytest.mark.parametrize("prec", [F.F64, F.F32])
def test_lattice_storage_matches_csr(prec, form):
    """3D Q1 lattice matrices: the operator applied from the per-cell factors
    or from the 27-point stencil storage equals the SELL-streamed CSR (SpMV,
    compute-residual and <sor> sweeps both ways), and the oracle's SpMV.  The
    stencil storage sums the same entries in the same order as the CSR row."""
    say(f"storage form={form} prec={prec}")
    n = 12
    rng = np.random.default_rng(17)
    kap = np.exp(rng.normal(scale=1.5, size=(n, n, n)))
    A, b = lat.assemble(3, 1, n, kap, prec=prec)
    B, _ = lat.assemble(3, 1, n, kap, prec=prec)
    assert A.storage(form) == form and B.storage(F.STORAGE_SELL) == F.STORAGE_SELL
    tol = 1e-13 if prec == F.F64 else 2e-6
    if form == F.STORAGE_STENCIL:
        tol = 1e-14 if prec == F.F64 else 1e-6
    y = rng.uniform(-1, 1, A.ncols)
    ya, yb = GpuVector.from_array(y, prec), GpuVector.from_array(y, prec)
    xa, xb = GpuVector(A.nrows, prec), GpuVector(A.nrows, prec)
    gemm(1.0, A, ya, 0.0, xa)
    gemm(1.0, B, yb, 0.0, xb)
    assert relerr(xa.numpy(), xb.numpy()) < tol
    rp, ci, v, _ = po.assemble(3, 1, n, kap)
    if prec == F.F64:
        assert relerr(xa.numpy(), po.spmv(rp, ci, v, y)) < 1e-13
    ra, rb = GpuVector.from_array(b.numpy(), prec), GpuVector.from_array(b.numpy(), prec)
    compute_residual(A, ya, ra, ra)
    compute_residual(B, yb, rb, rb)
    assert relerr(ra.numpy(), rb.numpy()) < tol
    from gnu_femlisp_amd.flgpu import check, lib
    bs = GpuVector.from_array(b.numpy(), prec)
    for bw in (0, 1):
        xs = [GpuVector.from_array(y, prec) for _ in range(2)]
        rs = [GpuVector(A.nrows, prec) for _ in range(2)]
        for M, xv, rv in ((A, xs[0], rs[0]), (B, xs[1], rs[1])):
            check(lib().flgpu_sor_iterate(M.h, 1.1, bw, xv.h, bs.h, rv.h))
        assert relerr(xs[0].numpy(), xs[1].numpy()) < tol
        assert relerr(rs[0].numpy(), rs[1].numpy()) < 10 * tol


def test_vector_ops():
    say("vector ops")
    rng = np.random.default_rng(9)
    for n in (0, 1, 1000, 300001):
        a, c = rng.normal(size=n), rng.normal(size=n)
        x, y = GpuVector.from_array(a), GpuVector.from_array(c)
        assert abs(dot(x, y) - a @ c) <= 1e-12 * max(1.0, np.abs(a) @ np.abs(c))
        assert abs(norm(x) - np.linalg.norm(a)) <= 1e-12 * max(1.0, np.linalg.norm(a))
        from gnu_femlisp_amd.matlisp import axpy, scal
        axpy(0.5, x, y)
        assert np.allclose(y.numpy(), 0.5 * a + c, rtol=1e-15, atol=1e-15)
        scal(-2.0, y)
        assert np.allclose(y.numpy(), -2 * (0.5 * a + c), rtol=1e-15, atol=1e-15)
    # dot is reproducible bit for bit
    x = GpuVector.from_array(rng.normal(size=100003))
    assert len({dot(x, x) for _ in range(5)}) == 1


def test_jacobi_and_sor_iterates():
    say("smoother iterates")
    dim, order, n = 2, 2, 5
    A, b = lat.assemble(dim, order, n)
    rp, ci, v = A.csr()
    M = sp.csr_matrix((v, ci, rp))
    d = M.diagonal()
    rng = np.random.default_rng(4)
    x0 = rng.normal(size=A.nrows)
    r0 = rng.normal(size=A.nrows)
    x, r = GpuVector.from_array(x0), GpuVector.from_array(r0)
    from gnu_femlisp_amd.flgpu import check, lib
    check(lib().flgpu_jacobi_iterate(A.h, 0.6, x.h, r.h))
    rr = r0 / d
    assert np.allclose(r.numpy(), rr, rtol=1e-15) and np.allclose(x.numpy(), x0 + 0.6 * rr, rtol=1e-15)
    # <sor> in (colour, index) order, restated row by row on the host
    lat.colour(A, dim, order, n)
    colour = lat.lattice_colours(dim, order, n)
    order_rows = np.lexsort((np.arange(A.nrows), colour))
    bb = rng.normal(size=A.nrows)
    xs = x0.copy()
    rs = np.zeros(A.nrows)
    for i in order_rows:
        corr = bb[i]
        for e in range(rp[i], rp[i + 1]):
            corr += (-xs[ci[e]]) * v[e]
        corr /= d[i]
        rs[i] = corr
        xs[i] += 1.2 * corr
    gx, gb, gr = GpuVector.from_array(x0), GpuVector.from_array(bb), GpuVector(A.nrows)
    check(lib().flgpu_sor_iterate(A.h, 1.2, 0, gx.h, gb.h, gr.h))
    assert relerr(gx.numpy(), xs) < 1e-13 and relerr(gr.numpy(), rs) < 1e-13


@pytest.mark.parametrize("form", [F.STORAGE_CELL, F.STORAGE_STENCIL, F.STORAGE_SELL])
@pytest.mark.parametrize("prec", [F.F64, F.F32])
@pytest.mark.parametrize("steps", [1, 4])
def test_jacobi_steps_match_product_iterator(form, prec, steps):
    """flgpu_jacobi_steps (residual fused into each step, out of place) equals the
    product-iterator of <jacobi> done the long way: r = b - A x, then the
    in-place jacobi iterate, per step; and, in FP64, a host restatement."""
    say(f"jacobi steps form={form} prec={prec} steps={steps}")
    from gnu_femlisp_amd.flgpu import check, lib
    n = 12
    rng = np.random.default_rng(23)
    kap = np.exp(rng.normal(scale=1.5, size=(n, n, n)))
    A, b = lat.assemble(3, 1, n, kap, prec=prec)
    assert A.storage(form) == form
    x0 = rng.uniform(-1, 1, A.nrows)
    x, tmp = GpuVector.from_array(x0, prec), GpuVector(A.nrows, prec)
    check(lib().flgpu_jacobi_steps(A.h, 0.7, steps, x.h, b.h, tmp.h))
    y, r = GpuVector.from_array(x0, prec), GpuVector(A.nrows, prec)
    for _ in range(steps):
        compute_residual(A, y, b, r)
        check(lib().flgpu_jacobi_iterate(A.h, 0.7, y.h, r.h))
    assert relerr(x.numpy(), y.numpy()) < (1e-14 if prec == F.F64 else 1e-6)
    if prec == F.F64:
        rp, ci, v = A.csr()
        M = sp.csr_matrix((v, ci, rp))
        xh, bh, d = x0.copy(), b.numpy(), M.diagonal()
        for _ in range(steps):
            xh = xh + 0.7 * (bh - M @ xh) / d
        assert relerr(x.numpy(), xh) < 1e-13
    with pytest.raises(F.FlgpuError):
        check(lib().flgpu_jacobi_steps(A.h, 0.7, 1, x.h, b.h, x.h))


@pytest.mark.parametrize("prec", [F.F64, F.F32])
@pytest.mark.parametrize("n", [6, 16, 32])
def test_jacobi_multi_step_passes_match_single_steps(prec, n):
    """Small stencil-form levels take several <jacobi> steps per launch (the iterate staged in
    LDS, a halo recomputed per block); every step count equals the same steps taken one
    launch at a time, bit for bit, and the vectors' roles (x in, x out) are kept."""
    say(f"jacobi multi-step n={n} prec={prec}")
    from gnu_femlisp_amd.flgpu import check, lib
    rng = np.random.default_rng(60 + n)
    kap = np.exp(rng.normal(scale=1.5, size=(n, n, n)))
    A, b = lat.assemble(3, 1, n, kap, prec=prec)
    assert A.storage(F.STORAGE_STENCIL) == F.STORAGE_STENCIL
    x0 = rng.uniform(-1, 1, A.nrows)
    for steps in range(1, 7):
        x, tmp = GpuVector.from_array(x0, prec), GpuVector(A.nrows, prec)
        check(lib().flgpu_jacobi_steps(A.h, 0.7, steps, x.h, b.h, tmp.h))
        y, t2 = GpuVector.from_array(x0, prec), GpuVector(A.nrows, prec)
        for _ in range(steps):
            check(lib().flgpu_jacobi_steps(A.h, 0.7, 1, y.h, b.h, t2.h))
        assert np.array_equal(x.numpy(), y.numpy()), steps


def _gpu_hier(dim, order, n0, levels, smoother, omega, pre, post, base, symmetric=False,
              kappa=None, prec=F.F64, mode=F.ASM_F64, gamma=1, transfer="geometric",
              additive=False, fas=False):
    sm = it.Jacobi(omega) if smoother == po.SMOOTH_JACOBI else it.SOR(omega)
    cls = it.GeometricFAS if fas else it.GeometricCS
    cs = cls(sm, pre, post, base, gamma=gamma, symmetric=symmetric, prec=prec,
             transfer=transfer, combination="additive" if additive else "multiplicative")
    return cs, it.Hierarchy.lattice(cs, dim, order, n0, levels, kappa=kappa, mode=mode)


@pytest.mark.parametrize("prec", [F.F64, F.F32])
def test_lattice_projection_matches_oracle(prec):
    """flgpu_lattice_projection (femlisp's projection-matrix: injection of the nodal
    functionals) equals the oracle's matrix bit for bit, and R I = Id on the GPU."""
    say(f"lattice projection prec={prec}")
    from gnu_femlisp_amd.flgpu import check, lib
    import ctypes
    for dim, order, nc in ((1, 1, 4), (2, 1, 3), (2, 2, 2), (3, 1, 4), (3, 2, 2)):
        h = ctypes.c_void_p()
        check(lib().flgpu_lattice_projection(default_context().h, dim, order, nc, prec,
                                             ctypes.byref(h)))
        R = GpuSparseMatrix(h, default_context())
        rp, ci, v = R.csr()
        orp, oci, ov = po.projection(dim, order, nc)
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(v, ov)
        P = lat.interpolation(dim, order, nc, prec=prec)
        xc = np.random.default_rng(nc).uniform(-1, 1, P.ncols)
        xc[po.boundary(dim, order, nc)] = 0  # Dirichlet columns of P are zero
        y = gemm(1.0, P, GpuVector.from_array(xc, prec), 0.0, GpuVector(P.nrows, prec))
        back = gemm(1.0, R, y, 0.0, GpuVector(P.ncols, prec)).numpy()
        assert np.array_equal(back, xc.astype(np.float32) if prec == F.F32 else xc)


@pytest.mark.parametrize("smoother", [po.SMOOTH_JACOBI, po.SMOOTH_SOR])
@pytest.mark.parametrize("gamma,additive", [(1, 0), (2, 0), (1, 1)])
@pytest.mark.parametrize("dim,order,n0,levels", [(3, 1, 1, 4), (2, 2, 1, 3)])
def test_fas_cycle_parity(smoother, gamma, additive, dim, order, n0, levels):
    """<geometric-fas> on the GPU: one cycle from zero against the oracle's FAS cycle; for the
    multiplicative combination also against the GPU's own correction scheme, which FAS equals
    up to rounding for a linear operator; the hierarchy's projections are the oracle's."""
    say(f"fas cycle smoother={smoother} gamma={gamma} additive={additive} dim={dim} "
        f"order={order}")
    omega = 0.6 if smoother == po.SMOOTH_JACOBI else 1.0
    base = 1 if levels > 3 else 0
    kw = dict(gamma=gamma, additive=bool(additive))
    cs, H = _gpu_hier(dim, order, n0, levels, smoother, omega, 2, 2, base, fas=True, **kw)
    oh = po.Hierarchy(dim, order, n0, levels)
    for l in range(levels - 1):
        rp, ci, v = H.level_matrix(l, 2).csr()
        orp, oci, ov = oh.projection(l)
        assert np.array_equal(ci, oci) and np.array_equal(v, ov)
    prm = po.params(smoother=smoother, omega=omega, ordering=po.ORDER_COLOUR, pre_steps=2,
                    post_steps=2, base_level=base, gamma=gamma, additive=additive, fas=1)
    rng = np.random.default_rng(21)
    ntop = n0 << (levels - 1)
    r = rng.uniform(-1, 1, oh.ndofs())
    r[po.boundary(dim, order, ntop)] = 0
    ref = oh.apply(prm, r)
    q = GpuVector(len(r))
    H.apply(q, GpuVector.from_array(r), GpuVector.from_array(r))
    assert relerr(q.numpy(), ref) < 1e-12
    if not additive:
        _, Hc = _gpu_hier(dim, order, n0, levels, smoother, omega, 2, 2, base, **kw)
        qc = GpuVector(len(r))
        Hc.apply(qc, GpuVector.from_array(r), GpuVector.from_array(r))
        assert relerr(q.numpy(), qc.numpy()) < 1e-12


def test_fas_preconditioned_cg_parity():
    """FAS as <cg>'s preconditioner: the oracle's iteration count, defects and solution."""
    say("fas cg")
    dim, order, n0, levels = 3, 1, 1, 4
    rng = np.random.default_rng(5)
    kap = np.exp(rng.normal(scale=1.0, size=(8, 8, 8)))
    cs, H = _gpu_hier(dim, order, n0, levels, po.SMOOTH_SOR, 1.0, 2, 2, 0, symmetric=True,
                      kappa=kap, fas=True)
    oh = po.Hierarchy(dim, order, n0, levels, kappa=kap)
    b = rng.uniform(-1, 1, oh.ndofs())
    b[po.boundary(dim, order, 8)] = 0
    prm = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_SOR, omega=1.0,
                    ordering=po.ORDER_COLOUR, pre_steps=2, post_steps=2, base_level=0,
                    symmetric=1, tol=1e-10, maxsteps=200, fas=1)
    st, xo, so, dno = oh.solve(prm, b)
    assert st == 0
    bb = it.solve(it.LinearSolver(it.CG(cs), success_if=("<", "reduction", 1e-10),
                                  failure_if=(">", "step", 200)),
                  {"matrix": H.top_matrix(), "rhs": GpuVector.from_array(b), "hierarchy": H})
    assert bb["status"] == "success" and bb["steps"] == so
    assert relerr(bb["solution"].numpy(), xo) < 1e-10
    assert np.allclose(bb["defnorms"], dno, rtol=1e-8)


@pytest.mark.parametrize("alias", [0, 1])
@pytest.mark.parametrize("smoother", [po.SMOOTH_JACOBI, po.SMOOTH_SOR])
def test_vcycle_operator_parity(alias, smoother):
    say(f"vcycle apply alias={alias} smoother={smoother}")
    dim, order, n0, levels = 3, 1, 1, 4
    omega = 0.6 if smoother == po.SMOOTH_JACOBI else 1.0
    cs, H = _gpu_hier(dim, order, n0, levels, smoother, omega, 2, 2, 1)
    oh = po.Hierarchy(dim, order, n0, levels)
    prm = po.params(smoother=smoother, omega=omega, ordering=po.ORDER_COLOUR, pre_steps=2,
                    post_steps=2, base_level=1, alias=alias)
    rng = np.random.default_rng(8)
    r = rng.uniform(-1, 1, oh.ndofs())
    r[po.boundary(dim, order, 8)] = 0
    ref = oh.apply(prm, r)
    q = GpuVector(len(r))
    w = GpuVector.from_array(r)
    if alias:
        H.apply(q, w, w)
    else:
        H.apply(q, w, GpuVector.from_array(r))
    assert relerr(q.numpy(), ref) < 1e-12


@pytest.mark.parametrize("prec", [F.F64, F.F32])
def test_vcycle_graph_replays_match_eager(prec):
    """The captured cycle, replayed several times with eager work and
    synchronisations in between, equals the eagerly launched cycle bit for
    bit on every replay (FP32 hierarchies applied to FP64 vectors too)."""
    say(f"graph replays prec={prec}")
    rng = np.random.default_rng(31)
    kap = np.exp(rng.normal(size=(16, 16, 16)))
    cs, H = _gpu_hier(3, 1, 1, 5, po.SMOOTH_SOR, 1.0, 2, 2, 1, symmetric=True, kappa=kap,
                      prec=prec, mode=F.ASM_MFMA if prec == F.F32 else F.ASM_F64)
    N = H.top_matrix().nrows
    r = rng.uniform(-1, 1, N)
    r[po.boundary(3, 1, 16)] = 0
    H.use_graphs(False)
    qe, w1, w2 = GpuVector(N), GpuVector.from_array(r), GpuVector.from_array(r)
    H.apply(qe, w1, w2)
    ref = qe.numpy()
    H.use_graphs(True)
    q, w1, w2 = GpuVector(N), GpuVector(N), GpuVector(N)
    other = GpuVector.from_array(rng.normal(size=N))
    from gnu_femlisp_amd.matlisp import axpy, fill
    for rep in range(4):
        fill(q, 0.0)
        w1.upload(r)
        w2.upload(r)
        H.apply(q, w1, w2)
        assert np.array_equal(q.numpy(), ref), rep
        axpy(0.5, w1, other)  # eager work between replays
        norm(other)


def test_known_answer_vcycle_rate_1d():
    """mg-cdr.lisp:85-87 on the GPU: CR 0.39407..."""
    say("known answer 1d")
    cs, H = _gpu_hier(1, 1, 1, 4, po.SMOOTH_JACOBI, 0.5, 1, 0, 2)
    A = H.top_matrix()
    bb = it.solve(it.LinearSolver(cs, success_if=("<", "defnorm", 1e-12), failure_if=(">", "step", 10)),
                  {"matrix": A, "rhs": H.rhs, "hierarchy": H})
    assert bb["steps"] == 11
    cr = bb["report"]["convergence-rate"]
    assert cr < 0.3941 and abs(cr - 0.39407) < 1e-5


def test_known_answer_tridiagonal_cg():
    say("known answer 3x3")
    k = json.load(open(os.path.join(GOLDEN, "femlisp_known_answers.json")))["tridiagonal_3x3"]
    A = GpuSparseMatrix.from_scipy(sp.csr_matrix(np.array(k["matrix"])))
    b = GpuVector.from_array(k["rhs"])
    bb = it.solve(it.LinearSolver(it.CG(), success_if=("<", "defnorm", 1e-12)),
                  {"matrix": A, "rhs": b})
    assert np.allclose(bb["solution"].numpy(), k["solution"], atol=1e-12)


SOLVE_CASES = [
    # dim, order, n0, levels, smoother, omega, pre, post, base, outer, symmetric
    (2, 1, 1, 6, po.SMOOTH_JACOBI, 0.6, 2, 2, 0, po.OUTER_CG, 0),   # config 1
    (3, 1, 1, 5, po.SMOOTH_SOR, 1.0, 2, 2, 0, po.OUTER_CG, 1),
    (3, 1, 1, 5, po.SMOOTH_SOR, 1.0, 2, 2, 1, po.OUTER_MG, 0),
    (3, 2, 1, 3, po.SMOOTH_JACOBI, 0.6, 2, 2, 0, po.OUTER_FCG, 0),
    (2, 2, 1, 4, po.SMOOTH_SOR, 1.0, 1, 1, 1, po.OUTER_CG, 1),
]


@pytest.mark.parametrize("case", SOLVE_CASES)
def test_solve_parity(case):
    dim, order, n0, levels, smoother, omega, pre, post, base, outer, sym = case
    say(f"solve {case}")
    rng = np.random.default_rng(12)
    ntop = n0 << (levels - 1)
    kap = np.exp(rng.normal(scale=1.0, size=(ntop,) * dim))
    oh = po.Hierarchy(dim, order, n0, levels, kappa=kap)
    b = rng.uniform(-1, 1, oh.ndofs())
    b[po.boundary(dim, order, ntop)] = 0
    prm = po.params(outer=outer, smoother=smoother, omega=omega, ordering=po.ORDER_COLOUR,
                    pre_steps=pre, post_steps=post, base_level=base, symmetric=sym, alias=0,
                    tol=1e-10, maxsteps=200)
    st, xo, so, dno = oh.solve(prm, b)
    assert st == 0
    cs, H = _gpu_hier(dim, order, n0, levels, smoother, omega, pre, post, base, symmetric=bool(sym),
                      kappa=kap)
    A = H.top_matrix()
    itn = {po.OUTER_MG: cs, po.OUTER_CG: it.CG(cs), po.OUTER_FCG: it.FCG(cs)}[outer]
    bb = it.solve(it.LinearSolver(itn, success_if=("<", "reduction", 1e-10),
                                  failure_if=(">", "step", 200)),
                  {"matrix": A, "rhs": GpuVector.from_array(b), "hierarchy": H})
    assert bb["status"] == "success"
    assert bb["steps"] == so
    assert relerr(bb["solution"].numpy(), xo) < 1e-10
    assert np.allclose(bb["defnorms"], dno, rtol=1e-8)


@pytest.mark.parametrize("smoother,omega,steps,base,outer,gamma,additive", [
    (po.SMOOTH_SOR, 1.0, 2, 0, po.OUTER_CG, 2, 0),
    (po.SMOOTH_JACOBI, 0.7, 4, 1, po.OUTER_FCG, 2, 0),
    (po.SMOOTH_JACOBI, 0.7, 2, 1, po.OUTER_MG, 0, 0),
    (po.SMOOTH_SOR, 1.0, 1, 1, po.OUTER_MG, 3, 0),
    (po.SMOOTH_JACOBI, 0.5, 2, 0, po.OUTER_FCG, 1, 1),
    (po.SMOOTH_SOR, 1.0, 1, 1, po.OUTER_FCG, 2, 1),
])
def test_gamma_solve_parity(smoother, omega, steps, base, outer, gamma, additive):
    """W-cycles (gamma 2, 3), smoothing only (gamma 0) and the :additive
    combination: the oracle's steps, defect history and solution."""
    say(f"gamma solve smoother={smoother} gamma={gamma} outer={outer} additive={additive}")
    dim, order, n0, levels = 3, 1, 1, 5
    rng = np.random.default_rng(29)
    ntop = n0 << (levels - 1)
    kap = np.exp(rng.normal(scale=1.0, size=(ntop,) * dim))
    oh = po.Hierarchy(dim, order, n0, levels, kappa=kap)
    b = rng.uniform(-1, 1, oh.ndofs())
    b[po.boundary(dim, order, ntop)] = 0
    sym = int(smoother == po.SMOOTH_SOR)
    maxsteps = 400 if gamma == 0 else 100
    prm = po.params(outer=outer, smoother=smoother, omega=omega, ordering=po.ORDER_COLOUR,
                    pre_steps=steps, post_steps=steps, base_level=base, symmetric=sym, alias=0,
                    tol=1e-8 if gamma == 0 else 1e-10, maxsteps=maxsteps, gamma=gamma,
                    additive=additive)
    st, xo, so, dno = oh.solve(prm, b)
    cs, H = _gpu_hier(dim, order, n0, levels, smoother, omega, steps, steps, base,
                      symmetric=bool(sym), kappa=kap, gamma=gamma, additive=bool(additive))
    itn = {po.OUTER_MG: cs, po.OUTER_CG: it.CG(cs), po.OUTER_FCG: it.FCG(cs)}[outer]
    bb = it.solve(it.LinearSolver(itn, success_if=("<", "reduction", prm.tol),
                                  failure_if=(">", "step", maxsteps)),
                  {"matrix": H.top_matrix(), "rhs": GpuVector.from_array(b), "hierarchy": H})
    assert bb["steps"] == so and (bb["status"] == "success") == (st == 0)
    assert np.allclose(bb["defnorms"], dno, rtol=1e-8)
    assert relerr(bb["solution"].numpy(), xo) < 1e-9


@pytest.mark.parametrize("dim,order,levels", [(3, 1, 4), (2, 2, 3), (2, 1, 5)])
def test_galerkin_hierarchy_parity(dim, order, levels):
    say(f"galerkin dim={dim} order={order} levels={levels}")
    ntop = 1 << (levels - 1)
    kap = np.exp(np.random.default_rng(3).normal(scale=1.5, size=(ntop,) * dim))
    oh = po.Hierarchy(dim, order, 1, levels, kappa=kap, galerkin=True)
    cs = it.GeometricCS(it.SOR(1.0), 1, 1, 0, galerkin=True)
    H = it.Hierarchy.lattice(cs, dim, order, 1, levels, kappa=kap)
    for l in range(levels):
        rp, ci, v = H.level_matrix(l).csr()
        orp, oci, ov = oh.matrix(l)
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
        assert rowrel(v, ov, orp) < 1e-12


@pytest.mark.parametrize("prec", [F.F64, F.F32])
def test_blackbox_interpolation_parity(prec):
    """Operator-dependent interpolation from a jumping coefficient: values
    against the oracle's, and the transpose (restriction) refreshed with them."""
    say(f"blackbox interpolation prec={prec}")
    n = 16
    rng = np.random.default_rng(31)
    kap = np.exp(rng.normal(scale=2.0, size=(n, n, n)))
    A, _ = lat.assemble(3, 1, n, kap, prec=prec)
    P = lat.interpolation(3, 1, n // 2, prec=prec)
    y = rng.uniform(-1, 1, P.nrows)
    before = gemm(1.0, P, GpuVector.from_array(y, prec), 0.0, GpuVector(P.ncols, prec),
                  job="tn").numpy()
    lat.blackbox_interpolation(A, P)
    prp, pci, pv = P.csr()
    orp, oci, ov, _ = po.assemble(3, 1, n, kap)
    if prec == F.F32:  # the oracle reads the same FP32-rounded matrix values
        ov = ov.astype(np.float32).astype(np.float64)
    brp, bci, bv = po.blackbox_interpolation(n, orp, oci, ov)
    assert np.array_equal(prp, brp) and np.array_equal(pci, bci)
    assert rowrel(pv, bv, brp) < (1e-13 if prec == F.F64 else 1e-6)
    after = gemm(1.0, P, GpuVector.from_array(y, prec), 0.0, GpuVector(P.ncols, prec),
                 job="tn").numpy()
    Pt = sp.csr_matrix((bv, bci, brp), shape=(P.nrows, P.ncols)).T
    assert relerr(after, Pt @ y) < (1e-13 if prec == F.F64 else 1e-5)
    assert relerr(after, before) > 1e-3  # the restriction follows the new values


@pytest.mark.parametrize("prec", [F.F64, F.F32])
@pytest.mark.parametrize("kind", ["geometric", "blackbox"])
@pytest.mark.parametrize("nc", [1, 2, 5, 8])
def test_structured_transfers_match_sell(prec, kind, nc):
    """3D Q1 lattice interpolation: the products with columns and row starts implied by the
    trilinear pattern (prolongation P y, restriction P^T r) equal the SELL products of the
    same matrix bit for bit, for several alpha / beta, and the CSR product to rounding."""
    say(f"structured transfers prec={prec} kind={kind} nc={nc}")
    rng = np.random.default_rng(40 + nc)
    P = lat.interpolation(3, 1, nc, prec=prec)
    if kind == "blackbox":
        n = 2 * nc
        A, _ = lat.assemble(3, 1, n, np.exp(rng.normal(scale=2.0, size=(n, n, n))), prec=prec)
        lat.blackbox_interpolation(A, P)
    assert P.storage() == F.STORAGE_STENCIL
    xc, zf = rng.uniform(-1, 1, P.ncols), rng.uniform(-1, 1, P.nrows)
    yf, zc = rng.uniform(-1, 1, P.nrows), rng.uniform(-1, 1, P.ncols)
    cases = [(1.0, 0.0), (1.0, 1.0), (-0.5, 2.0)]

    def products():
        out = []
        for alpha, beta in cases:
            out.append(gemm(alpha, P, GpuVector.from_array(xc, prec), beta,
                            GpuVector.from_array(zf, prec)).numpy())
            out.append(gemm(alpha, P, GpuVector.from_array(yf, prec), beta,
                            GpuVector.from_array(zc, prec), job="tn").numpy())
        return out
    implied = products()
    assert P.storage(F.STORAGE_SELL) == F.STORAGE_SELL
    sell = products()
    for a, b in zip(implied, sell):
        assert np.array_equal(a, b)
    rp, ci, v = P.csr()
    Ph = sp.csr_matrix((v, ci, rp), shape=(P.nrows, P.ncols))
    tol = 1e-14 if prec == F.F64 else 1e-6
    for k, (alpha, beta) in enumerate(cases):
        assert relerr(implied[2 * k], alpha * (Ph @ xc) + beta * zf) < tol
        assert relerr(implied[2 * k + 1], alpha * (Ph.T @ yf) + beta * zc) < tol


def test_blackbox_hierarchy_solve_parity():
    """Black-box transfers + Galerkin levels: level matrices against the
    oracle's, then CG with the V-cycle: the oracle's iteration count, defect
    history and solution."""
    say("blackbox hierarchy solve")
    dim, order, n0, levels = 3, 1, 2, 4
    ntop = n0 << (levels - 1)
    from gnu_femlisp_amd.fields import inclusions
    kap = inclusions(ntop, seed=5, rmin=1.5, rmax=3.0)
    oh = po.Hierarchy(dim, order, n0, levels, kappa=kap, transfer="blackbox")
    cs = it.GeometricCS(it.SOR(1.0), 2, 2, 0, symmetric=True, transfer="blackbox")
    H = it.Hierarchy.lattice(cs, dim, order, n0, levels, kappa=kap)
    for l in range(levels):
        rp, ci, v = H.level_matrix(l).csr()
        orp, oci, ov = oh.matrix(l)
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
        assert rowrel(v, ov, orp) < 1e-12
        # the fine matrix is exactly symmetric, so the Galerkin levels are made exactly
        # symmetric too (upper triangle mirrored): the stencil kernels rely on it
        Ml = sp.csr_matrix((v, ci, rp))
        assert (Ml != Ml.T).nnz == 0, l
    rng = np.random.default_rng(4)
    b = rng.uniform(-1, 1, oh.ndofs())
    b[po.boundary(dim, order, ntop)] = 0
    prm = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_SOR, omega=1.0,
                    ordering=po.ORDER_COLOUR, pre_steps=2, post_steps=2, base_level=0,
                    symmetric=1, alias=0, tol=1e-10, maxsteps=200)
    st, xo, so, dno = oh.solve(prm, b)
    assert st == 0
    bb = it.solve(it.LinearSolver(it.CG(cs), success_if=("<", "reduction", 1e-10),
                                  failure_if=(">", "step", 200)),
                  {"matrix": H.top_matrix(), "rhs": GpuVector.from_array(b), "hierarchy": H})
    assert bb["status"] == "success" and bb["steps"] == so
    assert relerr(bb["solution"].numpy(), xo) < 1e-10
    assert np.allclose(bb["defnorms"], dno, rtol=1e-8)


@pytest.mark.parametrize("prec", [F.F64, F.F32])
def test_lattice_galerkin_equals_csr_product(prec):
    """3D Q1: a black-box hierarchy's Galerkin levels (the product with every position taken
    from the lattices, one coarse entry per thread) equal the CSR product of the same P and A
    (flgpu_galerkin_product, which searches the rows) bit for bit, apart from the unit diagonal
    the hierarchy puts on Dirichlet rows"""
    say(f"lattice galerkin vs csr prec={prec}")
    from gnu_femlisp_amd.fields import inclusions
    from gnu_femlisp_amd.flgpu import check, lib
    n0, levels = 4, 4
    ntop = n0 << (levels - 1)
    kap = inclusions(ntop, seed=9, rmin=1.5, rmax=4.0)
    cs = it.GeometricCS(it.Jacobi(0.9), 2, 2, 0, prec=prec, transfer="blackbox")
    H = it.Hierarchy.lattice(cs, 3, 1, n0, levels, kappa=kap)
    for l in range(levels - 1):
        n = n0 << l
        Af, P = H.level_matrix(l + 1), H.level_matrix(l, interp=True)
        Ac, _ = lat.assemble(3, 1, n, prec=prec)  # the coarse lattice pattern
        check(lib().flgpu_galerkin_product(P.h, Af.h, Ac.h))
        rp, ci, v = H.level_matrix(l).csr()
        grp, gci, gv = Ac.csr()
        assert np.array_equal(rp, grp) and np.array_equal(ci, gci)
        rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        M = n + 1
        z, y, x = rows // (M * M), (rows // M) % M, rows % M
        dirichlet = ((z == 0) | (z == n) | (y == 0) | (y == n) | (x == 0) | (x == n)) & (ci == rows)
        assert np.array_equal(v[~dirichlet], gv[~dirichlet]), l
        assert np.all(v[dirichlet] == 1.0) and np.all(gv[dirichlet] == 0.0), l


def test_blackbox_mixed_fcg_reassemble():
    """FP32 black-box hierarchy inside FP64 FCG reaches a 1e-10 true residual
    with fewer steps than the geometric transfers on a 1e4-contrast problem;
    after re-assembly for another coefficient the transfers follow it (the
    solve equals a freshly built hierarchy's)."""
    say("blackbox mixed fcg")
    from gnu_femlisp_amd.fields import inclusions
    n0, levels = 2, 5
    n = n0 << (levels - 1)
    steps = {}
    for tr in ("geometric", "blackbox"):
        kap = inclusions(n, seed=11, rmin=2.0, rmax=4.0)
        cs = it.GeometricCS(it.SOR(1.0), 2, 2, 0, symmetric=True, prec=F.F32, transfer=tr)
        H = it.Hierarchy.lattice(cs, 3, 1, n0, levels, kappa=kap, mode=F.ASM_MFMA)
        A, b = lat.assemble(3, 1, n, kap, prec=F.F64)
        bb = it.solve(it.LinearSolver(it.FCG(cs), success_if=("<", "reduction", 1e-10),
                                      failure_if=(">", "step", 300), true_residual=True),
                      {"matrix": A, "rhs": b, "hierarchy": H})
        assert bb["status"] == "success"
        assert bb["true_residual"] / np.linalg.norm(b.numpy()) < 1e-10
        steps[tr] = bb["steps"]
    assert steps["blackbox"] < steps["geometric"]
    kap2 = inclusions(n, seed=12, rmin=2.0, rmax=4.0)
    H.reassemble(kap2)
    A2, b2 = lat.assemble(3, 1, n, kap2, prec=F.F64)
    H2 = it.Hierarchy.lattice(cs, 3, 1, n0, levels, kappa=kap2, mode=F.ASM_MFMA)
    res = []
    for h in (H, H2):
        bb = it.solve(it.LinearSolver(it.FCG(cs), success_if=("<", "reduction", 1e-10),
                                      failure_if=(">", "step", 300), true_residual=True),
                      {"matrix": A2, "rhs": b2, "hierarchy": h})
        res.append((bb["steps"], bb["defnorms"]))
    assert res[0][0] == res[1][0]
    assert np.allclose(res[0][1], res[1][1], rtol=1e-6)


def test_known_answer_galerkin_rate_3d():
    """mg-cdr.lisp:100 on the GPU: Galerkin GS V(1,0), CR < 0.15."""
    say("known answer galerkin")
    cs = it.GeometricCS(it.GaussSeidel(), 1, 0, 1, galerkin=True)
    H = it.Hierarchy.lattice(cs, 3, 1, 1, 3)
    bb = it.solve(it.LinearSolver(cs, success_if=("<", "defnorm", 1e-12), failure_if=(">", "step", 20)),
                  {"matrix": H.top_matrix(), "rhs": H.rhs, "hierarchy": H})
    assert bb["report"]["convergence-rate"] < 0.15


def test_alias_mode_matches_oracle():
    """femlisp's literal (iterate q w w): same (stalling) defect history."""
    say("alias cg")
    oh = po.Hierarchy(2, 1, 1, 5)
    prm = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_JACOBI, omega=0.6, pre_steps=2,
                    post_steps=2, base_level=0, alias=1, tol=1e-10, maxsteps=8)
    st, xo, so, dno = oh.solve(prm)
    cs, H = _gpu_hier(2, 1, 1, 5, po.SMOOTH_JACOBI, 0.6, 2, 2, 0)
    bb = it.solve(it.LinearSolver(it.CG(cs, alias=True), failure_if=(">", "step", 8)),
                  {"matrix": H.top_matrix(), "rhs": H.rhs, "hierarchy": H})
    assert bb["steps"] == so
    assert np.allclose(bb["defnorms"], dno, rtol=1e-9)


def test_mixed_precision_fcg():
    """FP32 V-cycle (BF16-MFMA assembled) inside FP64 FCG reaches 1e-10 true
    residual; the FP64 solution agrees with the FP64 oracle to 1e-9."""
    say("mixed fcg")
    dim, order, n0, levels = 3, 1, 1, 5
    rng = np.random.default_rng(21)
    kap = np.exp(rng.normal(scale=1.0, size=(16,) * 3))
    A, b = lat.assemble(dim, order, 16, kap, prec=F.F64)
    cs, H = _gpu_hier(dim, order, n0, levels, po.SMOOTH_SOR, 1.0, 2, 2, 1, symmetric=True,
                      kappa=kap, prec=F.F32, mode=F.ASM_MFMA)
    bb = it.solve(it.LinearSolver(it.FCG(cs), success_if=("<", "reduction", 1e-10),
                                  failure_if=(">", "step", 100), true_residual=True),
                  {"matrix": A, "rhs": b, "hierarchy": H})
    assert bb["status"] == "success"
    assert bb["true_residual"] / np.linalg.norm(b.numpy()) < 1e-10
    oh = po.Hierarchy(dim, order, n0, levels, kappa=kap)
    prm = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_SOR, ordering=po.ORDER_COLOUR,
                    pre_steps=2, post_steps=2, base_level=1, symmetric=1, tol=1e-12, maxsteps=100)
    _, xo, _, _ = oh.solve(prm)
    assert relerr(bb["solution"].numpy(), xo) < 1e-9
