"""Pins the CPU oracle (oracle/fl_oracle.c) before it is trusted as the
checker: femlisp's own known answers, closed forms and independent libraries.
Each check is chosen so that a plausible slip (dropped term, wrong sign or
index, transposed operand) fails it."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def known():
    with open(os.path.join(GOLDEN, "femlisp_known_answers.json")) as f:
        return json.load(f)


def csr(rp, ci, v):
    return sp.csr_matrix((v, ci, rp), shape=(len(rp) - 1, len(rp) - 1))


# ---- femlisp's asserted answers -------------------------------------------------

def test_vcycle_rate_1d_jacobi():
    """mg-cdr.lisp:85-87: 1D Q1 level 3, base-level 2, Jacobi 0.5, V(1,0),
    success-if (or (< :defnorm 1e-12) (> :step 10)) -> CR 0.39407..."""
    k = known()["vcycle_cr_1d_q1_l3_base2_jacobi05"]
    h = po.Hierarchy(1, 1, 1, 4)
    p = po.params(outer=po.OUTER_MG, smoother=po.SMOOTH_JACOBI, omega=0.5, pre_steps=1,
                  post_steps=0, base_level=2, stop=po.STOP_DEFNORM, tol=1e-12, maxsteps=10)
    st, x, steps, dn = h.solve(p)
    assert steps == 11
    cr = (dn[-1] / dn[0]) ** (1.0 / steps)
    assert cr < k["assert_below"]
    assert abs(cr - k["value"]) < 1e-5


def test_vcycle_rate_2d_gs():
    """mg-cdr.lisp:108-109: 2D Q1 level 3, Gauss-Seidel V(1,0) -> CR < 0.18."""
    k = known()["vcycle_cr_2d_q1_l3_gs"]
    h = po.Hierarchy(2, 1, 1, 4)
    for ordering in (po.ORDER_INDEX, po.ORDER_COLOUR):
        p = po.params(outer=po.OUTER_MG, smoother=po.SMOOTH_SOR, omega=1.0, ordering=ordering,
                      pre_steps=1, post_steps=0, base_level=1, stop=po.STOP_DEFNORM, tol=1e-12,
                      maxsteps=20)
        st, x, steps, dn = h.solve(p)
        assert (dn[-1] / dn[0]) ** (1.0 / steps) < k["assert_below"]


def test_vcycle_rate_3d_galerkin():
    """mg-cdr.lisp:100: 3D Q1 level 2 with Galerkin coarse operators, GS
    V(1,0), base-level 1 -> CR < 0.15."""
    k = known()["vcycle_cr_3d_q1_l2_galerkin"]
    h = po.Hierarchy(3, 1, 1, 3, galerkin=True)
    for ordering in (po.ORDER_INDEX, po.ORDER_COLOUR):
        p = po.params(outer=po.OUTER_MG, smoother=po.SMOOTH_SOR, omega=1.0, ordering=ordering,
                      pre_steps=1, post_steps=0, base_level=1, stop=po.STOP_DEFNORM, tol=1e-12,
                      maxsteps=20)
        st, x, steps, dn = h.solve(p)
        assert (dn[-1] / dn[0]) ** (1.0 / steps) < k["assert_below"]


def test_galerkin_equals_rediscretisation_for_constant_coefficient():
    """Trilinear interpolation is the natural injection of the coarse Q1 space,
    so P^T A P equals the re-discretised coarse matrix when kappa is constant;
    with a variable kappa it is a different (Galerkin) operator."""
    hr = po.Hierarchy(3, 1, 1, 4)
    hg = po.Hierarchy(3, 1, 1, 4, galerkin=True)
    for l in range(3):
        a, g = hr.matrix(l)[2], hg.matrix(l)[2]
        assert np.abs(a - g).max() < 1e-13 * np.abs(a).max()
    kap = np.exp(np.random.default_rng(0).normal(size=(8, 8, 8)))
    hk = po.Hierarchy(3, 1, 1, 4, kappa=kap, galerkin=True)
    rp, ci, v = hk.matrix(2)
    _, _, vr = po.assemble(3, 1, 4, kap)[:3]
    assert np.abs(v - vr).max() > 1e-3


def _bbox(n, kap):
    rp, ci, v, _ = po.assemble(3, 1, n, kap)
    prp, pci, pv = po.blackbox_interpolation(n, rp, ci, v)
    return csr(prp, pci, pv).tocsr(), n + 1, n // 2 + 1


def test_blackbox_constant_coefficient_is_trilinear():
    """For a constant coefficient the collapsed 27-point rows give weights 1/2
    per odd axis: the operator-dependent interpolation equals femlisp's
    trilinear one, pattern and values."""
    for n in (4, 8, 16):
        prp, pci, pv = po.blackbox_interpolation(n, *po.assemble(3, 1, n)[:3])
        trp, tci, tv = po.interpolation(3, 1, n // 2)
        assert np.array_equal(prp, trp) and np.array_equal(pci, tci)
        assert np.abs(pv - tv).max() < 1e-15


def test_blackbox_layered_coefficient_closed_form():
    """kappa constant on layers across axis 2: a fine node with only that
    coordinate odd sits between layers kappa_L and kappa_R, and flux
    continuity gives the weight kappa_L / (kappa_L + kappa_R) on its lower
    coarse neighbour (away from the boundary)."""
    n = 16
    lay = np.exp(np.random.default_rng(5).normal(scale=2.0, size=n))
    kap = np.broadcast_to(lay, (n, n, n)).copy()
    P, M, Mc = _bbox(n, kap)
    checked = 0
    for z in range(2, M - 2, 2):
        for y in range(2, M - 2, 2):
            for x in range(3, M - 3, 2):
                i = (z * M + y) * M + x
                lo = ((z // 2) * Mc + y // 2) * Mc + (x - 1) // 2
                kl, kr = lay[x - 1], lay[x]
                assert abs(P[i, lo] - kl / (kl + kr)) < 1e-13
                assert abs(P[i, lo + 1] - kr / (kl + kr)) < 1e-13
                checked += 1
    assert checked > 100


def test_blackbox_preserves_constants():
    """A row of A away from Dirichlet nodes sums to zero, so every fine node
    whose coordinates all lie in [2, M-3] interpolates constants exactly (its
    weights sum to 1), whatever the coefficient; boundary rows are zero."""
    n = 16
    kap = np.exp(np.random.default_rng(6).normal(scale=2.5, size=(n, n, n)))
    P, M, Mc = _bbox(n, kap)
    s = np.asarray(P.sum(axis=1)).ravel()
    c = np.indices((M, M, M)).reshape(3, -1)
    inner = np.all((c >= 2) & (c <= M - 3), axis=0)
    bnd = np.any((c == 0) | (c == M - 1), axis=0)
    assert np.abs(s[inner] - 1.0).max() < 1e-13
    assert np.abs(s[bnd]).max() == 0.0
    assert P.data.min() > -1e-14  # M-matrix rows give non-negative weights


def test_blackbox_hierarchy_beats_geometric_on_inclusions():
    """1e4-contrast inclusions (the deploy coefficient) at 32^3: the Galerkin
    hierarchy with black-box transfers needs fewer CG steps than femlisp's
    re-discretised geometric hierarchy."""
    from gnu_femlisp_amd.fields import inclusions
    kap = inclusions(32)
    prm = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_SOR, omega=1.0,
                    ordering=po.ORDER_COLOUR, pre_steps=2, post_steps=2, base_level=0,
                    symmetric=1, tol=1e-10, maxsteps=300)
    steps = {}
    for tr in ("geometric", "blackbox"):
        h = po.Hierarchy(3, 1, 2, 5, kappa=kap, transfer=tr)
        st, x, steps[tr], dn = h.solve(prm)
        assert st == 0
    assert steps["blackbox"] < steps["geometric"]


def _cycle_matrix(h, prm):
    N = h.ndofs()
    return np.stack([h.apply(prm, e) for e in np.eye(N)], axis=1)


def test_gamma_cycles():
    """femlisp's gamma (multigrid.lisp:268-273) pinned by what it must reduce to:
    gamma = 0 is smoothing only (pre + post damped Jacobi from zero, restated
    with scipy); on two levels every gamma is the V-cycle (the coarse level is
    the base); and with many coarse cycles the three-level W-cycle approaches
    the two-grid method with an exact coarse solve."""
    h = po.Hierarchy(1, 1, 2, 4)
    rp, ci, v = h.matrix(3)
    A = csr(rp, ci, v).toarray()
    d = np.diag(A)
    base = dict(outer=po.OUTER_MG, smoother=po.SMOOTH_JACOBI, omega=0.6, pre_steps=2,
                post_steps=1)
    # gamma 0: x = post(pre(0)) with 3 Jacobi steps on b
    M0 = _cycle_matrix(h, po.params(**base, base_level=0, gamma=0))
    S = np.eye(len(d)) - 0.6 * (A / d[:, None])
    Mj = np.zeros_like(A)
    for _ in range(3):
        Mj = S @ Mj + 0.6 * np.diag(1 / d)
    assert np.abs(M0 - Mj).max() < 1e-14
    # two levels (base 2 of levels 2, 3): gamma does not matter
    M1 = _cycle_matrix(h, po.params(**base, base_level=2, gamma=1))
    M2 = _cycle_matrix(h, po.params(**base, base_level=2, gamma=2))
    assert np.array_equal(M1, M2)
    # W-cycle with many coarse cycles on levels 3, 2, 1 -> exact solve on level 2
    Mw = _cycle_matrix(h, po.params(**base, base_level=1, gamma=60))
    Mw1 = _cycle_matrix(h, po.params(**base, base_level=1, gamma=1))
    assert np.abs(Mw - M1).max() < 1e-10 * np.abs(M1).max()
    assert np.abs(Mw1 - M1).max() > 1e-6 * np.abs(M1).max()


def test_additive_combination():
    """:additive (multigrid.lisp:262-287) restricts the residual before smoothing, so on
    two levels its cycle is the smoothing-only cycle plus P A_c^-1 P^T exactly, with
    P and A_c read back from the hierarchy; the multiplicative cycle is not."""
    for dim, order, n0 in ((1, 1, 4), (2, 1, 2), (2, 2, 1)):
        h = po.Hierarchy(dim, order, n0, 2)
        base = dict(outer=po.OUTER_MG, smoother=po.SMOOTH_JACOBI, omega=0.6, pre_steps=2,
                    post_steps=1, base_level=0)
        Madd = _cycle_matrix(h, po.params(**base, additive=1))
        Msm = _cycle_matrix(h, po.params(**base, gamma=0))
        Mmul = _cycle_matrix(h, po.params(**base))
        prp, pci, pv = h.interp(0)
        P = sp.csr_matrix((pv, pci, prp), shape=(h.ndofs(1), h.ndofs(0))).toarray()
        crp, cci, cv = h.matrix(0)
        Ac = csr(crp, cci, cv).toarray()
        C = P @ np.linalg.inv(Ac) @ P.T
        assert np.abs(Madd - (Msm + C)).max() < 1e-12 * np.abs(Madd).max()
        assert np.abs(Mmul - (Msm + C)).max() > 1e-6 * np.abs(Madd).max()


def test_fas_equals_correction_scheme_for_linear_problems():
    """Brandt's FAS (<geometric-fas>, geomg.lisp:174-194; restrict/prolongate at
    multigrid.lisp:226-247) approximates the solution on every level.  For a linear
    operator its multiplicative cycle is the correction scheme's up to rounding, whatever the
    smoother, the number of coarse cycles, the dimension or the order: the coarse solution
    starts from the projected fine one with rhs_c = I^T r + A_c R x, and the correction
    prolongated is x_c - R x.  Dropping the A_c R x term, or projecting anything but the
    coincident fine node, breaks the equality."""
    rng = np.random.default_rng(12)
    for dim, order, n0, levels in ((1, 1, 2, 4), (2, 1, 2, 3), (2, 2, 1, 3), (3, 1, 1, 3)):
        h = po.Hierarchy(dim, order, n0, levels)
        b = rng.uniform(-1, 1, h.ndofs())
        for smoother, omega in ((po.SMOOTH_JACOBI, 0.6), (po.SMOOTH_SOR, 1.0)):
            for gamma in (1, 2):
                base = dict(outer=po.OUTER_MG, smoother=smoother, omega=omega,
                            ordering=po.ORDER_COLOUR, pre_steps=2, post_steps=1, base_level=0,
                            gamma=gamma)
                x_cs = h.apply(po.params(**base), b)
                x_fas = h.apply(po.params(**base, fas=1), b)
                assert np.linalg.norm(x_fas - x_cs) < 1e-12 * np.linalg.norm(x_cs), \
                    (dim, order, smoother, gamma)
        # as a preconditioner of <cg>: the same steps and solution
        prm = dict(outer=po.OUTER_CG, smoother=po.SMOOTH_SOR, omega=1.0,
                   ordering=po.ORDER_COLOUR, pre_steps=2, post_steps=2, base_level=0,
                   symmetric=1, tol=1e-10, maxsteps=200)
        st0, x0, s0, _ = h.solve(po.params(**prm), b)
        st1, x1, s1, _ = h.solve(po.params(**prm, fas=1), b)
        assert st0 == st1 == 0 and s0 == s1
        assert np.linalg.norm(x1 - x0) < 1e-9 * np.linalg.norm(x0)


def test_fas_projection_is_nodal_injection():
    """FAS-R is femlisp's projection-matrix: the injection of the dof functionals
    (fetransfer.lisp:201-213).  Pinned by the reference's own identity R I = Id
    (sparseas.lisp:436-448 checks P I = Id) and by what injection means: R applied to a
    function sampled on the fine lattice gives its samples at the coarse nodes exactly."""
    for dim, order, nc in ((1, 1, 4), (2, 1, 3), (2, 2, 2), (3, 1, 2), (3, 2, 1)):
        rp, ci, v = po.projection(dim, order, nc)
        Nc, Nf = po.ndofs(dim, order, nc), po.ndofs(dim, order, 2 * nc)
        R = sp.csr_matrix((v, ci, rp), shape=(Nc, Nf))
        irp, ici, iv = po.interpolation(dim, order, nc)
        I = sp.csr_matrix((iv, ici, irp), shape=(Nf, Nc))
        RI = (R @ I).toarray()
        # interior identity exactly; Dirichlet rows of I are zeroed (constraints.lisp:58-72)
        inner = np.flatnonzero(~po.boundary(dim, order, nc))
        assert np.array_equal(RI[np.ix_(inner, inner)], np.eye(len(inner)))

        def sample(n):
            M = order * n + 1
            g = np.meshgrid(*([np.arange(M) / (M - 1)] * dim), indexing="ij")
            return np.ravel(np.sin(3 * g[0]) + sum((k + 2) * g[k] ** 2 for k in range(1, dim)))
        assert np.array_equal(R @ sample(2 * nc), sample(nc))
    h = po.Hierarchy(3, 1, 1, 3)
    for l in range(2):
        a, b = h.projection(l), po.projection(3, 1, 1 << l)
        assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_fe_value_1d():
    """ellsys-fe.lisp:249-256: 1D Q1 -u''=1 on level 2 gives u(0.5) = 0.125 exactly."""
    rp, ci, v, b = po.assemble(1, 1, 4)
    u = sla.spsolve(csr(rp, ci, v).tocsc(), b)
    assert u[2] == known()["fe_value_1d_q1_l2_at_half"]["value"]


def test_centre_values_converge():
    """model-problem.lisp:39-87,184-199: exact centre values of -Lap u = 1 are
    h-convergence targets; Q1 error must fall ~4x per refinement."""
    ex = known()["exact_centre_values"]
    for dim, ns in ((2, (8, 16, 32)), (3, (4, 8, 16))):
        errs = []
        for n in ns:
            rp, ci, v, b = po.assemble(dim, 1, n)
            u = sla.spsolve(csr(rp, ci, v).tocsc(), b)
            M = n + 1
            centre = sum((n // 2) * M ** k for k in range(dim))
            errs.append(abs(u[centre] - ex[str(dim)]))
        assert errs[-1] < errs[0]
        ratios = [errs[i] / errs[i + 1] for i in range(len(errs) - 1)]
        assert all(3.0 < r < 5.0 for r in ratios), ratios


# ---- closed forms of the pattern ----------------------------------------------------

@pytest.mark.parametrize("dim,order,n", [(1, 1, 5), (2, 1, 7), (3, 1, 4), (2, 2, 3), (3, 2, 3)])
def test_pattern_closed_forms(dim, order, n):
    rp, ci = po.pattern(dim, order, n)
    N = (order * n + 1) ** dim
    assert len(rp) == N + 1
    assert rp[-1] == ((3 * n + 1) if order == 1 else (8 * n + 1)) ** dim
    A = sp.csr_matrix((np.ones(len(ci)), ci, rp), shape=(N, N))
    assert (A != A.T).nnz == 0  # symmetric pattern
    for r in range(N):
        row = ci[rp[r]:rp[r + 1]]
        assert np.all(np.diff(row) > 0) and r in row


def test_pattern_row_counts_q1_3d():
    rp, ci = po.pattern(3, 1, 4)
    cnt = np.diff(rp).reshape(5, 5, 5)
    assert cnt[2, 2, 2] == 27 and cnt[0, 2, 2] == 18 and cnt[0, 0, 2] == 12 and cnt[0, 0, 0] == 8


def test_pattern_row_counts_q2_3d():
    rp, ci = po.pattern(3, 2, 3)
    cnt = np.diff(rp).reshape(7, 7, 7)
    assert cnt[2, 2, 2] == 125 and cnt[1, 2, 2] == 75 and cnt[1, 1, 2] == 45 and cnt[1, 1, 1] == 27


# ---- quadrature / basis / element ------------------------------------------------

@pytest.mark.parametrize("s", [1, 2, 3, 4, 5])
def test_gauss_rule_against_numpy(s):
    p, w = po.gauss_rule(s)
    x, wx = np.polynomial.legendre.leggauss(s)
    # femlisp's order: points (1 - x_i)/2 for ascending roots
    assert np.allclose(p, (1 - x) / 2, atol=1e-15, rtol=0)
    assert np.allclose(w, wx / 2, atol=1e-15, rtol=0)
    for k in range(2 * s):
        assert abs(np.dot(w, p ** k) - 1.0 / (k + 1)) < 1e-14


def test_lagrange_basis_partition_and_nodes():
    for order in (1, 2):
        for t in np.linspace(0, 1, 7):
            phi, dphi = po.lagrange_1d(order, t)
            assert abs(phi.sum() - 1) < 1e-15 and abs(dphi.sum()) < 1e-13
        for a in range(order + 1):
            phi, _ = po.lagrange_1d(order, a / order)
            assert np.allclose(phi, np.eye(order + 1)[a], atol=1e-15)


def test_element_matrix_1d_closed_form():
    h = 0.25
    K, b = po.element_matrix(1, 1, h)
    assert np.allclose(K, np.array([[1, -1], [-1, 1]]) / h, rtol=1e-15, atol=1e-14)
    assert np.allclose(b, [h / 2, h / 2], rtol=1e-15)
    K, b = po.element_matrix(1, 2, h)
    assert np.allclose(K, np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / (3 * h), rtol=1e-14)
    assert np.allclose(b, h * np.array([1 / 6, 2 / 3, 1 / 6]), rtol=1e-14)


def test_element_matrix_tensor_product():
    """d-dim stiffness = sum over axes of K1 (x) M1 (x) M1 with the 1D mass
    matrix M1 evaluated by the same Gauss rule (exact for these degrees)."""
    for order in (1, 2):
        h = 0.5
        K1, _ = po.element_matrix(1, order, h)
        s = order + 1
        p, w = po.gauss_rule(s)
        phis = np.array([po.lagrange_1d(order, t)[0] for t in p])
        M1 = h * (phis.T * w) @ phis
        K1u = K1 * h  # unit-length stiffness scaled below
        for dim in (2, 3):
            K, _ = po.element_matrix(dim, order, h)
            ref = 0
            for axis in range(dim):
                mats = [M1 / h] * dim
                mats[axis] = K1u
                t = mats[0]
                for m in mats[1:]:
                    t = np.kron(t, m)
                ref = ref + t * h ** (dim - 2)
            assert np.allclose(K, ref, rtol=1e-13, atol=1e-14)


def test_q1_stencils():
    """2D: centre 8/3, neighbours -1/3; 3D (spacing h): centre 8h/3, axis 0,
    face-diagonal -h/6, corner -h/12; interior rows sum to zero."""
    rp, ci, v, b = po.assemble(2, 1, 8)
    A = csr(rp, ci, v).toarray()
    M = 9
    c = 4 * M + 4
    assert abs(A[c, c] - 8 / 3) < 1e-14
    for d in (-M - 1, -M, -M + 1, -1, 1, M - 1, M, M + 1):
        assert abs(A[c, c + d] + 1 / 3) < 1e-14
    n = 8
    h = 1.0 / n
    rp, ci, v, b = po.assemble(3, 1, n)
    A = csr(rp, ci, v)
    M = n + 1
    c = 4 * M * M + 4 * M + 4
    row = dict(zip(ci[rp[c]:rp[c + 1]], v[rp[c]:rp[c + 1]]))
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                k = abs(dx) + abs(dy) + abs(dz)
                want = {0: 8 * h / 3, 1: 0.0, 2: -h / 6, 3: -h / 12}[k]
                assert abs(row[c + dz * M * M + dy * M + dx] - want) < 1e-15
    assert abs(b[c] - h ** 3) < 1e-16
    assert abs(sum(row.values())) < 1e-15


def test_dirichlet_elimination():
    rp, ci, v, b = po.assemble(2, 2, 3)
    bd = po.boundary(2, 2, 3)
    A = csr(rp, ci, v)
    for r in np.where(bd)[0]:
        row = dict(zip(ci[rp[r]:rp[r + 1]], v[rp[r]:rp[r + 1]]))
        assert row[r] == 1.0 and all(val == 0.0 for c, val in row.items() if c != r)
        assert b[r] == 0.0
    Ad = A.toarray()
    assert np.all(Ad[np.ix_(~bd, bd)] == 0)
    assert np.allclose(Ad, Ad.T, atol=1e-14)
    assert np.all(np.linalg.eigvalsh(Ad) > 0)


def test_kappa_sampling():
    """Piecewise-constant kappa: scaling kappa by c scales interior rows by c;
    a kappa on a finer grid sampled at coarse quadrature points."""
    rng = np.random.default_rng(1)
    kap = np.exp(rng.normal(size=(4, 4)))
    rp, ci, v1, _ = po.assemble(2, 1, 4, kap)
    _, _, v2, _ = po.assemble(2, 1, 4, 3.0 * kap)
    bd = po.boundary(2, 1, 4)
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    inner = ~bd[rows] & ~bd[ci]
    assert np.allclose(v2[inner], 3.0 * v1[inner], rtol=1e-14)
    # constant kappa on a finer coefficient grid == kappa = const
    _, _, v3, _ = po.assemble(2, 1, 4, np.full((16, 16), 2.0))
    _, _, v4, _ = po.assemble(2, 1, 4, np.full((4, 4), 2.0))
    assert np.allclose(v3, v4, rtol=1e-15)


# ---- interpolation -----------------------------------------------------------------

@pytest.mark.parametrize("dim,order,nc", [(1, 1, 4), (2, 1, 3), (3, 1, 2), (1, 2, 3), (2, 2, 2), (3, 2, 2)])
def test_interpolation(dim, order, nc):
    rp, ci, v = po.interpolation(dim, order, nc)
    assert rp[-1] == ((3 * nc + 1) if order == 1 else (8 * nc + 1)) ** dim
    Mc, Mf = order * nc + 1, 2 * order * nc + 1
    P = sp.csr_matrix((v, ci, rp), shape=(Mf ** dim, Mc ** dim))
    # interpolation reproduces polynomials of the element degree that vanish
    # on the boundary (sparseas.lisp:401-448: x(1-x) reproduced to 1e-10)
    def f(coords):
        return np.prod([c * (1 - c) for c in coords], axis=0) * (
            1 + (0.3 * coords[0] if order == 2 else 0))
    def grid(M):
        g = np.indices((M,) * dim).reshape(dim, -1) / (M - 1)
        return g
    if order == 1:
        # a product of hats is reproduced only at coarse nodes; test linear
        # interpolation property on the midpoint of each coarse edge instead
        uc = np.random.default_rng(0).normal(size=Mc ** dim)
        bdc = po.boundary(dim, order, nc)
        uc[bdc] = 0
        uf = P @ uc
        fine = np.indices((Mf,) * dim).reshape(dim, -1)
        even = np.all(fine % 2 == 0, axis=0)
        coarse_idx = sum((fine[k][even] // 2) * Mc ** (dim - 1 - k) for k in range(dim))
        assert np.allclose(uf[even], uc[coarse_idx])
    else:
        uc = np.prod([c * (1 - c) for c in grid(Mc)], axis=0)
        uf = np.prod([c * (1 - c) for c in grid(Mf)], axis=0)
        bdf = po.boundary(dim, order, 2 * nc)
        assert np.allclose((P @ uc)[~bdf], uf[~bdf], atol=1e-14)
    # Dirichlet rows / columns zeroed but stored
    bdf = po.boundary(dim, order, 2 * nc)
    bdc = po.boundary(dim, order, nc)
    rows = np.repeat(np.arange(Mf ** dim), np.diff(rp))
    assert np.all(v[bdf[rows] | bdc[ci]] == 0.0)


# ---- the CG preconditioner aliasing of femlisp ---------------------------------------

def test_cg_alias_quirk():
    """femlisp's <cg> calls the preconditioner as (iterate q w w)
    (krylow.lisp:118-121); with Jacobi's in-place gesv! (linit.lisp:234) the
    V(2,2) preconditioner then sees a clobbered rhs and CG stalls, while the
    same V-cycle on separate vectors converges fast."""
    h = po.Hierarchy(2, 1, 1, 6)
    res = {}
    for alias in (1, 0):
        p = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_JACOBI, omega=0.6, pre_steps=2,
                      post_steps=2, base_level=0, alias=alias, tol=1e-10, maxsteps=60)
        res[alias] = h.solve(p)
    assert res[1][0] == 1  # failure: step limit
    assert res[0][0] == 0 and res[0][2] <= 10


def test_fcg_equals_cg_for_spd_preconditioner():
    h = po.Hierarchy(3, 1, 1, 4)
    kw = dict(smoother=po.SMOOTH_JACOBI, omega=0.6, pre_steps=2, post_steps=2, base_level=1,
              alias=0, tol=1e-10, maxsteps=50)
    _, x1, s1, d1 = h.solve(po.params(outer=po.OUTER_CG, **kw))
    _, x2, s2, d2 = h.solve(po.params(outer=po.OUTER_FCG, **kw))
    assert s1 == s2
    assert np.allclose(d1, d2, rtol=1e-6)


def test_solution_matches_direct_solve():
    rng = np.random.default_rng(3)
    kap = np.exp(rng.normal(scale=1.5, size=(8, 8, 8)))
    h = po.Hierarchy(3, 1, 1, 4, kappa=kap)
    p = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_SOR, omega=1.0, ordering=po.ORDER_COLOUR,
                  pre_steps=2, post_steps=2, symmetric=1, base_level=0, tol=1e-12, maxsteps=100)
    st, x, steps, dn = h.solve(p)
    rp, ci, v = h.matrix()
    u = sla.spsolve(csr(rp, ci, v).tocsc(), h.rhs())
    assert st == 0
    assert np.linalg.norm(x - u) / np.linalg.norm(u) < 1e-10


# ---- <cg> restart-cycle (krylow.lisp:68-70,101-105,128-131) -------------------------------

def _steepest_descent(A, b, steps):
    """textbook steepest descent from zero with the true residual each step:
    x += (r.r / r.Ar) r, r = b - A x; returns the residual norms ||r_0..r_steps||"""
    x = np.zeros_like(b)
    out = []
    for _ in range(steps + 1):
        r = b - A @ x
        out.append(np.linalg.norm(r))
        x = x + (r @ r) / (r @ (A @ r)) * r
    return np.array(out)


def test_cg_restart_cycle_one_is_steepest_descent():
    """With restart-cycle 1 every step restarts from the recomputed residual: plain <cg>
    becomes steepest descent (a dropped restart or a missing residual refresh differs)."""
    h = po.Hierarchy(2, 1, 1, 4)
    A = csr(*h.matrix())
    b = h.rhs()
    p = po.params(outer=po.OUTER_CG_PLAIN, tol=1e-30, maxsteps=11, restart_cycle=1)
    st, x, steps, dn = h.solve(p, b)
    assert st == 1 and steps == 12
    ref = _steepest_descent(A, b, 12)
    assert np.allclose(dn, ref[: len(dn)], rtol=1e-10, atol=0)
    # without the restart the same solver is CG, which is much faster here
    _, _, _, dcg = h.solve(po.params(outer=po.OUTER_CG_PLAIN, tol=1e-30, maxsteps=11), b)
    assert dcg[-1] < 1e-3 * dn[-1]


@pytest.mark.parametrize("outer", [po.OUTER_CG, po.OUTER_CG_PLAIN])
def test_cg_restart_is_cg_on_the_residual_equation(outer):
    """After a restart cycle of k steps, CG starts afresh on A d = b - A x_k: two cycles
    give x_k + (k CG steps on the residual equation from zero)."""
    k = 3
    h = po.Hierarchy(3, 1, 1, 4)
    A = csr(*h.matrix())
    b = h.rhs()
    kw = dict(outer=outer, smoother=po.SMOOTH_JACOBI, omega=0.6, pre_steps=2, post_steps=2,
              base_level=1, tol=1e-30)
    # maxsteps m stops after m + 1 steps
    _, xk, sk, _ = h.solve(po.params(maxsteps=k - 1, **kw), b)
    _, d, _, _ = h.solve(po.params(maxsteps=k - 1, **kw), b - A @ xk)
    _, x2k, s2k, dn = h.solve(po.params(maxsteps=2 * k - 1, restart_cycle=k, **kw), b)
    assert sk == k and s2k == 2 * k
    assert np.linalg.norm(x2k - (xk + d)) <= 1e-12 * np.linalg.norm(x2k)
    # the defect observed after the cycle is the recomputed true residual
    assert np.isclose(dn[k], np.linalg.norm(b - A @ xk), rtol=1e-12)
    # a cycle longer than the run changes nothing
    _, x1, _, d1 = h.solve(po.params(maxsteps=2 * k - 1, **kw), b)
    _, x2, _, d2 = h.solve(po.params(maxsteps=2 * k - 1, restart_cycle=50, **kw), b)
    assert np.array_equal(x1, x2) and np.array_equal(d1, d2)


def test_cg_restart_recomputes_the_residual():
    """Past convergence the recurrence residual falls far below rounding while the true one
    stalls near 1e-16: the defect after a completed cycle is the recomputed b - A x."""
    h = po.Hierarchy(2, 1, 1, 4)
    A = csr(*h.matrix())
    b = h.rhs()
    k = 40
    _, _, _, d0 = h.solve(po.params(outer=po.OUTER_CG_PLAIN, tol=1e-300, maxsteps=k - 1), b)
    _, xk, _, _ = h.solve(po.params(outer=po.OUTER_CG_PLAIN, tol=1e-300, maxsteps=k - 1), b)
    _, _, _, d1 = h.solve(po.params(outer=po.OUTER_CG_PLAIN, tol=1e-300, maxsteps=2 * k,
                                    restart_cycle=k), b)
    true_k = np.linalg.norm(b - A @ xk)
    assert d0[k] < 1e-20 * true_k  # the recurrence has drifted off
    assert np.isclose(d1[k], true_k, rtol=1e-2)  # rounding-level value, summation order
    assert np.array_equal(d1[:k], d0[:k])


# ---- <safe-linear-solver> (linsolve.lisp:45-56,75-89) --------------------------------------

def test_safe_linear_solver_falls_back_to_lu():
    """On failure the iteration becomes <lu> and the loop goes on with the same step count:
    one <lu> step on the residual at hand solves the system."""
    h = po.Hierarchy(2, 1, 1, 6)  # 2D Q1 32^2, 1089 DOFs
    A = csr(*h.matrix())
    b = h.rhs()
    kw = dict(outer=po.OUTER_CG, smoother=po.SMOOTH_JACOBI, omega=0.6, pre_steps=2,
              post_steps=2, base_level=0, tol=1e-10, maxsteps=2)
    st0, x0, s0, d0 = h.solve(po.params(**kw), b)
    st, x, steps, dn = h.solve(po.params(fallback=po.FALLBACK_LU, **kw), b)
    assert st0 == 1 and s0 == 3
    assert st == 0 and steps == s0 + 1
    assert np.array_equal(dn[: s0 + 1], d0)  # the failed iteration's history is kept
    u = sla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - u) <= 1e-12 * np.linalg.norm(u)
    # the <lu> step used the recurrence residual: x = x_3 + A^-1 r_3
    assert dn[-1] < 1e-10 * dn[0]


def test_safe_linear_solver_fallback_failure():
    """A second failure is femlisp's 'The fallback iteration has failed.' error."""
    h = po.Hierarchy(2, 1, 1, 3)
    p = po.params(outer=po.OUTER_CG_PLAIN, tol=1e-300, maxsteps=1, fallback=po.FALLBACK_LU)
    st, x, steps, dn = h.solve(p)
    assert st == 2 and steps == 3


# ---- vertex-centred block smoothers (blockit.lisp:113-203, geoblock.lisp:65-157) ----------

def _geometric_blocks(dim, order, n):
    """Blocks from the definition: vertex v and every non-Dirichlet DOF whose carrying cell
    has v in its closure.  A DOF at lattice point m is carried by the cell that extends along
    the axes where m is not on a vertex plane; that cell contains v when m_k = v_k on those
    planes and |m_k - v_k| < order elsewhere.  Returned as {vertex: set of keys} and colours."""
    M = order * n + 1
    bd = po.boundary(dim, order, n)
    pts = np.array(np.unravel_index(np.arange(M ** dim), (M,) * dim)).T
    blocks, colour = {}, {}
    for a in np.ndindex(*(n + 1,) * dim):
        v = np.array(a) * order
        key = int(np.ravel_multi_index(tuple(v), (M,) * dim))
        on_plane = pts % order == 0
        inside = np.where(on_plane, pts == v, np.abs(pts - v) < order).all(axis=1)
        ks = set(np.flatnonzero(inside & ~bd).tolist()) | {key}
        blocks[key] = ks
        colour[key] = int(sum((a[k] & 1) << (dim - 1 - k) for k in range(dim)))
    return blocks, colour


@pytest.mark.parametrize("dim,order,n", [(2, 1, 4), (2, 2, 3), (3, 1, 3), (3, 2, 2)])
def test_vertex_blocks_match_the_geometric_definition(dim, order, n):
    """flo_vertex_blocks against the blocks built from cell closures, the colour order, and
    the independence that makes any order within a colour give the same iterate."""
    keys = po.vertex_blocks(dim, order, n)
    geo, colour = _geometric_blocks(dim, order, n)
    assert keys.shape == ((n + 1) ** dim, (2 * order - 1) ** dim)
    W = keys.shape[1]
    centres = keys[:, W // 2]
    assert sorted(centres.tolist()) == sorted(geo)
    for row, c in zip(keys, centres):
        assert set(row[row >= 0].tolist()) == geo[int(c)]
    cols = [colour[int(c)] for c in centres]
    assert cols == sorted(cols)
    rp, ci, _, _ = po.assemble(dim, order, n)
    A = csr(rp, ci, np.ones(len(ci))).toarray() != 0
    for col in set(cols):
        mine = [keys[i][keys[i] >= 0] for i in range(len(cols)) if cols[i] == col]
        for i in range(len(mine)):
            for j in range(i + 1, len(mine)):
                assert not A[np.ix_(mine[i], mine[j])].any()


def _brute_ssc(A, blocks, colour, x, b, omega, rng):
    """<ssc> iterate from its definition, blocks colour by colour in random order within a
    colour, with numpy's dense solve; with omega = 1 the block's residual must vanish."""
    x = x.copy()
    for col in sorted(set(colour.values())):
        vs = [v for v in blocks if colour[v] == col]
        for v in rng.permutation(vs):
            ks = np.array(sorted(blocks[v]))
            r = b[ks] - A[ks] @ x
            x[ks] += omega * np.linalg.solve(A[np.ix_(ks, ks)], r)
            if omega == 1.0:
                assert np.abs(b[ks] - A[ks] @ x).max() < 1e-12 * np.abs(b).max()
    return x


@pytest.mark.parametrize("dim,n", [(2, 4), (3, 2), (3, 3)])
@pytest.mark.parametrize("omega", [1.0, 0.8])
def test_ssc_matches_the_definition(dim, n, omega):
    """One Q2 <geometric-ssc> iterate against the brute-force definition.  x is random at
    the Dirichlet nodes too, so a block that wrongly holds a Dirichlet key other than its
    vertex changes the result."""
    rng = np.random.default_rng(dim * 10 + n)
    rp, ci, v, _ = po.assemble(dim, 2, n, kappa=np.exp(rng.normal(size=n ** dim)))
    A = csr(rp, ci, v).toarray()
    N = A.shape[0]
    x0, b = rng.normal(size=N), rng.normal(size=N)
    x, _ = po.block_iterate(rp, ci, v, dim, 2, n, True, omega, x0, b, np.zeros(N))
    geo, colour = _geometric_blocks(dim, 2, n)
    ref = _brute_ssc(A, geo, colour, x0, b, omega, rng)
    assert np.abs(x - ref).max() < 1e-12 * np.abs(ref).max()
    # backwards: the colours last to first
    xb, _ = po.block_iterate(rp, ci, v, dim, 2, n, True, omega, x0, b, np.zeros(N),
                             backwards=True)
    rev = {k: (1 << dim) - 1 - c for k, c in colour.items()}
    refb = _brute_ssc(A, geo, rev, x0, b, omega, rng)
    assert np.abs(xb - refb).max() < 1e-12 * np.abs(refb).max()


@pytest.mark.parametrize("dim,n", [(2, 4), (3, 2)])
def test_psc_matches_the_definition(dim, n):
    """<geometric-psc>: x += damp sum_B R_B^T A_BB^-1 r_B with the residual as given."""
    rng = np.random.default_rng(5)
    rp, ci, v, _ = po.assemble(dim, 2, n, kappa=np.exp(rng.normal(size=n ** dim)))
    A = csr(rp, ci, v).toarray()
    N = A.shape[0]
    x0, b = rng.normal(size=N), rng.normal(size=N)
    r0 = b - A @ x0
    x, r = po.block_iterate(rp, ci, v, dim, 2, n, False, 0.7, x0, b, r0)
    assert np.array_equal(r, r0)
    ref = x0.copy()
    for ks in _geometric_blocks(dim, 2, n)[0].values():
        ks = np.array(sorted(ks))
        ref[ks] += 0.7 * np.linalg.solve(A[np.ix_(ks, ks)], r0[ks])
    assert np.abs(x - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_q1_block_smoothers_are_point_smoothers(dim):
    """For Q1 every vertex block is one DOF, so <geometric-ssc> is <sor> in colour order and
    <geometric-psc> is <jacobi> with the same damping: whole V-cycles agree exactly."""
    rng = np.random.default_rng(dim)
    n0, lv = (2, 4) if dim == 2 else (1, 3)
    nt = n0 << (lv - 1)
    h = po.Hierarchy(dim, 1, n0, lv, kappa=np.exp(rng.normal(size=nt ** dim)))
    b = rng.normal(size=h.ndofs())
    for blk, point in ((po.SMOOTH_SSC, po.SMOOTH_SOR), (po.SMOOTH_PSC, po.SMOOTH_JACOBI)):
        for steps in (1, 2):
            kw = dict(omega=0.9, ordering=po.ORDER_COLOUR, pre_steps=steps, post_steps=steps,
                      symmetric=1)
            x1 = h.apply(po.params(smoother=blk, **kw), b)
            x2 = h.apply(po.params(smoother=point, **kw), b)
            assert np.array_equal(x1, x2)


@pytest.mark.parametrize("dim,n0,levels", [(2, 1, 5), (3, 1, 3)])
def test_ssc_vcycle_contracts_for_q2(dim, n0, levels):
    """femlisp picks <geometric-ssc> for order > 1 (geomg.lisp:58-71); its own test asserts a
    V-cycle rate below 0.06 for 2D order 4 (mg-cdr.lisp:102-103), which this Q1/Q2 oracle
    cannot run.  Q2 with V(1,1) must contract at least that fast here; the measured rates
    are 0.007 (2D, 16 cells per axis) and 0.010 (3D, 4 cells per axis)."""
    h = po.Hierarchy(dim, 2, n0, levels)
    p = po.params(outer=po.OUTER_MG, smoother=po.SMOOTH_SSC, omega=1.0, pre_steps=1,
                  post_steps=1, stop=po.STOP_REDUCTION, tol=1e-10, maxsteps=30)
    st, x, steps, dn = h.solve(p)
    assert st == 0
    rate = (dn[-1] / dn[0]) ** (1.0 / steps)
    assert rate < 0.06, rate


# ---- host threads ---------------------------------------------------------------------------

@pytest.mark.parametrize("smoother,order", [(po.SMOOTH_SOR, 1), (po.SMOOTH_JACOBI, 1),
                                            (po.SMOOTH_SSC, 2)])
def test_oracle_results_do_not_depend_on_threads(smoother, order):
    """the CPU baseline's parallel loops (rows, one colour's rows or blocks, fixed dot ranges)
    give the 1-thread results bit for bit"""
    n0, levels = (1, 5) if order == 1 else (1, 4)
    rng = np.random.default_rng(9)
    ntop = n0 << (levels - 1)
    kap = np.exp(rng.normal(size=(ntop,) * 3))
    h = po.Hierarchy(3, order, n0, levels, kappa=kap)
    p = po.params(outer=po.OUTER_CG, smoother=smoother, omega=1.0 if smoother else 0.6,
                  ordering=po.ORDER_COLOUR, pre_steps=2, post_steps=2, base_level=1,
                  symmetric=1, tol=1e-10, maxsteps=100)
    out = []
    for t in (1, 0):
        po.set_threads(t)
        out.append(h.solve(p))
    po.set_threads(0)
    (s1, x1, n1, d1), (s2, x2, n2, d2) = out
    assert s1 == s2 == 0 and n1 == n2
    assert np.array_equal(x1, x2) and np.array_equal(d1, d2)
