"""Linear iterations and solvers with femlisp's interface (FL.ITERATION,
FL.MULTIGRID, FL.GEOMG), executed by libflgpu.so.

Classes and their femlisp counterparts:
  Jacobi(damp)                <jacobi>          src/iteration/linit.lisp:200-252
  SOR(omega), GaussSeidel()   <sor>/<gauss-seidel> with a colour :compare
                                                src/iteration/linit.lisp:258-385
  GeometricSSC(omega)         <geometric-ssc>   src/iteration/geoblock.lisp:47-49,57-59,
                                                src/iteration/blockit.lisp:80-82,113-203
  GeometricPSC(damp)          <geometric-psc>   src/iteration/geoblock.lisp:43-45,53-55,
                                                src/iteration/blockit.lisp:76-78,113-203
  select_smoother(order)      select-smoother   src/iteration/geomg.lisp:58-71
  GeometricCS(...)            <geometric-cs>    src/iteration/geomg.lisp:109-155
  GeometricFAS(...)           <geometric-fas>   src/iteration/geomg.lisp:174-194
  CG(preconditioner)          <cg>              src/iteration/krylow.lisp:67-132
  FCG(preconditioner)         flexible CG (no femlisp counterpart; equals <cg>
                              for a fixed SPD preconditioner)
  LinearSolver(iteration, success_if, failure_if)
                              <linear-solver>   src/iteration/linsolve.lisp:40-73
  make_iterator(linit, mat)   make-iterator     src/iteration/linit.lisp:65-67
  solve(solver, blackboard)   solve             src/iteration/solve.lisp:46-71

success_if / failure_if take femlisp's forms in tuple spelling:
("<", "reduction", 1e-10), ("<", "defnorm", 1e-12), (">", "step", 100).
"""
import ctypes
import math

import numpy as np

from . import flgpu as _f
from .flgpu import F32, F64, MgParams, SolveParams, SolveReport, check, lib, ptr
from .matlisp import GpuSparseMatrix, GpuVector, compute_residual


class LinearIteration:
    damp = 1.0


class Jacobi(LinearIteration):
    def __init__(self, damp=1.0):
        self.damp = damp


class SOR(LinearIteration):
    """<sor> swept in (colour, index) order: femlisp's :compare ordering."""

    def __init__(self, omega=1.0):
        self.omega = omega


class GaussSeidel(SOR):
    def __init__(self):
        super().__init__(1.0)


class GeometricSSC(LinearIteration):
    """<geometric-ssc>: successive subspace correction over femlisp's vertex-centred blocks
    (a mesh vertex and the non-Dirichlet DOFs of the cells around it), swept by colour; per
    block the residual is updated and x_B += omega A_BB^-1 r_B.  Lattice matrices only."""

    def __init__(self, omega=1.0):
        self.omega = omega


class GeometricPSC(LinearIteration):
    """<geometric-psc>: the same blocks corrected additively from one residual,
    x += damp sum_B A_BB^-1 r_B.  The blocks overlap, so damp must be small (about
    2^-dim)."""

    def __init__(self, damp=1.0):
        self.damp = damp


def geometric_ssc(**kw):
    """femlisp's (geometric-ssc ...) constructor (geoblock.lisp:57-59)."""
    return GeometricSSC(**kw)


def geometric_psc(**kw):
    """femlisp's (geometric-psc ...) constructor (geoblock.lisp:53-55)."""
    return GeometricPSC(**kw)


def select_smoother(order):
    """select-smoother (geomg.lisp:58-71): <gauss-seidel> for order 1, <geometric-ssc>
    for higher order."""
    return GaussSeidel() if order == 1 else GeometricSSC()


class LU(LinearIteration):
    """<lu>: the coarse-grid iteration and the fallback of a SafeLinearSolver
    (dense FP64 inverse formed on the GPU)."""


class GeometricCS(LinearIteration):
    """<geometric-cs>: correction scheme, :multiplicative or :additive
    combination (multigrid.lisp:67-70,262-287); gamma coarse cycles per visit
    (1 V-cycle, 2 W-cycle, multigrid.lisp:268-273).

    levels: either a lattice description dict(dim=, order=, n0=, nlevels=,
    kappa=, mode=) assembled on the GPU, or explicit (A_list, I_list) of
    GpuSparseMatrix coarsest first (femlisp's :a-vec/:i-vec).

    Without a smoother, lattice hierarchies take select_smoother(order) and explicit
    matrices <gauss-seidel>.
    """

    def __init__(self, smoother=None, pre_steps=1, post_steps=1, base_level=0, gamma=1,
                 symmetric=False, prec=F64, combination="multiplicative", galerkin=False,
                 transfer="geometric"):
        if combination not in ("multiplicative", "additive"):
            raise ValueError("combination must be 'multiplicative' or 'additive'")
        self.combination = combination
        if gamma < 0:
            raise ValueError("gamma must be >= 0")
        self.gamma = gamma
        self.smoother = smoother
        self.pre_steps, self.post_steps = pre_steps, post_steps
        self.base_level = base_level
        self.symmetric = symmetric
        self.prec = prec
        self.galerkin = galerkin
        if transfer not in ("geometric", "blackbox"):
            raise ValueError("transfer must be 'geometric' or 'blackbox'")
        self.transfer = transfer

    def params(self, order=1):
        sm = self.smoother or select_smoother(order)
        if isinstance(sm, Jacobi):
            kind, omega = _f.SMOOTH_JACOBI, sm.damp
        elif isinstance(sm, SOR):
            kind, omega = _f.SMOOTH_SOR, sm.omega
        elif isinstance(sm, GeometricSSC):
            kind, omega = _f.SMOOTH_SSC, sm.omega
        elif isinstance(sm, GeometricPSC):
            kind, omega = _f.SMOOTH_PSC, sm.damp
        else:
            raise TypeError(f"no GPU smoother for {type(sm).__name__}")
        return MgParams(kind, omega, self.pre_steps, self.post_steps, self.base_level,
                        int(self.symmetric), self.prec, int(self.galerkin),
                        _f.TRANSFER_BLACKBOX if self.transfer == "blackbox"
                        else _f.TRANSFER_GEOMETRIC, int(self.gamma),
                        int(self.combination == "additive"), int(self.fas))

    fas = False


class GeometricFAS(GeometricCS):
    """<geometric-fas> (geomg.lisp:174-194): Brandt's FAS, the same cycle as GeometricCS
    except that every coarse level approximates the solution: it starts from R x (the
    projection-matrix, injection of the nodal functionals) with rhs I^T r + A_c R x and
    prolongates x_c - R x (multigrid.lisp:226-247).  Lattice hierarchies only."""

    fas = True


def geometric_cs(**kw):
    """femlisp's (geometric-cs ...) constructor (geomg.lisp:152-155)."""
    return GeometricCS(**kw)


def fas(**kw):
    """femlisp's (fas ...) constructor (geomg.lisp:180-182)."""
    return GeometricFAS(**kw)


class CG(LinearIteration):
    """<cg> (krylow.lisp:67-132); restart_cycle as femlisp's slot: every
    restart_cycle steps the residual is recomputed and the recurrence restarts."""

    def __init__(self, preconditioner=None, alias=False, restart_cycle=None):
        self.preconditioner = preconditioner
        self.alias = alias
        self.restart_cycle = restart_cycle


class FCG(CG):
    pass


class Hierarchy:
    """flgpu_mg: the multilevel decomposition of a <geometric-cs>."""

    def __init__(self, handle, ctx, keep=()):
        self.h = handle
        self.ctx = ctx
        self._keep = keep

    @classmethod
    def from_matrices(cls, mgit, A_list, I_list):
        ctx = A_list[0].ctx
        n = len(A_list)
        Aarr = (ctypes.c_void_p * n)(*[a.h.value if hasattr(a.h, "value") else a.h for a in A_list])
        Iarr = (ctypes.c_void_p * max(n - 1, 1))(
            *[i.h.value if hasattr(i.h, "value") else i.h for i in I_list] or [None])
        h = ctypes.c_void_p()
        prm = mgit.params()
        check(lib().flgpu_mg_create(ctx.h, n, Aarr, Iarr, ctypes.byref(prm), ctypes.byref(h)))
        return cls(h, ctx)

    @classmethod
    def lattice(cls, mgit, dim, order, n0, nlevels, kappa=None, f=1.0, mode=_f.ASM_F64,
                ctx=None):
        from .lattice import _kappa_vec
        from .matlisp import default_context
        ctx = ctx or default_context()
        kv, nk = _kappa_vec(kappa, ctx)
        h, hb = ctypes.c_void_p(), ctypes.c_void_p()
        prm = mgit.params(order)
        check(lib().flgpu_mg_create_lattice(ctx.h, dim, order, n0, nlevels, kv.h if kv else None,
                                            nk, f, mode, ctypes.byref(prm), ctypes.byref(h),
                                            ctypes.byref(hb)))
        self = cls(h, ctx, keep=(kv,))
        self.rhs = GpuVector.wrap(hb, ctx)
        return self

    def top_matrix(self):
        """The hierarchy's top-level matrix (owned by the hierarchy, which the
        returned object keeps alive)."""
        h = ctypes.c_void_p()
        check(lib().flgpu_mg_top_matrix(self.h, ctypes.byref(h)))
        m = GpuSparseMatrix(h, self.ctx, owner=False)
        m._parent = self
        return m

    def level_matrix(self, level, interp=False):
        """the level's matrix (interp False / 0), its interpolation to the next level
        (True / 1) or, in a FAS hierarchy, its projection from the next level (2)"""
        h = ctypes.c_void_p()
        check(lib().flgpu_mg_level_matrix(self.h, level, int(interp), ctypes.byref(h)))
        m = GpuSparseMatrix(h, self.ctx, owner=False)
        m._parent = self
        return m

    def apply(self, x, b, r):
        check(lib().flgpu_mg_apply(self.h, x.h, b.h, r.h))
        return x

    def reassemble(self, kappa):
        """New coefficient on every level of a lattice hierarchy (same pattern;
        transfers, Galerkin levels and the coarse inverse follow)."""
        from .lattice import _kappa_vec
        kv, nk = _kappa_vec(kappa, self.ctx)
        check(lib().flgpu_mg_reassemble(self.h, kv.h if kv else None, nk))
        self._keep = tuple(self._keep) + (kv,)
        return self

    def use_graphs(self, enable=True):
        check(lib().flgpu_mg_use_graphs(self.h, int(enable)))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().flgpu_mg_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Iterator:
    """<iterator> (src/iteration/linit.lisp:48-63): iterate(x, b, r)."""

    def __init__(self, matrix, iterate, residual_before, residual_after, initialize=None):
        self.matrix = matrix
        self.iterate = iterate
        self.residual_before = residual_before
        self.residual_after = residual_after
        self.initialize = initialize


def _product(mat, step, steps, residual_before, residual_after):
    """product-iterator (src/iteration/linit.lisp:94-120)."""
    def it(x, b, r):
        for _ in range(steps):
            step(x, b, r)
            if steps > 1 and residual_before and not residual_after:
                compute_residual(mat, x, b, r)
        return x
    return it


def make_iterator(linit, mat, steps=1, hierarchy=None):
    if isinstance(linit, Jacobi):
        def step(x, b, r):
            check(lib().flgpu_jacobi_iterate(mat.h, linit.damp, x.h, r.h))
        return Iterator(mat, _product(mat, step, steps, True, False), True, False)
    if isinstance(linit, SOR):
        def step(x, b, r):
            check(lib().flgpu_sor_iterate(mat.h, linit.omega, 0, x.h, b.h, r.h))
        return Iterator(mat, _product(mat, step, steps, False, False), False, False)
    if isinstance(linit, (GeometricSSC, GeometricPSC)):
        ssc = isinstance(linit, GeometricSSC)
        kind = _f.SMOOTH_SSC if ssc else _f.SMOOTH_PSC
        w = linit.omega if ssc else linit.damp

        def step(x, b, r):
            check(lib().flgpu_block_iterate(mat.h, kind, w, 0, x.h, b.h, r.h))
        # <block-iteration> iterators: residual-before t, residual-after nil (blockit.lisp:125,203)
        return Iterator(mat, _product(mat, step, steps, True, False), True, False)
    if isinstance(linit, GeometricCS):
        if hierarchy is None:
            raise ValueError("a <geometric-cs> iterator needs its hierarchy")
        return Iterator(mat, lambda x, b, r: hierarchy.apply(x, b, r), True, False)
    raise TypeError(f"no GPU iterator for {type(linit).__name__}")


class LinearSolver:
    def __init__(self, iteration, success_if=("<", "reduction", 1e-10),
                 failure_if=(">", "step", 100), true_residual=False, learned=False,
                 stall_window=0):
        self.iteration = iteration
        self.success_if = success_if
        self.failure_if = failure_if
        self.true_residual = true_residual
        self.learned = learned
        self.stall_window = stall_window

    def params(self):
        it = self.iteration
        _, key, tol = self.success_if
        stop = _f.STOP_REDUCTION if key == "reduction" else _f.STOP_DEFNORM
        maxsteps = int(self.failure_if[2]) if self.failure_if else 1 << 30
        if isinstance(it, FCG):
            outer, alias = _f.OUTER_FCG, 0
        elif isinstance(it, CG):
            outer = _f.OUTER_CG if it.preconditioner is not None else _f.OUTER_CG_PLAIN
            alias = int(it.alias)
        elif isinstance(it, GeometricCS):
            outer, alias = _f.OUTER_MG, 0
        else:
            raise TypeError(f"unsupported iteration {type(it).__name__}")
        restart = int(getattr(it, "restart_cycle", None) or 0)
        fallback = self.fallback_kind()
        return SolveParams(outer, alias, stop, float(tol), maxsteps, int(self.true_residual),
                           int(self.learned), int(self.stall_window), restart, fallback)

    def fallback_kind(self):
        return _f.FALLBACK_NONE


class SafeLinearSolver(LinearSolver):
    """<safe-linear-solver> (linsolve.lisp:45-56,75-89): on failure the iteration
    becomes the fallback, <lu> by default, and the loop goes on; a second failure
    is an error (status "fallback-failed")."""

    def __init__(self, iteration, fallback=None, **kw):
        super().__init__(iteration, **kw)
        self.fallback = LU() if fallback is None else fallback

    def fallback_kind(self):
        if isinstance(self.fallback, LU):
            return _f.FALLBACK_LU
        raise TypeError(f"no GPU fallback for {type(self.fallback).__name__}")


def solve(solver, blackboard):
    """solve on a blackboard with :matrix, :rhs and (for multigrid) :hierarchy.

    Returns the blackboard with :solution, :status, :steps, :defnorms and a
    :report holding :convergence-rate = reduction^(1/steps)
    (src/iteration/solve.lisp:157-165).
    """
    A = blackboard["matrix"]
    b = blackboard["rhs"]
    mg = blackboard.get("hierarchy")
    x = blackboard.get("solution") or A.make_domain_vector()
    prm = solver.params()
    rep = SolveReport()
    cap = prm.maxsteps + 3 if prm.maxsteps < (1 << 20) else 100000
    dn = np.zeros(cap)
    st = lib().flgpu_solve(mg.h if mg else None, A.h, ctypes.byref(prm), b.h, x.h,
                           ctypes.byref(rep), ptr(dn), cap)
    if st < 0:
        check(st)
    steps = rep.steps
    red = rep.final_defnorm / rep.initial_defnorm if rep.initial_defnorm else 0.0
    blackboard.update({
        "solution": x,
        "status": {0: "success", 1: "failure"}.get(st, "fallback-failed"),
        "fallback_used": bool(rep.fallback_used),
        "steps": steps,
        "defnorms": dn[: min(steps + 1, cap)].copy(),
        "defnorm": rep.final_defnorm,
        "true_residual": rep.true_residual_norm,
        "fell_back": bool(rep.fell_back),
        "seconds": rep.seconds,
        "report": {"convergence-rate": red ** (1.0 / steps) if steps > 0 and red > 0 else None,
                   "steps": steps},
    })
    return blackboard
