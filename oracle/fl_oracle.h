/*
 * fl_oracle.h -- CPU restatement of femlisp's solver hot path for Lagrange
 * FE discretisations on uniformly refined box meshes.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the CPU
 * baseline; it is never linked into, loaded by or called from the product
 * library (libflgpu.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Every function cites the femlisp file:line whose semantics it restates.
 * Arithmetic is IEEE double throughout, as in femlisp (<sparse-matrix> is
 * double-float, src/matlisp/sparse-matrix.lisp:56-60).
 *
 * Canonical DOF numbering: femlisp keys DOFs by mesh cells and has no integer
 * index; the canonical index used here is the rank of the DOF's midpoint in
 * femlisp's sort-lexicographically order (src/mesh/meshgen.lisp:200-217,
 * coordinate 0 most significant), i.e. the lattice index
 *   idx = i_0 * M^(d-1) + ... + i_{d-1},  M = order*n + 1 nodes per axis.
 */
#ifndef FL_ORACLE_H
#define FL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- lattice (src/mesh/meshgen.lisp:126-138,180-183; src/mesh/refine.lisp:148-149) */
long flo_lattice_ndofs(int dim, int order, int n);
long flo_lattice_nnz(int dim, int order, int n);
/* CSR pattern of femlisp's assembled matrix: every pair of DOFs sharing a
 * cell, explicit zeros of Dirichlet rows/columns kept, columns ascending.
 * rowptr[N+1], colidx[nnz]. Returns 0. */
int flo_lattice_pattern(int dim, int order, int n, int *rowptr, int *colidx);
/* 1 for DOFs on the external boundary (src/discretization/system-fe.lisp:37-129) */
int flo_lattice_boundary(int dim, int order, int n, unsigned char *is_bdry);

/* ---- quadrature and basis (src/discretization/quadrature.lisp:122-215,
 *      src/discretization/lagrange.lisp:41-202) */
int flo_gauss_rule(int s, double *pts, double *wts);
void flo_lagrange_1d(int order, double t, double *phi, double *dphi);

/* Coefficient sampling: kappa is piecewise constant on an nk^dim grid of
 * cells of the unit cube (NULL -> 1).  A quadrature point at global
 * coordinate (c + g) * H on a level with n cells per axis (H = 1/n) samples
 * the coefficient cell c*(nk/n) + floor(g*(nk/n)) along each axis; nk must be
 * a multiple of n.  See flo_assemble. */

/* Element matrix and load vector of one cell of width H for diffusion kappa*I
 * sampled per quadrature point (kappa_qp[s^dim]) and source f
 * (src/discretization/ellsys-fe.lisp:120-127,183-190,
 *  src/discretization/fe.lisp:502-535).  Local DOF order: lattice order inside
 * the cell (coordinate 0 most significant).  Ke[nloc*nloc], be[nloc]. */
int flo_element_matrix(int dim, int order, double H, const double *kappa_qp,
                       double f, double *Ke, double *be);

/* Global assembly on the level with n cells per axis, followed by Dirichlet
 * elimination A <- S A S + P, b <- S b (zero boundary values)
 * (src/discretization/fedisc.lisp:108-268, constraints.lisp:323-411).
 * vals[nnz] in flo_lattice_pattern order; rhs[N] may be NULL. */
int flo_assemble(int dim, int order, int n, const double *kappa, int nk,
                 double f, double *vals, double *rhs);

/* Constrained :local interpolation matrix from the level with nc cells per
 * axis to the level with 2*nc cells: rows fine DOFs, columns coarse DOFs,
 * entries below 1e-10 dropped, fine Dirichlet rows and coarse Dirichlet
 * columns zeroed but kept (src/discretization/sparseas.lisp:232-272,
 * fetransfer.lisp:73-138, constraints.lisp:58-72,
 * src/matlisp/sparse-matrix.lisp:635-653).  Call with colidx == NULL to get
 * nnz returned; rowptr is always filled. */
long flo_interpolation(int dim, int order, int nc, int *rowptr, int *colidx,
                       double *vals);

/* FAS projection (femlisp's projection-matrix, src/discretization/sparseas.lisp:278-297, with
 * compute-local-pmatrix, fetransfer.lisp:201-213: the injection of the dof functionals)
 * from the level with 2*nc cells per axis to the level with nc: rows coarse DOFs, one entry
 * 1.0 each, at the fine DOF of the same point (lattice index 2c per axis).  Returns nnz
 * (the coarse DOF count); rowptr, colidx, vals may be NULL. */
long flo_projection(int dim, int order, int nc, int *rowptr, int *colidx, double *vals);

/* Operator-dependent interpolation for 3D Q1 (not in femlisp; Dendy's
 * black-box prolongation for 27-point stencils): the values of the pattern of
 * flo_interpolation(3, 1, nf/2, prowptr, pcolidx, .) computed from the fine
 * matrix (nf cells per axis, lattice pattern).  A fine node with s odd
 * coordinates solves its row, collapsed onto its odd axes, for itself, using
 * the rows of nodes with fewer odd coordinates.  Fine Dirichlet rows and
 * coarse Dirichlet columns are zero.  Returns 0, or -1 for odd nf. */
int flo_blackbox_interpolation(int nf, const int *rowptr, const int *colidx,
                               const double *vals, const int *prowptr,
                               const int *pcolidx, double *pvals);

/* ---- algebra (src/matlisp/sparse-matrix.lisp:378-442) */
/* x <- beta*x + alpha*A*y  (gemm-nn!) */
void flo_spmv(int nrows, const int *rowptr, const int *colidx,
              const double *vals, double alpha, const double *y, double beta,
              double *x);
/* x <- beta*x + alpha*A^T*y  (gemm-tn!), A is nrows x ncols */
void flo_spmv_t(int nrows, int ncols, const int *rowptr, const int *colidx,
                const double *vals, double alpha, const double *y,
                double beta, double *x);
/* transpose CSR (ncols+1 rowptr of the result); columns ascending */
void flo_csr_transpose(int nrows, int ncols, const int *rowptr,
                       const int *colidx, const double *vals, int *trowptr,
                       int *tcolidx, double *tvals);

/* ---- block smoothers (src/iteration/blockit.lisp:41-203, geoblock.lisp:37-157) */

/* vertex-centred blocks of the lattice with n cells per axis: (n+1)^dim blocks of
 * W = (2 order - 1)^dim slots, keys[b*W + s] the DOF of vertex b's slot s (lattice offset
 * d - (order - 1), d in lexicographic order) or -1 (outside, or Dirichlet other than the
 * vertex); blocks in colour order (parity of the vertex coordinates).  keys may be NULL;
 * returns the number of blocks, -1 on bad arguments. */
long flo_vertex_blocks(int dim, int order, int n, int *keys);
/* one <geometric-ssc> (ssc = 1, omega) or <geometric-psc> (ssc = 0, damp = omega) iterate on
 * the lattice matrix (x, b, r as femlisp's iterator closure; <psc> needs r = b - A x),
 * blocks in flo_vertex_blocks order or reversed (backwards = 1). */
int flo_block_iterate(int nrows, const int *rowptr, const int *colidx, const double *vals,
                      int dim, int order, int n, int ssc, double omega, int backwards,
                      double *x, const double *b, double *r);

/* Host threads of the parallel loops (n <= 0: all cores); the results do not depend on it:
 * rows, one colour's <sor> rows and one colour's blocks run in parallel, each with its own
 * sequential arithmetic, and dots add fixed index ranges in a fixed order. */
void flo_set_threads(int n);
int flo_get_threads(void);

/* ---- multigrid hierarchy + solvers (src/iteration/ multigrid, linit, krylow) */
typedef struct flo_mg flo_mg;

/* Hierarchy on the unit cube with n0 cells per axis on level 0 and
 * `levels` levels (top level has n0*2^(levels-1) cells per axis);
 * re-discretisation on every level (src/iteration/geomg.lisp:109-142). */
flo_mg *flo_mg_create(int dim, int order, int n0, int levels,
                      const double *kappa, int nk, double f);
/* as flo_mg_create; galerkin_p = 1 builds the coarse levels as P^T A P
 * (geomg.lisp:122-123 :galerkin-p t, amg.lisp:247-280) instead of
 * re-discretising them; galerkin_p = 2 (3D Q1 only) does the same with the
 * operator-dependent interpolation of flo_blackbox_interpolation, computed
 * from each finer level's matrix, top level first */
flo_mg *flo_mg_create_ex(int dim, int order, int n0, int levels,
                         const double *kappa, int nk, double f, int galerkin_p);
void flo_mg_destroy(flo_mg *mg);
long flo_mg_ndofs(const flo_mg *mg, int level);
long flo_mg_nnz(const flo_mg *mg, int level);
/* copies of the level matrix / interpolation (I_level: level -> level+1) */
int flo_mg_get_matrix(const flo_mg *mg, int level, int *rowptr, int *colidx,
                      double *vals);
long flo_mg_interp_nnz(const flo_mg *mg, int level);
int flo_mg_get_interp(const flo_mg *mg, int level, int *rowptr, int *colidx,
                      double *vals);
int flo_mg_get_rhs(const flo_mg *mg, double *rhs);
/* copy of the FAS projection R_level (level+1 -> level), flo_projection's layout */
int flo_mg_get_projection(const flo_mg *mg, int level, int *rowptr, int *colidx,
                          double *vals);

enum { FLO_OUTER_MG = 0, FLO_OUTER_CG = 1, FLO_OUTER_CG_PLAIN = 2,
       FLO_OUTER_FCG = 3 };
enum { FLO_SMOOTH_JACOBI = 0, FLO_SMOOTH_SOR = 1, FLO_SMOOTH_SSC = 2, FLO_SMOOTH_PSC = 3 };
enum { FLO_ORDER_INDEX = 0, FLO_ORDER_COLOUR = 1 };
enum { FLO_STOP_REDUCTION = 0, FLO_STOP_DEFNORM = 1 };

typedef struct flo_params {
  int outer;        /* FLO_OUTER_* */
  int smoother;     /* FLO_SMOOTH_* */
  double omega;     /* <jacobi> damp / <sor> omega / <geometric-ssc> omega /
                       <geometric-psc> damp */
  int ordering;     /* FLO_ORDER_* for <sor> */
  int pre_steps, post_steps;
  int base_level;
  int symmetric;    /* 1: post-smoothing sweeps the ordering backwards */
  int alias;        /* 1: preconditioner called as (iterate q w w) */
  int stop;         /* FLO_STOP_* */
  double tol;
  int maxsteps;     /* failure-if (> :step maxsteps) */
  int gamma;        /* coarse cycles per visit: 1 V-cycle, 2 W-cycle, 0 smoothing only
                       (<mg-iteration> gamma, multigrid.lisp:60,268-273) */
  int additive;     /* 0 :multiplicative, 1 :additive combination (multigrid.lisp:67-70,
                       262-287): restrict first, then smooth; post-smooth, then prolongate */
  int fas;          /* 0 correction scheme, 1 Brandt's FAS (<geometric-fas>, geomg.lisp:174-194;
                       restrict/prolongate methods multigrid.lisp:226-247) */
  int restart_cycle; /* <cg> restart-cycle (krylow.lisp:68-70,101-105,128-131): 0 never, else
                       every restart_cycle steps the residual is recomputed from x and the
                       recurrence restarts from it */
  int fallback;     /* 0 <linear-solver>; FLO_FALLBACK_LU: <safe-linear-solver>
                       (linsolve.lisp:45-56,75-89) whose fallback iteration, taken on failure,
                       is <lu> (linit.lisp:132-162) on the top-level matrix */
} flo_params;

enum { FLO_FALLBACK_NONE = 0, FLO_FALLBACK_LU = 1 };

/* Solve A x = b from x = 0 with femlisp's <linear-solver> loop
 * (src/iteration/iterate.lisp:239-250, solve.lisp:126-165,
 * linsolve.lisp:58-73).  defnorms[0..steps] receives ||r_k||_2 (the norm
 * femlisp observes: recurrence residual for CG, true residual otherwise).
 * Returns 0 on success, 1 on failure (step limit), 2 when the fallback
 * iteration of a <safe-linear-solver> has failed too (femlisp signals an
 * error, linsolve.lisp:84-85), <0 on argument errors. */
int flo_mg_solve(flo_mg *mg, const flo_params *p, const double *b, double *x,
                 int *steps, double *defnorms);

/* One application of the preconditioner as femlisp's CG performs it:
 * x <- 0, then the V-cycle iterate on (x, b, r) where r is b itself when
 * p->alias, else a separate copy of b.  Used to probe the operator. */
int flo_mg_apply(flo_mg *mg, const flo_params *p, const double *b, double *x);

#ifdef __cplusplus
}
#endif
#endif
