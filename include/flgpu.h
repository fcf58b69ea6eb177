/*
 * flgpu.h -- C-ABI of libflgpu.so, the MI355X (gfx950) implementation of
 * femlisp's solver hot path for Lagrange FE discretisations.
 *
 * Conventions (those of femlisp's existing foreign bindings, e.g.
 * interface/superlu.c:51-88 reached through src/alien/superlu.lisp:57-75 and
 * fl.port:def-function / foreign-call, src/basic/port.lisp:401-547):
 *   - every entry point returns an int status: 0 = ok, < 0 = argument or
 *     runtime error, > 0 = numerical failure; flgpu_last_error() returns the
 *     message of the last failing call on the calling thread;
 *   - host arrays are caller-owned and only valid during the call (femlisp
 *     pins them for the call's duration only): the library copies them and
 *     never keeps host pointers;
 *   - host values are always double (femlisp's double-float arrays), indices
 *     int32 (src/matlisp/compressed.lisp:51-53); device storage may be FP64
 *     or FP32 (FLGPU_F64 / FLGPU_F32);
 *   - device memory is owned by opaque handles that the caller destroys;
 *   - calls on distinct contexts may run concurrently from any thread; calls
 *     on one context are serialised by the context's lock and run in order
 *     on the context's HIP stream.
 *
 * Vector arguments follow femlisp's argument order: gemm!-style calls take
 * (alpha A y beta x) and update x; iterators take (x b r) and update x (and
 * r where femlisp does).
 */
#ifndef FLGPU_H
#define FLGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define FLGPU_F64 0
#define FLGPU_F32 1

typedef struct flgpu_ctx flgpu_ctx;
typedef struct flgpu_vec flgpu_vec;
typedef struct flgpu_mat flgpu_mat;
typedef struct flgpu_mg flgpu_mg;
typedef struct flgpu_unet flgpu_unet;

/* ---- errors / context ---------------------------------------------------- */

/* Message of the last failing call on this thread (cf. the int status that
 * c_superlu returns, interface/superlu.c:51-88). */
int flgpu_last_error(char *buf, int len);

/* Context = one device + one HIP stream + a lock.  Replaces femlisp's
 * per-call foreign library state (src/alien/alien.lisp:42-48). */
int flgpu_ctx_create(int device, flgpu_ctx **out);
int flgpu_ctx_destroy(flgpu_ctx *ctx);
int flgpu_ctx_synchronize(flgpu_ctx *ctx);
/* name of the device, number of CUs, bytes of device memory */
int flgpu_ctx_info(flgpu_ctx *ctx, char *name, int len, int *cus, long *mem);
/* HIP events on the context's stream around a caller's region of work:
 * start records the first event, stop records the second, waits for it and
 * returns the elapsed milliseconds (femlisp's measure-time-for-block,
 * src/basic/mflop.lisp:89-96, for device work). */
int flgpu_ctx_timer_start(flgpu_ctx *ctx);
int flgpu_ctx_timer_stop(flgpu_ctx *ctx, double *ms);

/* ---- vectors (src/matlisp/vector.lisp:91-152,224-326;
 *      src/matlisp/sparse-vector.lisp:248-280) -------------------------------- */

/* zero-initialised, like make-domain-vector-for (src/matlisp/matrix.lisp:337) */
int flgpu_vec_create(flgpu_ctx *ctx, long n, int prec, flgpu_vec **out);
int flgpu_vec_destroy(flgpu_vec *v);
int flgpu_vec_info(const flgpu_vec *v, long *n, int *prec);
int flgpu_vec_upload(flgpu_vec *v, const double *host);
int flgpu_vec_download(const flgpu_vec *v, double *host);
/* fill! / x<-0 */
int flgpu_vec_fill(flgpu_vec *v, double value);
/* copy! x y : y <- x (converts precision) */
int flgpu_vec_copy(const flgpu_vec *x, flgpu_vec *y);
/* scal! a x : x <- a x */
int flgpu_vec_scal(double a, flgpu_vec *x);
/* axpy! a x y : y <- a x + y */
int flgpu_vec_axpy(double a, const flgpu_vec *x, flgpu_vec *y);
/* dot x y; deterministic fixed-order reduction in FP64 */
int flgpu_vec_dot(const flgpu_vec *x, const flgpu_vec *y, double *out);
/* norm x 2 / l2-norm */
int flgpu_vec_norm2(const flgpu_vec *x, double *out);
/* flgpu_vec_dot without any allocation: no femlisp counterpart; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).  The same
 * kernels, grid and result as flgpu_vec_dot, for any length; the 1024 per-block partials and
 * the result go to a fixed 8.2 KB array each device holds for the life of the process, not
 * to the context's buffers.  Calls that overlap on one device (from two contexts) race on
 * that array. */
int flgpu_vec_dot_noalloc(const flgpu_vec *x, const flgpu_vec *y, double *out);

/* ---- sparse matrices (src/matlisp/sparse-matrix.lisp:378-442) ------------ */

/* Upload a CSR matrix (rows ascending, columns ascending within a row), e.g.
 * the result of femlisp's sparse-matrix->ccs applied to the transpose
 * (src/matlisp/sparselu.lisp:226-333) with keys numbered by
 * sort-lexicographically (src/mesh/meshgen.lisp:214-217).  prec selects the
 * device storage.  Explicit zeros are kept. */
int flgpu_csr_upload(flgpu_ctx *ctx, int nrows, int ncols, long nnz,
                     const int *rowptr, const int *colidx, const double *vals,
                     int prec, flgpu_mat **out);
/* replace the values of an existing pattern (vals[nnz] in CSR order) */
int flgpu_csr_update_values(flgpu_mat *A, const double *vals);
/* any of the output arrays may be NULL */
int flgpu_csr_download(const flgpu_mat *A, int *rowptr, int *colidx,
                       double *vals);
int flgpu_mat_info(const flgpu_mat *A, int *nrows, int *ncols, long *nnz,
                   int *prec);
/* copy of A stored in another precision */
int flgpu_mat_convert(const flgpu_mat *A, int prec, flgpu_mat **out);
/* Storage the kernels stream (the CSR is always kept):
 *   FLGPU_STORAGE_CELL     3D Q1 lattice matrix with a coefficient constant
 *                          per cell: A applied from one factor per cell
 *                          (A_ij = Kref-entry x sum of the factors of the cells
 *                          shared by i and j), no matrix values streamed;
 *   FLGPU_STORAGE_STENCIL  3D Q1 lattice matrix: the 27 values of each row
 *                          (SoA), columns implied by the lattice;
 *   FLGPU_STORAGE_SELL     any matrix: SELL-64 values and column indices.
 * Lattice matrices get the richest form that applies.  max_form caps it
 * (forms are only ever dropped); *form receives the form in use.  New
 * values drop the per-cell form; a caller's colouring drops both lattice
 * forms (their sweeps follow the lattice colouring). */
#define FLGPU_STORAGE_SELL 0
#define FLGPU_STORAGE_STENCIL 1
#define FLGPU_STORAGE_CELL 2
int flgpu_mat_storage(flgpu_mat *A, int max_form, int *form);
/* Value planes the stencil-form kernels stream per row (no femlisp counterpart: byte
 * accounting for measurements): 27, or 14 when the stencil is exactly symmetric
 * (A(i, i + d) == A(i + d, i)) and the 13 lower coefficients are read from the mirrored upper
 * planes of the neighbour rows; 0 when A is not in the stencil form. */
int flgpu_mat_stencil_planes(const flgpu_mat *A, int *planes);
int flgpu_mat_destroy(flgpu_mat *A);

/* gemm-nn! alpha A y beta x : x <- beta x + alpha A y
 * (sparse-matrix.lisp:393-405; src/matlisp/matrix.lisp:213-240). */
int flgpu_gemm_nn(double alpha, const flgpu_mat *A, const flgpu_vec *y,
                  double beta, flgpu_vec *x);
/* gemm-tn! alpha A y beta x : x <- beta x + alpha A^T y
 * (sparse-matrix.lisp:406-433); the transpose is built on first use. */
int flgpu_gemm_tn(double alpha, const flgpu_mat *A, const flgpu_vec *y,
                  double beta, flgpu_vec *x);
/* galerkin-product (src/iteration/amg.lisp:247-280): the values of Ac
 * become P^T A P, accumulated in FP64 into Ac's existing pattern (an entry
 * of the product outside that pattern is an error).  P is fine x coarse. */
int flgpu_galerkin_product(const flgpu_mat *P, const flgpu_mat *A,
                           flgpu_mat *Ac);
/* compute-residual A x b r : r <- b - A x; b may be the same vector as r, in
 * which case r <- r - A x exactly as femlisp's copy!+gemm! does
 * (src/iteration/linit.lisp:73-80). */
int flgpu_compute_residual(const flgpu_mat *A, const flgpu_vec *x,
                           const flgpu_vec *b, flgpu_vec *r);

/* ---- smoothers (src/iteration/linit.lisp) ----------------------------------- */

/* Ordering of <sor> sweeps: rows sorted by (colour, index), i.e. <sor> with a
 * :compare function (linit.lisp:258-263,293-295).  Rows of one colour must
 * not couple.  colour[nrows] in [0, ncolours). */
int flgpu_mat_set_colouring(flgpu_mat *A, int ncolours, const int *colour);
/* greedy colouring of the pattern (rows visited in index order) when the
 * caller has none; returns the number of colours in *ncolours. */
int flgpu_mat_colour_greedy(flgpu_mat *A, int *ncolours);

/* one <jacobi> iterate (linit.lisp:220-250): r_i <- r_i / A_ii (gesv! in
 * place), x_i <- x_i + damp r_i.  Needs an up-to-date residual. */
int flgpu_jacobi_iterate(const flgpu_mat *A, double damp, flgpu_vec *x,
                         flgpu_vec *r);
/* `steps` <jacobi> steps as femlisp's product-iterator runs them for a
 * smoother (linit.lisp:94-120: residual-before t, residual-after nil): each
 * step recomputes r = b - A x and sets x <- x + damp r_i / A_ii, in one pass
 * over A, out of place through tmp (n values; x, b, tmp distinct).  The
 * residual is not returned. */
int flgpu_jacobi_steps(const flgpu_mat *A, double damp, int steps, flgpu_vec *x,
                       const flgpu_vec *b, flgpu_vec *tmp);
/* one <sor> sweep (linit.lisp:288-323) in colour order (backwards = 1
 * sweeps the colours last to first): per row corr = b_i - sum_j A_ij x_j,
 * corr /= A_ii, r_i <- corr, x_i += omega corr.  b may be r. */
int flgpu_sor_iterate(const flgpu_mat *A, double omega, int backwards,
                      flgpu_vec *x, const flgpu_vec *b, flgpu_vec *r);

/* ---- block smoothers (src/iteration/blockit.lisp:41-203,
 *      src/iteration/geoblock.lisp:37-157) --------------------------------- */

#define FLGPU_SMOOTH_SSC 2 /* <geometric-ssc>: successive subspace correction */
#define FLGPU_SMOOTH_PSC 3 /* <geometric-psc>: parallel subspace correction */

/* One iterate of femlisp's vertex-centred block smoother on a lattice matrix
 * (flgpu_lattice_assemble, a level of a lattice hierarchy, or a caller's matrix in lattice
 * numbering after flgpu_lattice_colouring): the block of a
 * mesh vertex v holds v and every DOF of a cell containing v that is not Dirichlet
 * (find-vertex-centered-block, geoblock.lisp:65-91) -- one DOF for Q1, 3^dim for Q2.
 * Blocks are swept by colour, the parity of the vertex coordinates (axis 0 most
 * significant; backwards = 1: last colour first), which is <ssc> with an :ordering slot that
 * sorts by colour; blocks of one colour do not couple, so the order inside a colour does
 * not matter.  Per block B (blockit.lisp:113-203, store-p t: inverses formed once):
 *   FLGPU_SMOOTH_SSC: r_B <- b_B - (A x)_B, x_B <- x_B + omega A_BB^-1 r_B;
 *   FLGPU_SMOOTH_PSC: x_B <- x_B + omega A_BB^-1 r_B (omega = <psc>'s damp) with the
 *   residual r = b - A x the caller provides (residual-before t); r is not written.
 * b may be r (femlisp's (iterate q w w)); x must differ from both. */
int flgpu_block_iterate(const flgpu_mat *A, int kind, double omega, int backwards,
                        flgpu_vec *x, const flgpu_vec *b, flgpu_vec *r);

/* ---- structured lattices: GPU assembly (src/discretization/fedisc.lisp:
 *      73-268, ellsys-fe.lisp:68-218, fe.lisp:235-244,502-535,
 *      constraints.lisp:323-411, sparseas.lisp:232-272) ------------------ */

/* Assemble -div(kappa grad u) = f with Lagrange Q`order` elements on the unit
 * `dim`-cube with n cells per axis, zero Dirichlet boundary, femlisp's full
 * sparsity pattern (all DOF pairs sharing a cell, Dirichlet rows/columns kept
 * as explicit zeros, identity diagonal), canonical lattice numbering.
 * kappa: NULL (kappa = 1) or a device vector of nk^dim per-cell values on a
 * grid of nk cells per axis (nk a multiple of n; a quadrature point samples
 * the coefficient cell it lies in, the cell on the upper side when it lies on
 * a face).  mode: FLGPU_ASM_F64 (FP64 FMA element kernel) or FLGPU_ASM_MFMA
 * (element contractions on MFMA: FP64 MFMA for prec F64, BF16 inputs with
 * FP32 accumulation for prec F32).  rhs may be NULL. */
#define FLGPU_ASM_F64 0
#define FLGPU_ASM_MFMA 1
int flgpu_lattice_assemble(flgpu_ctx *ctx, int dim, int order, int n,
                           const flgpu_vec *kappa, int nk, double f, int prec,
                           int mode, flgpu_mat **A, flgpu_vec **rhs);
/* Re-assemble the values of an existing lattice matrix for a new kappa. */
int flgpu_lattice_reassemble(flgpu_mat *A, const flgpu_vec *kappa, int nk,
                             int mode);
/* Constrained :local interpolation from the lattice with nc cells per axis to
 * the one with 2 nc cells (fetransfer.lisp:73-138, constraints.lisp:58-72):
 * rows fine DOFs, columns coarse DOFs; fine Dirichlet rows and coarse
 * Dirichlet columns zeroed but kept. */
int flgpu_lattice_interpolation(flgpu_ctx *ctx, int dim, int order, int nc,
                                int prec, flgpu_mat **P);
/* FAS projection from the lattice with 2 nc cells per axis to the one with nc
 * cells (femlisp's projection-matrix, sparseas.lisp:278-297, whose local matrix
 * for Lagrange elements is the injection of the dof functionals,
 * fetransfer.lisp:201-213): rows coarse DOFs, one entry 1 each, at the fine DOF
 * of the same point. */
int flgpu_lattice_projection(flgpu_ctx *ctx, int dim, int order, int nc, int prec,
                             flgpu_mat **R);
/* Operator-dependent ("black-box") interpolation for 3D Q1, no femlisp
 * counterpart: it takes the place of flgpu_lattice_interpolation's values
 * (femlisp's constrained :local interpolation, fetransfer.lisp:73-138) when
 * the coefficient jumps.  P must be a lattice interpolation from n/2 to n
 * cells per axis (its pattern), A the fine matrix on n cells (n even, lattice
 * pattern).  Dendy's black-box prolongation (J. Comput. Phys. 48 (1982)
 * 366-386) for 27-point stencils: a fine node with s odd coordinates solves
 * its row of A, summed onto its odd axes, for itself, using the rows of nodes
 * with fewer odd coordinates; fine Dirichlet rows and coarse Dirichlet
 * columns are zero.  For a constant coefficient it equals femlisp's trilinear
 * interpolation. */
int flgpu_lattice_blackbox_interpolation(const flgpu_mat *A, flgpu_mat *P);
/* lattice colouring for <sor>: coordinate parity (Q1) / coordinate mod 3 (Q2).  A caller's
 * matrix (flgpu_csr_upload, flgpu_system_read) in canonical lattice numbering is recorded as
 * that lattice's matrix, which the block smoothers (flgpu_block_iterate) need; a matrix of
 * another lattice is rejected. */
int flgpu_lattice_colouring(flgpu_mat *A, int dim, int order, int n);
/* The element kernel alone (discretize-locally for a batch of cells,
 * ellsys-fe.lisp:68-127): element matrices of cells first_cell ..
 * first_cell+ncells-1 (lattice cell order), local DOFs in lattice order
 * inside the cell, out[ncells][nloc][nloc] as double. */
int flgpu_lattice_element_matrices(flgpu_ctx *ctx, int dim, int order, int n,
                                   const flgpu_vec *kappa, int nk, int prec,
                                   int mode, long first_cell, long ncells,
                                   double *out);

/* ---- system files (next to femlisp's export sparse-matrix->ccs,
 *      src/matlisp/sparselu.lisp:235-333) --------------------------------- */

/* Read a system file: magic "FLSYS01\0", int32 orientation (0 rows / 1 columns
 * compressed), int32 flags (1: rhs follows, 2: key permutation follows), int64
 * nrows, ncols, nnz, int32 starts[major + 1], int32 indices[nnz], float64
 * values[nnz], [float64 rhs[nrows]], [int32 perm[nrows]] (little endian).
 * perm[k] is the canonical index of the file's row and column k (the rank of
 * its DOF midpoint in femlisp's sort-lexicographically order,
 * src/mesh/meshgen.lisp:214-217).  The matrix is returned in canonical
 * numbering as CSR with ascending columns; *rhs (if rhs is not NULL) receives
 * the permuted right-hand side or NULL when the file has none. */
int flgpu_system_read(flgpu_ctx *ctx, const char *path, int prec, flgpu_mat **A,
                      flgpu_vec **rhs);
/* Write A (rows compressed, no permutation) and, if not NULL, rhs. */
int flgpu_system_write(const char *path, const flgpu_mat *A, const flgpu_vec *rhs);

/* ---- geometric multigrid (src/iteration/multigrid.lisp:55-393,
 *      geomg.lisp:109-142) ------------------------------------------------- */

#define FLGPU_SMOOTH_JACOBI 0
#define FLGPU_SMOOTH_SOR 1

typedef struct flgpu_mg_params {
  int smoother;    /* FLGPU_SMOOTH_JACOBI (damped), FLGPU_SMOOTH_SOR (colour order), or the
                      block smoothers FLGPU_SMOOTH_SSC / FLGPU_SMOOTH_PSC (lattice matrices;
                      femlisp's default for order > 1, geomg.lisp:58-71) */
  double omega;    /* <jacobi> damp / <sor> omega / <ssc> omega / <psc> damp */
  int pre_steps;   /* product-iterator steps, linit.lisp:94-120 */
  int post_steps;
  int base_level;  /* coarse-grid solve level: dense FP64 LU/inverse */
  int symmetric;   /* 1: post-smoothing sweeps the colours backwards */
  int prec;        /* storage/arithmetic of the V-cycle: FLGPU_F64 / FLGPU_F32 */
  int galerkin;    /* lattice hierarchies: 1 = coarse levels P^T A P
                      (geomg.lisp:122-123 :galerkin-p t), 0 = re-discretised */
  int transfer;    /* lattice hierarchies: FLGPU_TRANSFER_GEOMETRIC (femlisp's
                      interpolation) or FLGPU_TRANSFER_BLACKBOX (3D Q1:
                      operator-dependent interpolation from each finer level's
                      matrix, coarse levels P^T A P whatever galerkin says) */
  int gamma;       /* coarse cycles per visit of a level (<mg-iteration> gamma,
                      multigrid.lisp:60,268-273): 1 V-cycle, 2 W-cycle, 0 smoothing
                      only; one when the coarse level is the base level */
  int additive;    /* 0 :multiplicative, 1 :additive combination (multigrid.lisp:
                      67-70,262-287): restrict first, then smooth the level; post-smooth,
                      then prolongate */
  int fas;         /* 0 correction scheme (<geometric-cs>), 1 Brandt's FAS
                      (<geometric-fas>, geomg.lisp:174-194): the coarse level starts
                      from R x with rhs I^T r + A_c R x and prolongates x_c - R x
                      (multigrid.lisp:226-247); R = flgpu_lattice_projection, so lattice
                      hierarchies only */
} flgpu_mg_params;

#define FLGPU_TRANSFER_GEOMETRIC 0
#define FLGPU_TRANSFER_BLACKBOX 1

/* Hierarchy from caller matrices: A[0..nlevels-1] level matrices (coarsest
 * first), I[0..nlevels-2] interpolations level l -> l+1 (rows = DOFs of
 * level l+1), exactly femlisp's :a-vec/:i-vec blackboard
 * (multigrid.lisp:139-141).  The matrices are copied (converted to
 * params->prec); the caller keeps ownership of its handles.  Matrices
 * without a colouring get a greedy one when the smoother is <sor>. */
int flgpu_mg_create(flgpu_ctx *ctx, int nlevels, flgpu_mat *const *A,
                    flgpu_mat *const *I, const flgpu_mg_params *params,
                    flgpu_mg **out);
/* Hierarchy assembled on the GPU on the unit cube: level l has n0 * 2^l cells
 * per axis, l = 0..nlevels-1; every level is re-discretised with the
 * coefficient sampled at its own quadrature points (geomg.lisp:122-126). */
int flgpu_mg_create_lattice(flgpu_ctx *ctx, int dim, int order, int n0,
                            int nlevels, const flgpu_vec *kappa, int nk,
                            double f, int asm_mode,
                            const flgpu_mg_params *params, flgpu_mg **out,
                            flgpu_vec **rhs);
int flgpu_mg_destroy(flgpu_mg *mg);
/* the top-level matrix of the hierarchy (owned by mg) */
int flgpu_mg_top_matrix(flgpu_mg *mg, flgpu_mat **A);
/* the matrix (I = 0), the interpolation to the next level (I = 1) or, with FAS,
 * the projection from the next level (I = 2) of one level, owned by mg (the
 * :a-vec / :i-vec / :fas-r-vec entries, multigrid.lisp:108-109,121) */
int flgpu_mg_level_matrix(flgpu_mg *mg, int level, int interp, flgpu_mat **A);
/* The <mg-iteration> iterator (multigrid.lisp:368-391): one V-cycle on
 * (x, b, r) with r = b - A x up to date; x is updated, r and b are used as
 * femlisp uses them (b may be the same vector as r: femlisp's <cg> calls
 * (iterate q w w), krylow.lisp:118-121).  Vectors may be FP64 with an FP32
 * hierarchy: then x <- x + V(r) with the cycle in FP32 (b must not be r). */
int flgpu_mg_apply(flgpu_mg *mg, flgpu_vec *x, flgpu_vec *b, flgpu_vec *r);
/* Replay cycles from HIP graphs (default 1): the first cycle on a vector
 * triple is captured and later cycles on it are one graph launch.  0 launches
 * every kernel separately. */
int flgpu_mg_use_graphs(flgpu_mg *mg, int enable);

/* Re-assemble every level of a lattice hierarchy for a new coefficient
 * (same pattern, transfers and colouring; coarse inverse recomputed). */
int flgpu_mg_reassemble(flgpu_mg *mg, const flgpu_vec *kappa, int nk);

/* ---- learned coarse correction (no femlisp counterpart) ------------------
 *
 * A 3-level 3D U-Net (16/32/64 channels, 3x3x3 kernels, BF16 weights, FP32
 * accumulation on MFMA; Adam) mapping a residual (and log kappa) on the n^3
 * grid of lattice nodes with coordinates < n to an error estimate.  Inside
 * flexible CG (flgpu_solve with learned = 1) the preconditioner becomes
 * z = V(r) + beta N(r - A V(r)), beta minimising the energy along N, when
 * that step adds at least 1% of the energy of the cycle's own step; the first
 * step where it does not goes on with z = V(r) and drops the network. */
int flgpu_unet_create(flgpu_ctx *ctx, int cin, unsigned long seed,
                      flgpu_unet **out);
int flgpu_unet_destroy(flgpu_unet *u);
int flgpu_unet_nparams(const flgpu_unet *u, long *np);
/* parameters and Adam moments as doubles (any of theta/m/v may be NULL on
 * get; m/v NULL on set resets them); the layout is per layer W[27][ci][co]
 * then bias[co] */
int flgpu_unet_get_params(const flgpu_unet *u, double *theta, double *m,
                          double *v, int *step);
int flgpu_unet_set_params(flgpu_unet *u, const double *theta, const double *m,
                          const double *v, int step);
/* x: FP32 [batch][n^3][cin], y: FP32 [batch][n^3] */
int flgpu_unet_forward(flgpu_unet *u, int batch, int n, const flgpu_vec *x,
                       flgpu_vec *y);
/* mean squared error against t and its gradient (grad[nparams], may be NULL) */
int flgpu_unet_loss_grad(flgpu_unet *u, int batch, int n, const flgpu_vec *x,
                         const flgpu_vec *t, double *loss, double *grad);
/* forward + backward + one Adam step; loss before the update */
int flgpu_unet_train_step(flgpu_unet *u, int batch, int n, const flgpu_vec *x,
                          const flgpu_vec *t, double lr, double *loss);
/* network input (target = 0: [r/rms(r), log(kappa)/4]) or target
 * (target = 1: e/rms(e)) of sample `sample` from a lattice vector of
 * (n+1)^3 values */
int flgpu_unet_pack(const flgpu_vec *r, const flgpu_vec *kappa, int n,
                    flgpu_vec *x, long sample, int target);
/* dst row k = src row idx[k] (rows of row_len values) */
int flgpu_vec_gather_rows(const flgpu_vec *src, long row_len, const long *idx,
                          int count, flgpu_vec *dst);
/* attach a trained network (and the n^3 cell coefficient it reads) */
int flgpu_mg_attach_unet(flgpu_mg *mg, flgpu_unet *u, const flgpu_vec *kappa,
                         int n);

/* ---- solvers (src/iteration/krylow.lisp:67-132, linsolve.lisp:58-73,
 *      solve.lisp:119-165, iterate.lisp:239-250) --------------------------- */

#define FLGPU_OUTER_MG 0       /* the V-cycle is the iteration */
#define FLGPU_OUTER_CG 1       /* <cg> with the V-cycle as preconditioner */
#define FLGPU_OUTER_CG_PLAIN 2 /* <cg> without preconditioner */
#define FLGPU_OUTER_FCG 3      /* flexible CG: beta = -(z.Ap)/(p.Ap), which equals Polak-Ribiere's
                                  z.(r - r_old)/(z_old.r_old) */
#define FLGPU_STOP_REDUCTION 0 /* success-if (< :reduction tol) */
#define FLGPU_STOP_DEFNORM 1   /* success-if (< :defnorm tol) */

typedef struct flgpu_solve_params {
  int outer;           /* FLGPU_OUTER_* */
  int alias;           /* 1: precondition as femlisp's <cg> does, (iterate q w w) */
  int stop;            /* FLGPU_STOP_* */
  double tol;
  int maxsteps;        /* failure-if (> :step maxsteps) */
  int true_residual;   /* 1: recompute r = b - A x in FP64 when the recurrence
                          residual meets tol, continue if the true one does not */
  int learned;         /* 1: add the learned coarse correction (FCG only) */
  int stall_window;    /* learned: fall back to plain GMG-FCG when the true
                          residual has not dropped by 10x in this many steps, or
                          as soon as the network's step would add less than 1% of the
                          cycle step's energy (that step goes on without it; FCG restarts
                          from the residual only if the network was added before) */
  int restart_cycle;   /* <cg>'s restart-cycle (krylow.lisp:68-70,101-105,128-131): 0 never;
                          else every restart_cycle CG/FCG steps the step leaves no
                          :residual-after, the loop recomputes r = b - A x and the next step
                          restarts the recurrence from it */
  int fallback;        /* FLGPU_FALLBACK_NONE: <linear-solver>; FLGPU_FALLBACK_LU:
                          <safe-linear-solver> (linsolve.lisp:45-56,75-89): on the first
                          failure the iteration becomes <lu> on A (linit.lisp:132-162:
                          x <- x + A^-1 r with the residual at hand, residual-after nil;
                          dense FP64 inverse formed on the GPU, at most 8192 unknowns) and the
                          loop goes on with the same step count; a second failure returns
                          status 2 (femlisp: "The fallback iteration has failed.") */
} flgpu_solve_params;

#define FLGPU_FALLBACK_NONE 0
#define FLGPU_FALLBACK_LU 1

typedef struct flgpu_solve_report {
  int status;                 /* 0 success, 1 failure (step limit) */
  int steps;
  double initial_defnorm;
  double final_defnorm;       /* the norm femlisp observes (recurrence for CG) */
  double true_residual_norm;  /* ||b - A x||_2 in FP64 at exit */
  int fell_back;              /* learned: 1 if the plain fallback took over */
  double seconds;             /* wall time of the iteration loop */
  int fallback_used;          /* <safe-linear-solver>: 1 if the fallback iteration ran */
} flgpu_solve_report;

/* Solve A x = b from x = 0 with femlisp's <linear-solver> loop.  mg may be
 * NULL for FLGPU_OUTER_CG_PLAIN.  defnorms (may be NULL) receives
 * ||r_k||_2 for k = 0..steps, at most max_defnorms values.  Returns 0 on
 * success, 1 on failure, 2 when a <safe-linear-solver>'s fallback failed too.
 * mg, b and x must belong to A's context. */
int flgpu_solve(flgpu_mg *mg, const flgpu_mat *A, const flgpu_solve_params *p,
                const flgpu_vec *b, flgpu_vec *x, flgpu_solve_report *report,
                double *defnorms, int max_defnorms);

#ifdef __cplusplus
}
#endif
#endif /* FLGPU_H */
