/*
 * fl_oracle.c -- CPU restatement of femlisp's solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker + CPU baseline).  Not part of the
 * product library; see fl_oracle.h.  Each routine names the femlisp source it
 * restates (paths relative to the femlisp tree).
 */
#include <omp.h>
#include "fl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FLO_MAXDIM 3

/* ------------------------------------------------------------------------ */
/* lattice                                                                   */
/* ------------------------------------------------------------------------ */

static long ipow(long b, int e) {
  long r = 1;
  while (e-- > 0) r *= b;
  return r;
}

long flo_lattice_ndofs(int dim, int order, int n) {
  return ipow((long)order * n + 1, dim);
}

/* Range of lattice coordinates that share a cell with coordinate i along one
 * axis: cells c with order*c <= i <= order*c+order.  A DOF couples with every
 * DOF of every cell containing it (the pattern femlisp's assemble-interior
 * creates, src/discretization/fedisc.lisp:144-163). */
static void axis_range(int order, int n, int i, int *lo, int *hi) {
  int cmin, cmax;
  if (i % order == 0) {
    cmin = i / order - 1;
    cmax = i / order;
    if (cmin < 0) cmin = 0;
    if (cmax > n - 1) cmax = n - 1;
  } else {
    cmin = cmax = i / order;
  }
  *lo = order * cmin;
  *hi = order * cmax + order;
}

long flo_lattice_nnz(int dim, int order, int n) {
  long per_axis = 0;
  int M = order * n + 1;
  for (int i = 0; i < M; i++) {
    int lo, hi;
    axis_range(order, n, i, &lo, &hi);
    per_axis += hi - lo + 1;
  }
  return ipow(per_axis, dim);
}

static void unrank(long idx, int dim, int M, int *c) {
  for (int k = dim - 1; k >= 0; k--) {
    c[k] = (int)(idx % M);
    idx /= M;
  }
}

static long rank(const int *c, int dim, int M) {
  long idx = 0;
  for (int k = 0; k < dim; k++) idx = idx * M + c[k];
  return idx;
}

int flo_lattice_pattern(int dim, int order, int n, int *rowptr, int *colidx) {
  int M = order * n + 1;
  long N = ipow(M, dim);
  long pos = 0;
  for (long r = 0; r < N; r++) {
    int c[FLO_MAXDIM], lo[FLO_MAXDIM], hi[FLO_MAXDIM], q[FLO_MAXDIM];
    unrank(r, dim, M, c);
    for (int k = 0; k < dim; k++) axis_range(order, n, c[k], &lo[k], &hi[k]);
    rowptr[r] = (int)pos;
    for (int k = 0; k < dim; k++) q[k] = lo[k];
    for (;;) {
      if (colidx) colidx[pos] = (int)rank(q, dim, M);
      pos++;
      int k = dim - 1;
      while (k >= 0 && q[k] == hi[k]) {
        q[k] = lo[k];
        k--;
      }
      if (k < 0) break;
      q[k]++;
    }
  }
  rowptr[N] = (int)pos;
  return 0;
}

/* Every DOF on an :external-boundary patch is constrained, without a
 * dimension filter (src/discretization/system-fe.lisp:43-44). */
int flo_lattice_boundary(int dim, int order, int n, unsigned char *is_bdry) {
  int M = order * n + 1;
  long N = ipow(M, dim);
  for (long r = 0; r < N; r++) {
    int c[FLO_MAXDIM];
    unrank(r, dim, M, c);
    unsigned char b = 0;
    for (int k = 0; k < dim; k++)
      if (c[k] == 0 || c[k] == M - 1) b = 1;
    is_bdry[r] = b;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* quadrature and Lagrange basis                                             */
/* ------------------------------------------------------------------------ */

static void legendre(int s, double x, double *p, double *dp) {
  double p0 = 1.0, p1 = x;
  if (s == 0) {
    *p = 1.0;
    *dp = 0.0;
    return;
  }
  for (int k = 2; k <= s; k++) {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *p = p1;
  *dp = s * (x * p1 - p0) / (x * x - 1.0);
}

/* s-point Gauss rule on [0,1] as femlisp builds it: Legendre roots x_i in
 * ascending order, points (1 - x_i)/2, weights w_i/2
 * (src/discretization/quadrature.lisp:149-176, gauss-rule-for-weight with
 * n = 0).  Roots by Newton iteration instead of femlisp's interval method
 * (quadrature.lisp:122-146); both converge to the double nearest the root up
 * to the last ulp. */
int flo_gauss_rule(int s, double *pts, double *wts) {
  if (s < 1 || s > 16) return -1;
  for (int i = 0; i < s; i++) {
    /* i-th root in ascending order */
    double x = -cos(M_PI * (i + 0.75) / (s + 0.5));
    for (int it = 0; it < 100; it++) {
      double p, dp;
      legendre(s, x, &p, &dp);
      double dx = p / dp;
      x -= dx;
      if (fabs(dx) < 1e-17) break;
    }
    double p, dp;
    legendre(s, x, &p, &dp);
    double w = 2.0 / ((1.0 - x * x) * dp * dp);
    pts[i] = (1.0 - x) / 2.0;
    wts[i] = w / 2.0;
  }
  return 0;
}

/* Uniform Lagrange nodes a/order on [0,1] (lagrange-coords-1d :uniform,
 * src/discretization/lagrange.lisp:41-45) and the nodal basis with its
 * derivative in product form. */
void flo_lagrange_1d(int order, double t, double *phi, double *dphi) {
  for (int a = 0; a <= order; a++) {
    double ta = (double)a / order;
    double v = 1.0;
    for (int b = 0; b <= order; b++)
      if (b != a) v *= (t - (double)b / order) / (ta - (double)b / order);
    phi[a] = v;
    double d = 0.0;
    for (int m = 0; m <= order; m++) {
      if (m == a) continue;
      double term = 1.0 / (ta - (double)m / order);
      for (int b = 0; b <= order; b++)
        if (b != a && b != m) term *= (t - (double)b / order) / (ta - (double)b / order);
      d += term;
    }
    if (dphi) dphi[a] = d;
  }
}

/* Local element kernel for an axis-aligned cube cell of width H
 * (discretize-locally, src/discretization/ellsys-fe.lisp:68-127,183-190):
 * for every quadrature point q, weight = w_q * |det J| (fe-cell-geometry,
 * src/discretization/fe.lisp:502-535), gradients G = Ghat J^-1 = Ghat / H,
 *   K_ij += weight * (G_j . kappa) . G_i,   b_k += weight * phi_k * f.
 * Quadrature rule: order+1 Gauss points per axis (quadrature-rule,
 * src/discretization/fe.lisp:235-244). */
typedef struct {
  int dim, nloc, nq;
  double weight[27];
  double vals[27][27];
  double grads[27][27][FLO_MAXDIM];
} elem_tab;

/* per-level tables: quadrature weights times |det J|, basis values and
 * physical gradients at the quadrature points of a cell of width H */
static void elem_tab_init(elem_tab *t, int dim, int order, double H) {
  int s = order + 1;
  double pts[16], wts[16];
  flo_gauss_rule(s, pts, wts);
  int nl1 = order + 1;
  t->dim = dim;
  t->nloc = (int)ipow(nl1, dim);
  t->nq = (int)ipow(s, dim);
  double phi[16][8], dphi[16][8];
  for (int q = 0; q < s; q++) flo_lagrange_1d(order, pts[q], phi[q], dphi[q]);
  double detJ = 1.0;
  for (int k = 0; k < dim; k++) detJ *= H;
  for (int q = 0; q < t->nq; q++) {
    int qc[FLO_MAXDIM];
    unrank(q, dim, s, qc);
    /* femlisp multiplies volume-at-point by the rule weight
     * (fe.lisp:527); the product rule weight is the product of the factor
     * weights (quadrature.lisp:178-184). */
    double wq = 1.0;
    for (int k = 0; k < dim; k++) wq *= wts[qc[k]];
    t->weight[q] = detJ * wq;
    for (int a = 0; a < t->nloc; a++) {
      int ac[FLO_MAXDIM];
      unrank(a, dim, nl1, ac);
      double v = 1.0;
      for (int k = 0; k < dim; k++) v *= phi[qc[k]][ac[k]];
      t->vals[q][a] = v;
      for (int m = 0; m < dim; m++) {
        double g = 1.0;
        for (int k = 0; k < dim; k++)
          g *= (k == m) ? dphi[qc[k]][ac[k]] : phi[qc[k]][ac[k]];
        t->grads[q][a][m] = g / H;
      }
    }
  }
}

static void elem_apply(const elem_tab *t, const double *kappa_qp, double f, double *Ke,
                       double *be) {
  int nloc = t->nloc, dim = t->dim;
  if (Ke) memset(Ke, 0, sizeof(double) * nloc * nloc);
  if (be) memset(be, 0, sizeof(double) * nloc);
  for (int q = 0; q < t->nq; q++) {
    double weight = t->weight[q];
    double kap = kappa_qp ? kappa_qp[q] : 1.0;
    if (Ke) {
      for (int i = 0; i < nloc; i++)
        for (int j = 0; j < nloc; j++) {
          double gg = 0.0;
          for (int m = 0; m < dim; m++) gg += t->grads[q][j][m] * kap * t->grads[q][i][m];
          Ke[i * nloc + j] += weight * gg;
        }
    }
    if (be)
      for (int k = 0; k < nloc; k++) be[k] += weight * t->vals[q][k] * f;
  }
}

int flo_element_matrix(int dim, int order, double H, const double *kappa_qp,
                       double f, double *Ke, double *be) {
  if (dim < 1 || dim > FLO_MAXDIM || order < 1 || order > 2) return -1;
  elem_tab t;
  elem_tab_init(&t, dim, order, H);
  elem_apply(&t, kappa_qp, f, Ke, be);
  return 0;
}

/* coefficient cell sampled by quadrature point g (in (0,1)) of cell c on a
 * level with n cells per axis; nk = coefficient grid cells per axis */
static int kappa_axis_index(int c, double g, int n, int nk) {
  int r = nk / n;
  return c * r + (int)floor(g * r);
}

static long row_offset(int dim, int order, int n, const int *P, const int *Q) {
  long off = 0;
  for (int k = 0; k < dim; k++) {
    int lo, hi;
    axis_range(order, n, P[k], &lo, &hi);
    off = off * (hi - lo + 1) + (Q[k] - lo);
  }
  return off;
}

/* Global assembly (assemble-interior, src/discretization/fedisc.lisp:108-206;
 * scatter global-local-matrix-operation, sparseif.lisp:308-385) followed by
 * Dirichlet elimination A <- S A S + P, b <- S(b - A r) + r with r = 0
 * (constraints.lisp:323-411, fedisc.lisp:244-253, geomg.lisp:128-136).
 * Cells are visited in lattice order; femlisp's cell order is hash order, so
 * the summation order of contributions agrees only up to rounding. */
int flo_assemble(int dim, int order, int n, const double *kappa, int nk,
                 double f, double *vals, double *rhs) {
  if (dim < 1 || dim > FLO_MAXDIM || order < 1 || order > 2 || n < 1) return -1;
  if (kappa && (nk % n) != 0) return -2;
  int M = order * n + 1;
  long N = ipow(M, dim);
  int *rowptr = (int *)malloc(sizeof(int) * (N + 1));
  int *colidx = (int *)malloc(sizeof(int) * flo_lattice_nnz(dim, order, n));
  flo_lattice_pattern(dim, order, n, rowptr, colidx);
  long nnz = rowptr[N];
  memset(vals, 0, sizeof(double) * nnz);
  if (rhs) memset(rhs, 0, sizeof(double) * N);
  int s = order + 1, nl1 = order + 1;
  int nloc = (int)ipow(nl1, dim), nq = (int)ipow(s, dim);
  double pts[16], wts[16];
  flo_gauss_rule(s, pts, wts);
  double H = 1.0 / n;
  double Ke[729], be[27], kq[27];
  long ncells = ipow(n, dim);
  elem_tab *tab = (elem_tab *)malloc(sizeof(elem_tab));
  elem_tab_init(tab, dim, order, H);
  for (long cell = 0; cell < ncells; cell++) {
    int cc[FLO_MAXDIM];
    unrank(cell, dim, n, cc);
    for (int q = 0; q < nq; q++) {
      int qc[FLO_MAXDIM];
      unrank(q, dim, s, qc);
      if (kappa) {
        int kc[FLO_MAXDIM];
        for (int k = 0; k < dim; k++) kc[k] = kappa_axis_index(cc[k], pts[qc[k]], n, nk);
        kq[q] = kappa[rank(kc, dim, nk)];
      } else {
        kq[q] = 1.0;
      }
    }
    elem_apply(tab, kq, f, Ke, be);
    for (int a = 0; a < nloc; a++) {
      int ac[FLO_MAXDIM], P[FLO_MAXDIM];
      unrank(a, dim, nl1, ac);
      for (int k = 0; k < dim; k++) P[k] = order * cc[k] + ac[k];
      long row = rank(P, dim, M);
      if (rhs) rhs[row] += be[a];
      for (int b = 0; b < nloc; b++) {
        int bc[FLO_MAXDIM], Q[FLO_MAXDIM];
        unrank(b, dim, nl1, bc);
        for (int k = 0; k < dim; k++) Q[k] = order * cc[k] + bc[k];
        vals[rowptr[row] + row_offset(dim, order, n, P, Q)] += Ke[a * nloc + b];
      }
    }
  }
  free(tab);
  unsigned char *bd = (unsigned char *)malloc(N);
  flo_lattice_boundary(dim, order, n, bd);
  for (long r = 0; r < N; r++) {
    for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
      int c = colidx[e];
      if (bd[r] || bd[c]) vals[e] = (bd[r] && c == r) ? 1.0 : 0.0;
    }
    if (rhs && bd[r]) rhs[r] = 0.0;
  }
  free(bd);
  free(rowptr);
  free(colidx);
  return 0;
}

/* 1-D interpolation row of fine coordinate F (fine level 2*nc cells):
 * coarse columns and weights = coarse basis evaluated at the fine node
 * (compute-local-imatrix, src/discretization/fetransfer.lisp:73-138), entries
 * below *ipt-tolerance* = 1e-10 dropped (fetransfer.lisp:41-43,130). */
static int interp_1d(int order, int nc, int F, int *cols, double *w) {
  int cell = F / (2 * order);
  if (cell > nc - 1) cell = nc - 1;
  double tau = (double)(F - 2 * order * cell) / (2.0 * order);
  double phi[8];
  flo_lagrange_1d(order, tau, phi, NULL);
  int cnt = 0;
  for (int a = 0; a <= order; a++) {
    if (fabs(phi[a]) < 1e-10) continue;
    cols[cnt] = order * cell + a;
    w[cnt] = phi[a];
    cnt++;
  }
  return cnt;
}

long flo_projection(int dim, int order, int nc, int *rowptr, int *colidx, double *vals) {
  int Mc = order * nc + 1, Mf = order * 2 * nc + 1;
  long Nc = ipow(Mc, dim);
  for (long c = 0; c < Nc; c++) {
    long q = c, f = 0, stride = 1;
    for (int k = 0; k < dim; k++) { /* the coarse node's lattice coordinates, doubled */
      f += 2 * (q % Mc) * stride;
      q /= Mc;
      stride *= Mf;
    }
    if (rowptr) rowptr[c] = (int)c;
    if (colidx) colidx[c] = (int)f;
    if (vals) vals[c] = 1.0;
  }
  if (rowptr) rowptr[Nc] = (int)Nc;
  return Nc;
}

long flo_interpolation(int dim, int order, int nc, int *rowptr, int *colidx,
                       double *vals) {
  int Mf = 2 * order * nc + 1, Mc = order * nc + 1;
  long Nf = ipow(Mf, dim);
  long pos = 0;
  for (long r = 0; r < Nf; r++) {
    int F[FLO_MAXDIM], cnt[FLO_MAXDIM], cols[FLO_MAXDIM][8], q[FLO_MAXDIM];
    double w[FLO_MAXDIM][8];
    unrank(r, dim, Mf, F);
    int fb = 0;
    for (int k = 0; k < dim; k++) {
      cnt[k] = interp_1d(order, nc, F[k], cols[k], w[k]);
      if (F[k] == 0 || F[k] == Mf - 1) fb = 1;
    }
    rowptr[r] = (int)pos;
    for (int k = 0; k < dim; k++) q[k] = 0;
    for (;;) {
      if (colidx) {
        int C[FLO_MAXDIM];
        double v = 1.0;
        int cb = 0;
        for (int k = 0; k < dim; k++) {
          C[k] = cols[k][q[k]];
          v *= w[k][q[k]];
          if (C[k] == 0 || C[k] == Mc - 1) cb = 1;
        }
        colidx[pos] = (int)rank(C, dim, Mc);
        /* remove-projection-range :row-p for fine Dirichlet rows and
         * :column-p for coarse Dirichlet columns (constraints.lisp:58-72) */
        vals[pos] = (fb || cb) ? 0.0 : v;
      }
      pos++;
      int k = dim - 1;
      while (k >= 0 && q[k] == cnt[k] - 1) {
        q[k] = 0;
        k--;
      }
      if (k < 0) break;
      q[k]++;
    }
  }
  rowptr[Nf] = (int)pos;
  return pos;
}

/* Operator-dependent ("black-box") interpolation for 3D Q1 lattices: not in
 * femlisp.  It restates J. E. Dendy's black-box multigrid prolongation
 * (J. Comput. Phys. 48 (1982) 366-386) for 27-point stencils.  Fine node i
 * with s odd coordinates is interpolated from the 2^s coarse nodes of the
 * coarse cell around it.  Its row of A is collapsed onto the odd axes
 * (coll[key] = sum of a_ij over the even-axis offsets), and the equation
 * coll[0] e_i + sum_key coll[key] e_(i+key) = 0 is solved for e_i.  The
 * neighbours' e are taken from their own interpolation rows: nodes with fewer
 * odd coordinates are done first (1, then 2, then 3).  Boundary nodes carry
 * no weight, which zeroes fine Dirichlet rows and coarse Dirichlet columns as
 * femlisp's constrained interpolation does (constraints.lisp:58-72).  The
 * values go into the pattern of flo_interpolation(3, 1, nf/2, ...). */
int flo_blackbox_interpolation(int nf, const int *rowptr, const int *colidx,
                               const double *vals, const int *prowptr,
                               const int *pcolidx, double *pvals) {
  if (nf < 2 || nf % 2) return -1;
  int M = nf + 1, Mc = nf / 2 + 1;
  long N = (long)M * M * M;
  double *W = (double *)calloc((size_t)N * 8, sizeof(double));
  if (!W) return -1;
  for (int pass = 0; pass <= 3; pass++)
    for (long i = 0; i < N; i++) {
      int c[3];
      unrank(i, 3, M, c);
      int odd[3] = {c[0] & 1, c[1] & 1, c[2] & 1};
      if (odd[0] + odd[1] + odd[2] != pass) continue;
      int bd = 0;
      for (int k = 0; k < 3; k++)
        if (c[k] == 0 || c[k] == M - 1) bd = 1;
      if (bd) continue; /* boundary nodes keep W = 0 */
      if (pass == 0) {
        W[i * 8] = 1.0;
        continue;
      }
      double coll[27];
      memset(coll, 0, sizeof(coll));
      for (int e = rowptr[i]; e < rowptr[i + 1]; e++) {
        int d[3], key = 0;
        unrank(colidx[e], 3, M, d);
        int ok = 1;
        for (int k = 0; k < 3; k++) {
          int o = d[k] - c[k];
          if (o < -1 || o > 1) ok = 0;
          key = key * 3 + (odd[k] ? o + 1 : 1);
        }
        if (ok) coll[key] += vals[e];
      }
      double a0 = coll[13];
      if (a0 == 0.0) continue;
      double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int key = 0; key < 27; key++) {
        if (key == 13 || coll[key] == 0.0) continue;
        int o[3] = {key / 9 - 1, (key / 3) % 3 - 1, key % 3 - 1}, nb[3];
        for (int k = 0; k < 3; k++) nb[k] = c[k] + o[k];
        long j = rank(nb, 3, M);
        double s = -coll[key] / a0;
        for (int b = 0; b < 8; b++) {
          double w = W[j * 8 + b];
          if (w == 0.0) continue;
          /* corner b of j (bits z,y,x over j's own base corner) as a corner of i */
          int tb = 0;
          for (int k = 0; k < 3; k++) {
            int bit = (b >> (2 - k)) & 1;
            int cc = (nb[k] >> 1) + bit - (c[k] >> 1);
            tb = tb * 2 + cc;
          }
          acc[tb] += s * w;
        }
      }
      for (int b = 0; b < 8; b++) W[i * 8 + b] = acc[b];
    }
  long Nfr = N;
  for (long i = 0; i < Nfr; i++) {
    int c[3];
    unrank(i, 3, M, c);
    for (int e = prowptr[i]; e < prowptr[i + 1]; e++) {
      int cc[3], b = 0, in = 1;
      unrank(pcolidx[e], 3, Mc, cc);
      for (int k = 0; k < 3; k++) {
        int bit = cc[k] - (c[k] >> 1);
        if (bit < 0 || bit > 1) in = 0;
        b = b * 2 + bit;
      }
      pvals[e] = in ? W[i * 8 + b] : 0.0;
    }
  }
  free(W);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* algebra                                                                   */
/* ------------------------------------------------------------------------ */

/* gemm-nn! on <sparse-matrix> (src/matlisp/sparse-matrix.lisp:378-405):
 * per row x_i <- beta*x_i, then for every block x_i += (alpha*y_j)*A_ij
 * (a 1x1 dgemm, src/matlisp/standard-matrix-blas.lisp:209-266). */
void flo_spmv(int nrows, const int *rowptr, const int *colidx,
              const double *vals, double alpha, const double *y, double beta,
              double *x) {
  /* rows in parallel, as femlisp's :parallel t row loop (sparse-matrix.lisp:397-405); every
   * row's sum keeps its order */
#pragma omp parallel for schedule(static) if (nrows > 4096)
  for (int i = 0; i < nrows; i++) {
    double acc = (beta == 0.0) ? 0.0 : beta * x[i];
    for (int e = rowptr[i]; e < rowptr[i + 1]; e++)
      acc += (alpha * y[colidx[e]]) * vals[e];
    x[i] = acc;
  }
}

/* gemm-tn! (sparse-matrix.lisp:406-433, column access):
 * x_c <- beta*x_c + sum_r (alpha*y_r) * A_rc */
void flo_spmv_t(int nrows, int ncols, const int *rowptr, const int *colidx,
                const double *vals, double alpha, const double *y,
                double beta, double *x) {
  for (int c = 0; c < ncols; c++) x[c] = (beta == 0.0) ? 0.0 : beta * x[c];
  for (int r = 0; r < nrows; r++)
    for (int e = rowptr[r]; e < rowptr[r + 1]; e++)
      x[colidx[e]] += (alpha * y[r]) * vals[e];
}

void flo_csr_transpose(int nrows, int ncols, const int *rowptr,
                       const int *colidx, const double *vals, int *trowptr,
                       int *tcolidx, double *tvals) {
  int *cnt = (int *)calloc(ncols + 1, sizeof(int));
  for (int e = 0; e < rowptr[nrows]; e++) cnt[colidx[e] + 1]++;
  for (int c = 0; c < ncols; c++) cnt[c + 1] += cnt[c];
  memcpy(trowptr, cnt, sizeof(int) * (ncols + 1));
  for (int r = 0; r < nrows; r++)
    for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
      int p = cnt[colidx[e]]++;
      tcolidx[p] = r;
      if (tvals) tvals[p] = vals[e];
    }
  free(cnt);
}

/* Host threads for the parallel loops (the CPU baseline times 1 and all cores). */
void flo_set_threads(int n) { omp_set_num_threads(n > 0 ? n : omp_get_num_procs()); }
int flo_get_threads(void) { return omp_get_max_threads(); }

/* dot (sparse-vector.lisp:260-264 with with-accumulators, parallel-adaptions.lisp:30-42):
 * partial sums over DOT_CHUNKS fixed index ranges, each in index order, added in range order.
 * The ranges do not depend on the number of threads, so neither does the result. */
#define DOT_CHUNKS 64
static double dot(long n, const double *x, const double *y) {
  double part[DOT_CHUNKS];
  long len = (n + DOT_CHUNKS - 1) / DOT_CHUNKS;
#pragma omp parallel for schedule(static) if (n > 65536)
  for (int c = 0; c < DOT_CHUNKS; c++) {
    double s = 0.0;
    long lo = c * len, hi = lo + len < n ? lo + len : n;
    for (long i = lo; i < hi; i++) s += x[i] * y[i];
    part[c] = s;
  }
  double s = 0.0;
  for (int c = 0; c < DOT_CHUNKS; c++) s += part[c];
  return s;
}

/* ------------------------------------------------------------------------ */
/* multigrid hierarchy                                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  int nrows, ncols;
  int *rowptr, *colidx;
  double *vals;
} csr_t;

typedef struct {
  int n;
  long N;
  csr_t A;
  csr_t I;  /* interpolation level -> level+1 (rows: level+1 DOFs) */
  csr_t IT; /* its transpose */
  csr_t R;  /* FAS projection level+1 -> level (flo_projection) */
  double *diag;
  unsigned char *bdry;
  double *lu; /* dense LU of A (base level), row-major */
  int lu_n;
  double *sol, *rhs, *res;
  int residual_p;
} level_t;

struct flo_mg {
  int dim, order, n0, levels;
  level_t *lv;
  double *rhs_top;
};

static void csr_free(csr_t *m) {
  free(m->rowptr);
  free(m->colidx);
  free(m->vals);
  memset(m, 0, sizeof(*m));
}

/* galerkin-product R A P with R = P^T (src/iteration/amg.lisp:247-280):
 * loop over fine rows l, entries A_lm, coarse k with P_lk, and P's row m:
 * C_kn += (P_lk A_lm) P_mn, into the coarse lattice pattern (the structural
 * product of the constrained P, whose zeroed entries are stored, with A is
 * exactly that pattern).  Then coarse Dirichlet elimination + P: the zeroed
 * columns of P leave constrained rows and columns zero, and the identity is
 * added on their diagonal (geomg.lisp:128-136). */
void flo_mg_destroy(flo_mg *mg);

static int galerkin(level_t *C, const level_t *F) {
  const csr_t *A = &F->A, *P = &C->I;
  csr_t *Ac = &C->A;
  long nnz = Ac->rowptr[Ac->nrows];
  memset(Ac->vals, 0, sizeof(double) * nnz);
  for (int l = 0; l < A->nrows; l++)
    for (int e = A->rowptr[l]; e < A->rowptr[l + 1]; e++) {
      int m = A->colidx[e];
      double alm = A->vals[e];
      for (int pk = P->rowptr[l]; pk < P->rowptr[l + 1]; pk++) {
        int k = P->colidx[pk];
        double w = P->vals[pk] * alm;
        for (int pn = P->rowptr[m]; pn < P->rowptr[m + 1]; pn++) {
          int n = P->colidx[pn];
          int lo = Ac->rowptr[k], hi = Ac->rowptr[k + 1] - 1;
          while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (Ac->colidx[mid] < n) lo = mid + 1;
            else hi = mid;
          }
          if (Ac->colidx[lo] != n) return -1;
          Ac->vals[lo] += w * P->vals[pn];
        }
      }
    }
  for (long r = 0; r < C->N; r++)
    if (C->bdry[r])
      for (int e = Ac->rowptr[r]; e < Ac->rowptr[r + 1]; e++)
        if (Ac->colidx[e] == r) Ac->vals[e] += 1.0;
  for (long r = 0; r < C->N; r++)
    for (int e = Ac->rowptr[r]; e < Ac->rowptr[r + 1]; e++)
      if (Ac->colidx[e] == r) C->diag[r] = Ac->vals[e];
  return 0;
}

flo_mg *flo_mg_create(int dim, int order, int n0, int levels,
                      const double *kappa, int nk, double f) {
  return flo_mg_create_ex(dim, order, n0, levels, kappa, nk, f, 0);
}

flo_mg *flo_mg_create_ex(int dim, int order, int n0, int levels,
                         const double *kappa, int nk, double f, int galerkin_p) {
  if (dim < 1 || dim > 3 || order < 1 || order > 2 || n0 < 1 || levels < 1) return NULL;
  flo_mg *mg = (flo_mg *)calloc(1, sizeof(flo_mg));
  mg->dim = dim;
  mg->order = order;
  mg->n0 = n0;
  mg->levels = levels;
  mg->lv = (level_t *)calloc(levels, sizeof(level_t));
  int ntop = n0 << (levels - 1);
  if (kappa && nk % ntop != 0) {
    free(mg->lv);
    free(mg);
    return NULL;
  }
  for (int l = 0; l < levels; l++) {
    level_t *L = &mg->lv[l];
    L->n = n0 << l;
    L->N = flo_lattice_ndofs(dim, order, L->n);
    long nnz = flo_lattice_nnz(dim, order, L->n);
    L->A.nrows = L->A.ncols = (int)L->N;
    L->A.rowptr = (int *)malloc(sizeof(int) * (L->N + 1));
    L->A.colidx = (int *)malloc(sizeof(int) * nnz);
    L->A.vals = (double *)malloc(sizeof(double) * nnz);
    flo_lattice_pattern(dim, order, L->n, L->A.rowptr, L->A.colidx);
    double *rhs = (l == levels - 1) ? (double *)malloc(sizeof(double) * L->N) : NULL;
    /* with galerkin-p only the top level is discretised (geomg.lisp:119-126) */
    if (!galerkin_p || l == levels - 1) flo_assemble(dim, order, L->n, kappa, nk, f, L->A.vals, rhs);
    if (rhs) mg->rhs_top = rhs;
    L->diag = (double *)malloc(sizeof(double) * L->N);
    for (long r = 0; r < L->N; r++)
      for (int e = L->A.rowptr[r]; e < L->A.rowptr[r + 1]; e++)
        if (L->A.colidx[e] == r) L->diag[r] = L->A.vals[e];
    L->bdry = (unsigned char *)malloc(L->N);
    flo_lattice_boundary(dim, order, L->n, L->bdry);
    if (l < levels - 1) {
      long Nf = flo_lattice_ndofs(dim, order, 2 * L->n);
      L->I.nrows = (int)Nf;
      L->I.ncols = (int)L->N;
      L->I.rowptr = (int *)malloc(sizeof(int) * (Nf + 1));
      long inz = flo_interpolation(dim, order, L->n, L->I.rowptr, NULL, NULL);
      L->I.colidx = (int *)malloc(sizeof(int) * inz);
      L->I.vals = (double *)malloc(sizeof(double) * inz);
      flo_interpolation(dim, order, L->n, L->I.rowptr, L->I.colidx, L->I.vals);
      L->IT.nrows = (int)L->N;
      L->IT.ncols = (int)Nf;
      L->IT.rowptr = (int *)malloc(sizeof(int) * (L->N + 1));
      L->IT.colidx = (int *)malloc(sizeof(int) * inz);
      L->IT.vals = (double *)malloc(sizeof(double) * inz);
      flo_csr_transpose((int)Nf, (int)L->N, L->I.rowptr, L->I.colidx, L->I.vals,
                        L->IT.rowptr, L->IT.colidx, L->IT.vals);
      L->R.nrows = (int)L->N;
      L->R.ncols = (int)Nf;
      L->R.rowptr = (int *)malloc(sizeof(int) * (L->N + 1));
      L->R.colidx = (int *)malloc(sizeof(int) * L->N);
      L->R.vals = (double *)malloc(sizeof(double) * L->N);
      flo_projection(dim, order, L->n, L->R.rowptr, L->R.colidx, L->R.vals);
    }
    L->sol = (double *)calloc(L->N, sizeof(double));
    L->rhs = (double *)calloc(L->N, sizeof(double));
    L->res = (double *)calloc(L->N, sizeof(double));
  }
  if (galerkin_p == 2 && (dim != 3 || order != 1)) {
    flo_mg_destroy(mg);
    return NULL;
  }
  if (galerkin_p)
    for (int l = levels - 2; l >= 0; l--) {
      level_t *C = &mg->lv[l];
      const level_t *F = &mg->lv[l + 1];
      if (galerkin_p == 2) {
        /* operator-dependent transfer from the finer level's (Galerkin) matrix */
        flo_blackbox_interpolation(F->n, F->A.rowptr, F->A.colidx, F->A.vals, C->I.rowptr,
                                   C->I.colidx, C->I.vals);
        flo_csr_transpose(C->I.nrows, C->I.ncols, C->I.rowptr, C->I.colidx, C->I.vals,
                          C->IT.rowptr, C->IT.colidx, C->IT.vals);
      }
      if (galerkin(C, F) != 0) {
        flo_mg_destroy(mg);
        return NULL;
      }
    }
  return mg;
}

void flo_mg_destroy(flo_mg *mg) {
  if (!mg) return;
  for (int l = 0; l < mg->levels; l++) {
    level_t *L = &mg->lv[l];
    csr_free(&L->A);
    csr_free(&L->I);
    csr_free(&L->IT);
    csr_free(&L->R);
    free(L->diag);
    free(L->bdry);
    free(L->lu);
    free(L->sol);
    free(L->rhs);
    free(L->res);
  }
  free(mg->rhs_top);
  free(mg->lv);
  free(mg);
}

long flo_mg_ndofs(const flo_mg *mg, int level) {
  if (level < 0 || level >= mg->levels) return -1;
  return mg->lv[level].N;
}
long flo_mg_nnz(const flo_mg *mg, int level) {
  if (level < 0 || level >= mg->levels) return -1;
  return mg->lv[level].A.rowptr[mg->lv[level].N];
}
int flo_mg_get_matrix(const flo_mg *mg, int level, int *rowptr, int *colidx,
                      double *vals) {
  if (level < 0 || level >= mg->levels) return -1;
  const level_t *L = &mg->lv[level];
  long nnz = L->A.rowptr[L->N];
  if (rowptr) memcpy(rowptr, L->A.rowptr, sizeof(int) * (L->N + 1));
  if (colidx) memcpy(colidx, L->A.colidx, sizeof(int) * nnz);
  if (vals) memcpy(vals, L->A.vals, sizeof(double) * nnz);
  return 0;
}
long flo_mg_interp_nnz(const flo_mg *mg, int level) {
  if (level < 0 || level >= mg->levels - 1) return -1;
  return mg->lv[level].I.rowptr[mg->lv[level].I.nrows];
}
int flo_mg_get_interp(const flo_mg *mg, int level, int *rowptr, int *colidx,
                      double *vals) {
  if (level < 0 || level >= mg->levels - 1) return -1;
  const csr_t *I = &mg->lv[level].I;
  long nnz = I->rowptr[I->nrows];
  if (rowptr) memcpy(rowptr, I->rowptr, sizeof(int) * (I->nrows + 1));
  if (colidx) memcpy(colidx, I->colidx, sizeof(int) * nnz);
  if (vals) memcpy(vals, I->vals, sizeof(double) * nnz);
  return 0;
}
int flo_mg_get_projection(const flo_mg *mg, int level, int *rowptr, int *colidx,
                          double *vals) {
  if (level < 0 || level >= mg->levels - 1) return -1;
  const csr_t *R = &mg->lv[level].R;
  if (rowptr) memcpy(rowptr, R->rowptr, sizeof(int) * (R->nrows + 1));
  if (colidx) memcpy(colidx, R->colidx, sizeof(int) * R->nrows);
  if (vals) memcpy(vals, R->vals, sizeof(double) * R->nrows);
  return 0;
}
int flo_mg_get_rhs(const flo_mg *mg, double *rhs) {
  memcpy(rhs, mg->rhs_top, sizeof(double) * mg->lv[mg->levels - 1].N);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* iterators (src/iteration/linit.lisp)                                      */
/* ------------------------------------------------------------------------ */

/* compute-residual r = b - A x: copy! b r, then gemm! -1 A x 1 r
 * (linit.lisp:73-80).  When b and r are the same vector the copy is a no-op
 * and r <- r - A x. */
static void compute_residual(const csr_t *A, const double *x, const double *b,
                             double *r) {
  if (r != b) memcpy(r, b, sizeof(double) * A->nrows);
  flo_spmv(A->nrows, A->rowptr, A->colidx, A->vals, -1.0, x, 1.0, r);
}

/* <jacobi> iterate on <sparse-matrix> (linit.lisp:220-250): per row,
 * gesv! overwrites r_i with A_ii^-1 r_i, then x_i += damp * r_i. */
static void jacobi_iterate(const level_t *L, double omega, double *x, double *r) {
#pragma omp parallel for schedule(static) if (L->N > 4096)
  for (long i = 0; i < L->N; i++) {
    r[i] = r[i] / L->diag[i];
    x[i] += omega * r[i];
  }
}

/* <sor> iterate on <sparse-matrix> (linit.lisp:288-323), store-p nil: for
 * each row key in the ordering, corr := r_i := b_i, corr += (-x_j) A_ij over
 * the row (diagonal included), gesv! corr /= A_ii, x_i += omega corr. */
static void sor_row(const level_t *L, double omega, long i, double *x, const double *b,
                    double *r) {
  const csr_t *A = &L->A;
  double corr = b[i];
  for (int e = A->rowptr[i]; e < A->rowptr[i + 1]; e++)
    corr += (-x[A->colidx[e]]) * A->vals[e];
  corr = corr / L->diag[i];
  r[i] = corr;
  x[i] += omega * corr;
}

/* seg (NULL for the index order): nseg runs of the ordering, one per colour; rows of one
 * colour do not couple, so each run's rows are updated in parallel with the result of the
 * sequential sweep (femlisp's <parallel-sor> idea, linit.lisp:325-376, on a colouring) */
static void sor_iterate(const level_t *L, double omega, const long *ordering,
                        const long *seg, int nseg, double *x, const double *b, double *r) {
  if (!seg) {
    for (long k = 0; k < L->N; k++) sor_row(L, omega, ordering[k], x, b, r);
    return;
  }
  for (int c = 0; c < nseg; c++) {
    long lo = seg[c], hi = seg[c + 1];
#pragma omp parallel for schedule(static) if (hi - lo > 4096)
    for (long k = lo; k < hi; k++) sor_row(L, omega, ordering[k], x, b, r);
  }
}

/* ordering: FLO_ORDER_INDEX = canonical index order; FLO_ORDER_COLOUR =
 * <sor> :compare by lattice colour (coordinate parity for Q1, coordinate
 * mod 3 for Q2), then index (linit.lisp:258-263,293-295). */
static long *make_ordering(int dim, int order, const level_t *L, int mode,
                           int backwards, long **seg_out, int *nseg_out) {
  long *ord = (long *)malloc(sizeof(long) * L->N);
  int M = order * L->n + 1;
  int m = order + 1; /* colours per axis */
  long *seg = NULL;
  int ncol = 0;
  if (mode == FLO_ORDER_INDEX) {
    for (long i = 0; i < L->N; i++) ord[i] = i;
  } else {
    ncol = (int)ipow(m, dim);
    seg = (long *)malloc(sizeof(long) * (ncol + 1));
    long pos = 0;
    for (int col = 0; col < ncol; col++) {
      seg[col] = pos;
      for (long i = 0; i < L->N; i++) {
        int c[FLO_MAXDIM];
        unrank(i, dim, M, c);
        int cc = 0;
        for (int k = 0; k < dim; k++) cc = cc * m + c[k] % m;
        if (cc == col) ord[pos++] = i;
      }
    }
    seg[ncol] = pos;
  }
  if (backwards) {
    for (long i = 0; i < L->N / 2; i++) {
      long t = ord[i];
      ord[i] = ord[L->N - 1 - i];
      ord[L->N - 1 - i] = t;
    }
    if (seg) { /* the runs in reverse order: run c starts where run ncol-1-c ended */
      long *rs = (long *)malloc(sizeof(long) * (ncol + 1));
      for (int c = 0; c <= ncol; c++) rs[c] = L->N - seg[ncol - c];
      free(seg);
      seg = rs;
    }
  }
  if (seg_out) *seg_out = seg;
  else free(seg);
  if (nseg_out) *nseg_out = ncol;
  return ord;
}

/* setup-blocks of <geometric-blocking-mixin> with :vertex-centered blocks (geoblock.lisp:
 * 113-157) and find-vertex-centered-block (:65-91).  One block per mesh vertex v (a lattice
 * point whose coordinates are multiples of order): v itself and every DOF of a cell containing
 * v that is not Dirichlet -- the lattice points v + d with |d_k| <= order - 1 (a point with
 * |d_k| = order lies on a cell without v).  Uniform meshes have no hanging nodes, so the
 * constraints-Q loop adds nothing.  keys[b * W + s], W = (2 order - 1)^dim, slot s = d in
 * lexicographic order (axis 0 most significant), -1 where v + d lies outside the lattice or is
 * Dirichlet (v itself is kept either way, as the reference keeps the vertex key).  Blocks come
 * sorted by colour -- the parity of v / order per axis, axis 0 most significant -- then by
 * index: blocks of one colour neither share a key nor read one another's keys through A, so any
 * order within a colour gives the same iterate.  Returns the number of blocks (n + 1)^dim;
 * keys may be NULL. */
long flo_vertex_blocks(int dim, int order, int n, int *keys) {
  if (dim < 1 || dim > FLO_MAXDIM || order < 1 || n < 1) return -1;
  int M = order * n + 1, h = order - 1, w1 = 2 * order - 1;
  int W = (int)ipow(w1, dim);
  long nv = ipow(n + 1, dim);
  if (!keys) return nv;
  unsigned char *bd = (unsigned char *)malloc((size_t)ipow(M, dim));
  flo_lattice_boundary(dim, order, n, bd);
  long pos = 0;
  for (int col = 0; col < (1 << dim); col++)
    for (long iv = 0; iv < nv; iv++) {
      int a[FLO_MAXDIM];
      unrank(iv, dim, n + 1, a);
      int cc = 0;
      for (int k = 0; k < dim; k++) cc = 2 * cc + (a[k] & 1);
      if (cc != col) continue;
      int *slot = keys + pos * W;
      for (int s = 0; s < W; s++) {
        int d[FLO_MAXDIM], c[FLO_MAXDIM], inside = 1, centre = 1;
        unrank(s, dim, w1, d);
        for (int k = 0; k < dim; k++) {
          c[k] = a[k] * order + d[k] - h;
          if (c[k] < 0 || c[k] >= M) inside = 0;
          if (d[k] != h) centre = 0;
        }
        long key = inside ? rank(c, dim, M) : -1;
        if (key >= 0 && !centre && bd[key]) key = -1;
        slot[s] = (int)key;
      }
      pos++;
    }
  free(bd);
  return nv;
}

/* A_ij of a CSR matrix with ascending columns (0 outside the pattern) */
static double csr_entry(const csr_t *A, int i, int j) {
  int lo = A->rowptr[i], hi = A->rowptr[i + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (A->colidx[mid] == j) return A->vals[mid];
    if (A->colidx[mid] < j) lo = mid + 1;
    else hi = mid - 1;
  }
  return 0.0;
}

/* gesv! (matrix.lisp:428-436, LAPACK dgesv): LU with partial pivoting of the row-major m x m
 * matrix a (destroyed), b overwritten by the solution */
static void dense_gesv(int m, double *a, double *b) {
  for (int k = 0; k < m; k++) {
    int p = k;
    for (int i = k + 1; i < m; i++)
      if (fabs(a[(size_t)i * m + k]) > fabs(a[(size_t)p * m + k])) p = i;
    if (p != k) {
      for (int j = 0; j < m; j++) {
        double t = a[(size_t)k * m + j];
        a[(size_t)k * m + j] = a[(size_t)p * m + j];
        a[(size_t)p * m + j] = t;
      }
      double t = b[k];
      b[k] = b[p];
      b[p] = t;
    }
    for (int i = k + 1; i < m; i++) {
      double l = a[(size_t)i * m + k] / a[(size_t)k * m + k];
      for (int j = k + 1; j < m; j++) a[(size_t)i * m + j] -= l * a[(size_t)k * m + j];
      b[i] -= l * b[k];
    }
  }
  for (int i = m - 1; i >= 0; i--) {
    double s = b[i];
    for (int j = i + 1; j < m; j++) s -= a[(size_t)i * m + j] * b[j];
    b[i] = s / a[(size_t)i * m + i];
  }
}

/* One <ssc> or <psc> iterate on a <sparse-matrix> (make-iterator of <block-iteration>,
 * blockit.lisp:113-203; store-p nil, no inner iteration), blocks as flo_vertex_blocks gives
 * them, first to last (backwards: last to first).  Per block B:
 *  <ssc> first updates the residual on B, r_k := b_k, r_k -= sum_j A_kj x_j (:156-161);
 *  both then solve A_BB y = r_B by gesv! on a copy (:179-185) and set x_B += omega y, where
 *  omega is <ssc>'s omega and <psc>'s damp (:139-142).
 * <psc> reads r as the caller left it (residual-before t) and never writes it. */
static void one_block(const csr_t *A, const int *blk, int W, int ssc, double omega,
                      double *x, const double *b, double *r, double *loc, double *y, int *ks);

/* the colour of block q: the parity of its vertex coordinates divided by the order (the vertex
 * is the centre slot of the key list) */
static int block_colour(const int *keys, long q, int W, int dim, int order, int n) {
  int M = order * n + 1, c[FLO_MAXDIM], cc = 0;
  unrank(keys[q * W + (W - 1) / 2], dim, M, c);
  for (int k = 0; k < dim; k++) cc = cc * 2 + (c[k] / order) % 2;
  return cc;
}

static void block_iterate(const csr_t *A, const int *keys, long nb, int W, int ssc,
                          double omega, int backwards, double *x, const double *b, double *r,
                          int dim, int order, int n) {
  /* runs of blocks of one colour (flo_vertex_blocks sorts by colour): their blocks neither
   * share a key nor read one another's keys, so each run is swept in parallel with the
   * result of the sequential sweep */
  long lo = 0;
  while (lo < nb) {
    long qlo = backwards ? nb - 1 - lo : lo;
    int col = block_colour(keys, qlo, W, dim, order, n);
    long hi = lo + 1;
    while (hi < nb && block_colour(keys, backwards ? nb - 1 - hi : hi, W, dim, order, n) == col)
      hi++;
#pragma omp parallel if (hi - lo > 256)
    {
      double *loc = (double *)malloc(sizeof(double) * W * W);
      double *y = (double *)malloc(sizeof(double) * W);
      int *ks = (int *)malloc(sizeof(int) * W);
#pragma omp for schedule(static)
      for (long q = lo; q < hi; q++)
        one_block(A, keys + (backwards ? nb - 1 - q : q) * W, W, ssc, omega, x, b, r, loc, y,
                  ks);
      free(loc);
      free(y);
      free(ks);
    }
    lo = hi;
  }
}

static void one_block(const csr_t *A, const int *blk, int W, int ssc, double omega,
                      double *x, const double *b, double *r, double *loc, double *y, int *ks) {
  {
    int m = 0;
    for (int s = 0; s < W; s++)
      if (blk[s] >= 0) ks[m++] = blk[s];
    if (ssc)
      for (int i = 0; i < m; i++) {
        int k = ks[i];
        double rk = b[k];
        for (int e = A->rowptr[k]; e < A->rowptr[k + 1]; e++)
          rk += -1.0 * (A->vals[e] * x[A->colidx[e]]);
        r[k] = rk;
      }
    for (int i = 0; i < m; i++) {
      y[i] = r[ks[i]];
      for (int j = 0; j < m; j++) loc[i * m + j] = csr_entry(A, ks[i], ks[j]);
    }
    dense_gesv(m, loc, y);
    for (int i = 0; i < m; i++) x[ks[i]] += omega * y[i];
  }
}

int flo_block_iterate(int nrows, const int *rowptr, const int *colidx, const double *vals,
                      int dim, int order, int n, int ssc, double omega, int backwards,
                      double *x, const double *b, double *r) {
  long nb = flo_vertex_blocks(dim, order, n, NULL);
  if (nb < 0 || flo_lattice_ndofs(dim, order, n) != nrows) return -1;
  int W = (int)ipow(2 * order - 1, dim);
  int *keys = (int *)malloc(sizeof(int) * nb * W);
  flo_vertex_blocks(dim, order, n, keys);
  csr_t A = {nrows, nrows, (int *)rowptr, (int *)colidx, (double *)vals};
  block_iterate(&A, keys, nb, W, ssc, omega, backwards, x, b, r, dim, order, n);
  free(keys);
  return 0;
}

/* dense LU without pivoting of the base-level matrix: femlisp factors the
 * coarse matrix once (<lu> store-p, linit.lisp:132-162; sparse-ldu,
 * src/matlisp/sparselu.lisp:61-126) and applies getrs!. */
static void coarse_factor(level_t *L) {
  int n = (int)L->N;
  L->lu_n = n;
  L->lu = (double *)calloc((size_t)n * n, sizeof(double));
  for (int r = 0; r < n; r++)
    for (int e = L->A.rowptr[r]; e < L->A.rowptr[r + 1]; e++)
      L->lu[(size_t)r * n + L->A.colidx[e]] = L->A.vals[e];
  double *a = L->lu;
  for (int k = 0; k < n; k++)
    for (int i = k + 1; i < n; i++) {
      double lik = a[(size_t)i * n + k] / a[(size_t)k * n + k];
      a[(size_t)i * n + k] = lik;
      if (lik != 0.0)
        for (int j = k + 1; j < n; j++) a[(size_t)i * n + j] -= lik * a[(size_t)k * n + j];
    }
}

static void coarse_solve(const level_t *L, double *r) {
  int n = L->lu_n;
  const double *a = L->lu;
  for (int i = 0; i < n; i++) {
    double s = r[i];
    for (int j = 0; j < i; j++) s -= a[(size_t)i * n + j] * r[j];
    r[i] = s;
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = r[i];
    for (int j = i + 1; j < n; j++) s -= a[(size_t)i * n + j] * r[j];
    r[i] = s / a[(size_t)i * n + i];
  }
}

typedef struct {
  flo_mg *mg;
  const flo_params *p;
  long **pre_ord, **post_ord;
  long **pre_seg, **post_seg; /* colour runs of the orderings (NULL: index order) */
  int nseg;
  int **blk;   /* vertex-centred blocks per level (block smoothers) */
  long *nblk;
  int blk_w;
  int top;
} mgrun_t;

/* product-iterator (linit.lisp:94-120): nr-steps sweeps, recomputing the
 * residual after each sweep when residual-before and not residual-after. */
static void smoother_run(mgrun_t *R, int l, int post) {
  level_t *L = &R->mg->lv[l];
  const flo_params *p = R->p;
  int steps = post ? p->post_steps : p->pre_steps;
  if (steps <= 0) return;
  if (p->smoother == FLO_SMOOTH_JACOBI) {
    /* residual-before t: ensure-mg-residual (multigrid.lisp:173-189) */
    if (!L->residual_p) {
      compute_residual(&L->A, L->sol, L->rhs, L->res);
      L->residual_p = 1;
    }
    for (int k = 0; k < steps; k++) {
      jacobi_iterate(L, p->omega, L->sol, L->res);
      if (steps > 1) compute_residual(&L->A, L->sol, L->rhs, L->res);
    }
  } else if (p->smoother == FLO_SMOOTH_SSC || p->smoother == FLO_SMOOTH_PSC) {
    /* <block-iteration> iterators have residual-before t (blockit.lisp:125) */
    if (!L->residual_p) {
      compute_residual(&L->A, L->sol, L->rhs, L->res);
      L->residual_p = 1;
    }
    for (int k = 0; k < steps; k++) {
      block_iterate(&L->A, R->blk[l], R->nblk[l], R->blk_w, p->smoother == FLO_SMOOTH_SSC,
                    p->omega, post && p->symmetric, L->sol, L->rhs, L->res, R->mg->dim,
                    R->mg->order, L->n);
      if (steps > 1) compute_residual(&L->A, L->sol, L->rhs, L->res);
    }
  } else {
    long *ord = post ? R->post_ord[l] : R->pre_ord[l];
    long *seg = post ? R->post_seg[l] : R->pre_seg[l];
    for (int k = 0; k < steps; k++)
      sor_iterate(L, p->omega, ord, seg, R->nseg, L->sol, L->rhs, L->res);
  }
  /* residual-after nil for every smoother */
  L->residual_p = 0;
}

/* FAS-R_{l-1} applied to level l's vector v (gemm! with the projection matrix, multigrid.lisp:
 * 230, 246): sign = 0: out = R v; sign = -1: out = out - R v */
static void fas_project(const flo_mg *mg, int l, const double *v, double *out, int sign) {
  const csr_t *R = &mg->lv[l - 1].R;
  flo_spmv(R->nrows, R->rowptr, R->colidx, R->vals, sign ? -1.0 : 1.0, v, sign ? 1.0 : 0.0, out);
}

/* lmgc with the correction scheme or FAS, :multiplicative or :additive, gamma coarse cycles
 * (multigrid.lisp:202-252,254-288). */
static void lmgc(mgrun_t *R, int l) {
  flo_mg *mg = R->mg;
  level_t *L = &mg->lv[l];
  if (l == R->p->base_level) {
    /* coarse-grid iterator <lu>: r <- A^-1 r, x += damp r */
    coarse_solve(L, L->res);
    for (long i = 0; i < L->N; i++) L->sol[i] += 1.0 * L->res[i];
    L->residual_p = 0;
    return;
  }
  if (!R->p->additive) smoother_run(R, l, 0);
  /* restrict: ensure-mg-residual, res_{l-1} = I^T res_l */
  if (!L->residual_p) {
    compute_residual(&L->A, L->sol, L->rhs, L->res);
    L->residual_p = 1;
  }
  level_t *C = &mg->lv[l - 1];
  flo_spmv(C->IT.nrows, C->IT.rowptr, C->IT.colidx, C->IT.vals, 1.0, L->res, 0.0, C->res);
  if (R->p->fas) {
    /* <fas> :after (multigrid.lisp:226-232): rhs_c = res_c, sol_c = FAS-R sol_l,
     * rhs_c += A_c sol_c, residual-p t */
    memcpy(C->rhs, C->res, sizeof(double) * C->N);
    fas_project(mg, l, L->sol, C->sol, 0);
    flo_spmv(C->A.nrows, C->A.rowptr, C->A.colidx, C->A.vals, 1.0, C->sol, 1.0, C->rhs);
  } else {
    /* <correction-scheme> :after: x_c = 0, rhs_c = res_c, residual-p t */
    memset(C->sol, 0, sizeof(double) * C->N);
    memcpy(C->rhs, C->res, sizeof(double) * C->N);
  }
  C->residual_p = 1;
  /* :additive smooths the fine level after the restriction, from the same residual */
  if (R->p->additive) smoother_run(R, l, 0);
  /* gamma coarse cycles, one when the coarse level is the base level, none for gamma 0; the
   * coarse residual is brought up to date between them (multigrid.lisp:268-273) */
  int reps = R->p->gamma == 0 ? 0 : (l - 1 == R->p->base_level ? 1 : R->p->gamma);
  for (int i = reps; i >= 1; i--) {
    lmgc(R, l - 1);
    if (i != 1 && !C->residual_p) {
      compute_residual(&C->A, C->sol, C->rhs, C->res);
      C->residual_p = 1;
    }
  }
  /* :additive post-smooths before the prolongation */
  if (R->p->additive) smoother_run(R, l, 1);
  if (R->p->fas) {
    /* <fas> prolongate (multigrid.lisp:241-247): res_c = sol_c - FAS-R sol_l,
     * sol_l += I res_c */
    memcpy(C->res, C->sol, sizeof(double) * C->N);
    fas_project(mg, l, L->sol, C->res, -1);
    flo_spmv(C->I.nrows, C->I.rowptr, C->I.colidx, C->I.vals, 1.0, C->res, 1.0, L->sol);
  } else {
    /* prolongate: sol_l += I sol_{l-1} */
    flo_spmv(C->I.nrows, C->I.rowptr, C->I.colidx, C->I.vals, 1.0, C->sol, 1.0, L->sol);
  }
  L->residual_p = 0; /* prolongate :after, residual-p nil */
  if (!R->p->additive) smoother_run(R, l, 1);
}

static void mgrun_init(mgrun_t *R, flo_mg *mg, const flo_params *p) {
  R->mg = mg;
  R->p = p;
  R->top = mg->levels - 1;
  R->pre_ord = (long **)calloc(mg->levels, sizeof(long *));
  R->post_ord = (long **)calloc(mg->levels, sizeof(long *));
  R->pre_seg = (long **)calloc(mg->levels, sizeof(long *));
  R->post_seg = (long **)calloc(mg->levels, sizeof(long *));
  R->nseg = 0;
  if (p->smoother == FLO_SMOOTH_SOR)
    for (int l = p->base_level + 1; l < mg->levels; l++) {
      R->pre_ord[l] = make_ordering(mg->dim, mg->order, &mg->lv[l], p->ordering, 0,
                                    &R->pre_seg[l], &R->nseg);
      if (p->symmetric) {
        R->post_ord[l] = make_ordering(mg->dim, mg->order, &mg->lv[l], p->ordering, 1,
                                       &R->post_seg[l], &R->nseg);
      } else {
        R->post_ord[l] = R->pre_ord[l];
        R->post_seg[l] = R->pre_seg[l];
      }
    }
  R->blk = (int **)calloc(mg->levels, sizeof(int *));
  R->nblk = (long *)calloc(mg->levels, sizeof(long));
  R->blk_w = (int)ipow(2 * mg->order - 1, mg->dim);
  if (p->smoother == FLO_SMOOTH_SSC || p->smoother == FLO_SMOOTH_PSC)
    for (int l = p->base_level + 1; l < mg->levels; l++) {
      int n = mg->lv[l].n;
      R->nblk[l] = flo_vertex_blocks(mg->dim, mg->order, n, NULL);
      R->blk[l] = (int *)malloc(sizeof(int) * R->nblk[l] * R->blk_w);
      flo_vertex_blocks(mg->dim, mg->order, n, R->blk[l]);
    }
  level_t *B = &mg->lv[p->base_level];
  if (!B->lu) coarse_factor(B);
}

static void mgrun_free(mgrun_t *R) {
  for (int l = 0; l < R->mg->levels; l++) {
    if (R->post_ord[l] != R->pre_ord[l]) free(R->post_ord[l]);
    free(R->pre_ord[l]);
    if (R->post_seg[l] != R->pre_seg[l]) free(R->post_seg[l]);
    free(R->pre_seg[l]);
    free(R->blk[l]);
  }
  free(R->pre_ord);
  free(R->post_ord);
  free(R->pre_seg);
  free(R->post_seg);
  free(R->blk);
  free(R->nblk);
}

/* the <mg-iteration> iterator (multigrid.lisp:368-391): top-level sol/rhs/res
 * are the caller's vectors, residual-p t. */
static void mg_iterate(mgrun_t *R, double *x, double *b, double *r) {
  flo_mg *mg = R->mg;
  level_t *T = &mg->lv[R->top];
  double *s0 = T->sol, *b0 = T->rhs, *r0 = T->res;
  T->sol = x;
  T->rhs = b;
  T->res = r;
  T->residual_p = 1;
  if (R->top == R->p->base_level) {
    coarse_solve(T, T->res);
    for (long i = 0; i < T->N; i++) T->sol[i] += T->res[i];
  } else {
    lmgc(R, R->top);
  }
  T->sol = s0;
  T->rhs = b0;
  T->res = r0;
}

/* preconditioner call as <cg> performs it (krylow.lisp:118-121):
 * (copy! r w) (x<-0 q) (funcall iterate q w w).  alias = 0 passes a second
 * copy of r as the residual vector instead. */
static void precond(mgrun_t *R, long N, const double *r, double *q, double *w,
                    double *w2) {
  memcpy(w, r, sizeof(double) * N);
  memset(q, 0, sizeof(double) * N);
  if (R->p->alias) {
    mg_iterate(R, q, w, w);
  } else {
    memcpy(w2, r, sizeof(double) * N);
    mg_iterate(R, q, w, w2);
  }
}

int flo_mg_apply(flo_mg *mg, const flo_params *p, const double *b, double *x) {
  if (p->base_level < 0 || p->base_level >= mg->levels) return -1;
  mgrun_t R;
  mgrun_init(&R, mg, p);
  long N = mg->lv[R.top].N;
  double *w = (double *)malloc(sizeof(double) * N);
  double *w2 = (double *)malloc(sizeof(double) * N);
  precond(&R, N, b, x, w, w2);
  free(w);
  free(w2);
  mgrun_free(&R);
  return 0;
}

/* <linear-solver> loop (iterate.lisp:239-250; solve.lisp:126-155;
 * linsolve.lisp:58-73; problem.lisp:109-117) around the iteration chosen by
 * p->outer: the V-cycle itself, <cg> with the V-cycle as preconditioner
 * (krylow.lisp:72-132), plain <cg>, or flexible CG (Polak-Ribiere beta).
 *
 * restart_cycle restates <cg>'s restart-cycle (krylow.lisp:101-105,128-131): the
 * iterator counts its steps, and when the count reaches a multiple of the cycle the
 * step returns no :residual-after flag, so the loop recomputes r = b - A x
 * (solve.lisp:139-145), and the next step restarts the recurrence from it.
 *
 * fallback restates <safe-linear-solver> (linsolve.lisp:45-56,75-89): when the
 * failure test holds the first time, the iterator is replaced by the fallback's,
 * <lu> (linit.lisp:132-162: r <- A^-1 r, x <- x + r, residual-after nil), and the
 * loop goes on with the same blackboard (step count, defnorms, the residual the
 * failed iteration left); a second failure is femlisp's "The fallback iteration has
 * failed." error, status 2 here. */
int flo_mg_solve(flo_mg *mg, const flo_params *p, const double *b, double *x,
                 int *steps_out, double *defnorms) {
  if (p->base_level < 0 || p->base_level >= mg->levels) return -1;
  mgrun_t R;
  mgrun_init(&R, mg, p);
  level_t *T = &mg->lv[R.top];
  long N = T->N;
  double *bb = (double *)malloc(sizeof(double) * N);
  memcpy(bb, b, sizeof(double) * N);
  double *r = (double *)calloc(N, sizeof(double));
  double *pv = (double *)calloc(N, sizeof(double));
  double *a = (double *)calloc(N, sizeof(double));
  double *w = (double *)calloc(N, sizeof(double));
  double *w2 = (double *)calloc(N, sizeof(double));
  double *q = (double *)calloc(N, sizeof(double));
  double *rold = (double *)calloc(N, sizeof(double));
  memset(x, 0, sizeof(double) * N);
  int residual_p = 0, step = 0, initialized = 0, status = 1;
  int count = 0;      /* <cg>'s step count for the restart cycle */
  int fallback_p = 0; /* the <lu> fallback iterator is in use */
  level_t F;          /* the top-level matrix's LU for the fallback */
  memset(&F, 0, sizeof F);
  double initial = -1.0, alpha = 0.0;
  for (;;) {
    /* intermediate :before */
    if (!residual_p) {
      compute_residual(&T->A, x, bb, r);
      residual_p = 1;
    }
    double defnorm = sqrt(dot(N, r, r));
    if (initial < 0) initial = defnorm;
    if (defnorms) defnorms[step] = defnorm;
    double reduction = (defnorm == 0.0) ? 0.0 : defnorm / initial;
    if ((p->stop == FLO_STOP_REDUCTION && reduction < p->tol) ||
        (p->stop == FLO_STOP_DEFNORM && defnorm < p->tol)) {
      status = 0;
      break;
    }
    if (step > p->maxsteps) {
      if (p->fallback == FLO_FALLBACK_NONE) {
        status = 1;
        break;
      }
      if (fallback_p) { /* "The fallback iteration has failed." */
        status = 2;
        break;
      }
      /* change to the fallback iteration; nothing else on the blackboard changes */
      fallback_p = 1;
      F = *T;
      F.lu = NULL;
      coarse_factor(&F);
    }
    if (fallback_p) { /* the <lu> iterate on the residual at hand */
      coarse_solve(&F, r);
      #pragma omp parallel for simd if (N > 65536)
      for (long i = 0; i < N; i++) x[i] += r[i];
      residual_p = 0;
      step++;
      continue;
    }
    /* <cg>'s restart at the start of a step after a completed cycle */
    if (p->restart_cycle > 0 && count > 0 && count % p->restart_cycle == 0) initialized = 0;
    /* next-step */
    switch (p->outer) {
      case FLO_OUTER_MG:
        mg_iterate(&R, x, bb, r);
        residual_p = 0;
        break;
      case FLO_OUTER_CG:
      case FLO_OUTER_CG_PLAIN: {
        int pc = (p->outer == FLO_OUTER_CG);
        if (!initialized) { /* restart */
          if (pc) precond(&R, N, r, pv, w, w2);
          else memcpy(pv, r, sizeof(double) * N);
          alpha = dot(N, pv, r);
          initialized = 1;
        }
        if (alpha != 0.0) {
          flo_spmv((int)N, T->A.rowptr, T->A.colidx, T->A.vals, 1.0, pv, 0.0, a);
          double beta = dot(N, a, pv);
          double lam = alpha / beta;
          #pragma omp parallel for simd if (N > 65536)
          for (long i = 0; i < N; i++) x[i] += lam * pv[i];
          #pragma omp parallel for simd if (N > 65536)
          for (long i = 0; i < N; i++) r[i] += (-lam) * a[i];
          if (pc) precond(&R, N, r, q, w, w2);
          else memcpy(q, r, sizeof(double) * N);
          double new_alpha = dot(N, q, r);
          double sc = new_alpha / alpha;
          #pragma omp parallel for simd if (N > 65536)
          for (long i = 0; i < N; i++) pv[i] = sc * pv[i] + q[i];
          alpha = new_alpha;
        }
        residual_p = !(p->restart_cycle > 0 && ++count % p->restart_cycle == 0);
        break;
      }
      case FLO_OUTER_FCG: {
        if (!initialized) {
          precond(&R, N, r, q, w, w2);
          memcpy(pv, q, sizeof(double) * N);
          alpha = dot(N, q, r);
          initialized = 1;
        }
        if (alpha != 0.0) {
          flo_spmv((int)N, T->A.rowptr, T->A.colidx, T->A.vals, 1.0, pv, 0.0, a);
          double lam = alpha / dot(N, pv, a);
          #pragma omp parallel for simd if (N > 65536)
          for (long i = 0; i < N; i++) x[i] += lam * pv[i];
          memcpy(rold, r, sizeof(double) * N);
          #pragma omp parallel for simd if (N > 65536)
          for (long i = 0; i < N; i++) r[i] += (-lam) * a[i];
          precond(&R, N, r, q, w, w2);
          double zr = dot(N, q, r);
          double beta = (zr - dot(N, q, rold)) / alpha;
          #pragma omp parallel for simd if (N > 65536)
          for (long i = 0; i < N; i++) pv[i] = q[i] + beta * pv[i];
          alpha = zr;
        }
        residual_p = !(p->restart_cycle > 0 && ++count % p->restart_cycle == 0);
        break;
      }
      default:
        status = -2;
        goto done;
    }
    step++;
  }
done:
  if (steps_out) *steps_out = step;
  free(F.lu);
  free(bb);
  free(r);
  free(pv);
  free(a);
  free(w);
  free(w2);
  free(q);
  free(rold);
  mgrun_free(&R);
  return status;
}
