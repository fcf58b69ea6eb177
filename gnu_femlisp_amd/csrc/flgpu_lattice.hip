// Structured lattices: femlisp's sparsity pattern, cell assembly with
// Dirichlet elimination, and the constrained :local interpolation, all built
// on the GPU for uniformly refined box meshes of the unit cube.
//
// Restated femlisp sources:
//   DOF keys / canonical order  src/discretization/sparseas.lisp:46-59,
//                               src/mesh/meshgen.lisp:200-217 (midpoint order)
//   pattern                     src/discretization/fedisc.lisp:144-163
//   element kernel              src/discretization/ellsys-fe.lisp:68-127,183-190,
//                               src/discretization/fe.lisp:235-244,502-535
//   quadrature                  src/discretization/quadrature.lisp:149-215
//   Lagrange basis              src/discretization/lagrange.lisp:41-202
//   Dirichlet elimination       src/discretization/constraints.lisp:323-411,
//                               src/discretization/system-fe.lisp:37-129
//   interpolation               src/discretization/sparseas.lisp:232-272,
//                               fetransfer.lisp:73-138, constraints.lisp:58-72
//
// Layout: a level with n cells per axis and Lagrange order p has M = p n + 1
// DOFs per axis, DOF index = i_0 M^(d-1) + ... + i_(d-1) (coordinate 0 most
// significant, femlisp's sort-lexicographically order of DOF midpoints).
// Every per-axis relation (which DOFs share a cell, interpolation weights) is
// a 1-D table; d-dimensional rows are tensor products of those tables, so row
// pointers have a closed form and no scan is needed.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "flgpu_internal.h"

namespace flgpu {

constexpr int kMaxDim = 3;

// ---- host-side reference element ----------------------------------------------

// s-point Gauss rule on [0,1] in femlisp's order: Legendre roots ascending,
// points (1 - x_i)/2, weights w_i / 2 (quadrature.lisp:149-176).
static void gauss_rule(int s, double *pts, double *wts) {
  for (int i = 0; i < s; i++) {
    double x = -std::cos(M_PI * (i + 0.75) / (s + 0.5));
    double p = 0, dp = 0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= s; k++) {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      p = (s == 1) ? x : p1;
      double pm1 = (s == 1) ? 1.0 : p0;
      dp = s * (x * p - pm1) / (x * x - 1.0);
      double dx = p / dp;
      x -= dx;
      if (std::fabs(dx) < 1e-17) break;
    }
    {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= s; k++) {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      p = (s == 1) ? x : p1;
      double pm1 = (s == 1) ? 1.0 : p0;
      dp = s * (x * p - pm1) / (x * x - 1.0);
    }
    pts[i] = (1.0 - x) / 2.0;
    wts[i] = (2.0 / ((1.0 - x * x) * dp * dp)) / 2.0;
  }
}

// uniform Lagrange basis of order p on [0,1] and its derivative
static void lagrange1d(int p, double t, double *phi, double *dphi) {
  for (int a = 0; a <= p; a++) {
    double ta = (double)a / p, v = 1.0, d = 0.0;
    for (int b = 0; b <= p; b++)
      if (b != a) v *= (t - (double)b / p) / (ta - (double)b / p);
    for (int m = 0; m <= p; m++) {
      if (m == a) continue;
      double term = 1.0 / (ta - (double)m / p);
      for (int b = 0; b <= p; b++)
        if (b != a && b != m) term *= (t - (double)b / p) / (ta - (double)b / p);
      d += term;
    }
    phi[a] = v;
    if (dphi) dphi[a] = d;
  }
}

static long ipow(long b, int e) {
  long r = 1;
  while (e-- > 0) r *= b;
  return r;
}

struct RefElement {
  int dim, order, s, nq, nl1, nloc, nloc2;
  double pts[8], wts[8];
  // [nq + 1][nloc*nloc]: row q < nq: w_q sum_m dphi_a/dx_m dphi_b/dx_m (unit cube); row nq:
  // the sum of rows 0 .. nq-1 in ascending q (the reference stiffness matrix)
  std::vector<double> T;
  std::vector<double> m;  // [nloc]: sum_q w_q phi_a(q)  (load vector of f = 1)
};

static RefElement ref_element(int dim, int order) {
  RefElement R;
  R.dim = dim;
  R.order = order;
  R.s = order + 1;  // quadrature-rule: (1+ order) points per axis
  R.nq = (int)ipow(R.s, dim);
  R.nl1 = order + 1;
  R.nloc = (int)ipow(R.nl1, dim);
  R.nloc2 = R.nloc * R.nloc;
  gauss_rule(R.s, R.pts, R.wts);
  double phi[8][8], dphi[8][8];
  for (int q = 0; q < R.s; q++) lagrange1d(order, R.pts[q], phi[q], dphi[q]);
  R.T.assign((size_t)(R.nq + 1) * R.nloc2, 0.0);
  R.m.assign(R.nloc, 0.0);
  std::vector<double> g((size_t)R.nloc * kMaxDim), v(R.nloc);
  for (int q = 0; q < R.nq; q++) {
    int qc[kMaxDim];
    for (int k = dim - 1, t = q; k >= 0; k--, t /= R.s) qc[k] = t % R.s;
    double wq = 1.0;
    for (int k = 0; k < dim; k++) wq *= R.wts[qc[k]];
    for (int a = 0; a < R.nloc; a++) {
      int ac[kMaxDim];
      for (int k = dim - 1, t = a; k >= 0; k--, t /= R.nl1) ac[k] = t % R.nl1;
      double val = 1.0;
      for (int k = 0; k < dim; k++) val *= phi[qc[k]][ac[k]];
      v[a] = val;
      for (int mm = 0; mm < dim; mm++) {
        double gg = 1.0;
        for (int k = 0; k < dim; k++) gg *= (k == mm) ? dphi[qc[k]][ac[k]] : phi[qc[k]][ac[k]];
        g[(size_t)a * kMaxDim + mm] = gg;
      }
    }
    for (int a = 0; a < R.nloc; a++) {
      R.m[a] += wq * v[a];
      for (int b = 0; b < R.nloc; b++) {
        double gg = 0.0;
        for (int mm = 0; mm < dim; mm++) gg += g[(size_t)a * kMaxDim + mm] * g[(size_t)b * kMaxDim + mm];
        R.T[(size_t)q * R.nloc2 + (size_t)a * R.nloc + b] = wq * gg;
      }
    }
  }
  R.T.resize((size_t)(R.nq + 1) * R.nloc2, 0.0);
  for (int b = 0; b < R.nloc2; b++) {
    double sum = 0.0;
    for (int q = 0; q < R.nq; q++) sum += R.T[(size_t)q * R.nloc2 + b];
    R.T[(size_t)R.nq * R.nloc2 + b] = sum;
  }
  return R;
}

// ---- 1-D tables --------------------------------------------------------------------

// Per-axis relation as CSR: for each row coordinate, ascending column
// coordinates and weights.
struct Table1D {
  int nr = 0, nc = 0;
  std::vector<int> ptr, col;
  std::vector<double> w;
};

// DOFs sharing a cell along one axis (the pattern)
static Table1D table_pattern(int order, int n) {
  Table1D t;
  int M = order * n + 1;
  t.nr = t.nc = M;
  t.ptr.push_back(0);
  for (int i = 0; i < M; i++) {
    int cmin, cmax;
    if (i % order == 0) {
      cmin = std::max(i / order - 1, 0);
      cmax = std::min(i / order, n - 1);
    } else {
      cmin = cmax = i / order;
    }
    for (int j = order * cmin; j <= order * cmax + order; j++) {
      t.col.push_back(j);
      t.w.push_back(1.0);
    }
    t.ptr.push_back((int)t.col.size());
  }
  return t;
}

// 1-D interpolation, coarse nc cells -> fine 2 nc cells: coarse basis at the
// fine node, entries below 1e-10 dropped (fetransfer.lisp:41-43,120-130)
static Table1D table_interp(int order, int nc) {
  Table1D t;
  int Mf = 2 * order * nc + 1, Mc = order * nc + 1;
  t.nr = Mf;
  t.nc = Mc;
  t.ptr.push_back(0);
  for (int F = 0; F < Mf; F++) {
    int cell = std::min(F / (2 * order), nc - 1);
    double tau = (double)(F - 2 * order * cell) / (2.0 * order);
    double phi[8];
    lagrange1d(order, tau, phi, nullptr);
    for (int a = 0; a <= order; a++) {
      if (std::fabs(phi[a]) < 1e-10) continue;
      t.col.push_back(order * cell + a);
      t.w.push_back(phi[a]);
    }
    t.ptr.push_back((int)t.col.size());
  }
  return t;
}

static Table1D table_transpose(const Table1D &t) {
  Table1D r;
  r.nr = t.nc;
  r.nc = t.nr;
  r.ptr.assign(t.nc + 1, 0);
  for (int c : t.col) r.ptr[c + 1]++;
  for (int c = 0; c < t.nc; c++) r.ptr[c + 1] += r.ptr[c];
  r.col.resize(t.col.size());
  r.w.resize(t.col.size());
  std::vector<int> fill(r.ptr.begin(), r.ptr.end() - 1);
  for (int i = 0; i < t.nr; i++)
    for (int e = t.ptr[i]; e < t.ptr[i + 1]; e++) {
      int p = fill[t.col[e]]++;
      r.col[p] = i;
      r.w[p] = t.w[e];
    }
  return r;
}

struct DevTable {
  int nr, nc, nnz;
  int *ptr, *col;
  double *w;
};

static DevTable upload_table(flgpu_ctx *ctx, const Table1D &t) {
  DevTable d;
  d.nr = t.nr;
  d.nc = t.nc;
  d.nnz = (int)t.col.size();
  d.ptr = (int *)dev_alloc(sizeof(int) * t.ptr.size());
  d.col = (int *)dev_alloc(sizeof(int) * std::max<size_t>(t.col.size(), 1));
  d.w = (double *)dev_alloc(sizeof(double) * std::max<size_t>(t.w.size(), 1));
  FLGPU_HIP(hipMemcpyAsync(d.ptr, t.ptr.data(), sizeof(int) * t.ptr.size(), hipMemcpyHostToDevice,
                           ctx->stream));
  if (!t.col.empty()) {
    FLGPU_HIP(hipMemcpyAsync(d.col, t.col.data(), sizeof(int) * t.col.size(),
                             hipMemcpyHostToDevice, ctx->stream));
    FLGPU_HIP(hipMemcpyAsync(d.w, t.w.data(), sizeof(double) * t.w.size(), hipMemcpyHostToDevice,
                             ctx->stream));
  }
  return d;
}

static void free_table(DevTable &d) {
  dev_free(d.ptr);
  dev_free(d.col);
  dev_free(d.w);
}

// ---- tensor-product CSR ----------------------------------------------------------

// rowptr of row (i_0..i_{d-1}) = sum_k (prod_{l<k} W(i_l)) * P(i_k) * nnz1^(d-1-k)
// with W(i) = ptr[i+1]-ptr[i], P(i) = ptr[i] (prefix of W), nnz1 = ptr[nr].
__global__ void k_tp_rowptr(long N, int dim, DevTable t, int *rowptr) {
  long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > N) return;
  if (r == N) {
    long tot = 1;
    for (int k = 0; k < dim; k++) tot *= t.nnz;
    rowptr[N] = (int)tot;
    return;
  }
  int c[kMaxDim];
  long q = r;
  for (int k = dim - 1; k >= 0; k--) {
    c[k] = (int)(q % t.nr);
    q /= t.nr;
  }
  long off = 0, pre = 1;
  for (int k = 0; k < dim; k++) {
    long post = 1;
    for (int l = k + 1; l < dim; l++) post *= t.nnz;
    off += pre * t.ptr[c[k]] * post;
    pre *= (t.ptr[c[k] + 1] - t.ptr[c[k]]);
  }
  rowptr[r] = (int)off;
}

// Fills columns (and values = product of 1-D weights) of every row.
// zero_rule: 0 = keep values; 1 = interpolation (zero rows whose row
// coordinate is on the boundary and columns on the boundary).
template <class T>
__global__ void k_tp_fill(long N, int dim, DevTable t, const int *rowptr, int *colidx, T *vals,
                          int zero_rule) {
  long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N) return;
  int c[kMaxDim], lo[kMaxDim], hi[kMaxDim], e[kMaxDim];
  long q = r;
  bool rb = false;
  for (int k = dim - 1; k >= 0; k--) {
    c[k] = (int)(q % t.nr);
    q /= t.nr;
    lo[k] = t.ptr[c[k]];
    hi[k] = t.ptr[c[k] + 1];
    e[k] = lo[k];
    if (c[k] == 0 || c[k] == t.nr - 1) rb = true;
  }
  long pos = rowptr[r];
  for (;;) {
    long col = 0;
    double w = 1.0;
    bool cb = false;
    for (int k = 0; k < dim; k++) {
      int cc = t.col[e[k]];
      col = col * t.nc + cc;
      w *= t.w[e[k]];
      if (cc == 0 || cc == t.nc - 1) cb = true;
    }
    colidx[pos] = (int)col;
    if (vals) vals[pos] = (zero_rule == 1 && (rb || cb)) ? (T)0 : (T)w;
    pos++;
    int k = dim - 1;
    while (k >= 0 && e[k] == hi[k] - 1) {
      e[k] = lo[k];
      k--;
    }
    if (k < 0) break;
    e[k]++;
  }
}

// ---- coefficient sampling -----------------------------------------------------------

// S[c][q] = H^(d-2) kappa(x_q): the quadrature point g of cell c samples the
// coefficient cell c*r + floor(g*r) per axis, r = nk/n.
__global__ void k_coef(long ncells, int dim, int n, int s, int nq, const double *__restrict__ g,
                       const double *__restrict__ kappa, int nk, double scale, double *S) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * nq) return;
  long cell = t / nq;
  int q = (int)(t % nq);
  int cc[kMaxDim], qc[kMaxDim];
  long cq = cell;
  for (int k = dim - 1; k >= 0; k--) {
    cc[k] = (int)(cq % n);
    cq /= n;
  }
  for (int k = dim - 1, u = q; k >= 0; k--, u /= s) qc[k] = u % s;
  double kap = 1.0;
  if (kappa) {
    int r = nk / n;
    long idx = 0;
    for (int k = 0; k < dim; k++) idx = idx * nk + (long)cc[k] * r + (int)floor(g[qc[k]] * r);
    kap = kappa[idx];
  }
  S[t] = scale * kap;
}

// ---- element contractions E[c][ab] = sum_q S[c][q] T[q][ab] ---------------------------

template <class TE>
__global__ void k_elem_fma(long ncells, int nq, int nloc2, const double *__restrict__ S,
                           const double *__restrict__ Tt, TE *E) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * nloc2) return;
  long cell = t / nloc2;
  int ab = (int)(t % nloc2);
  double acc = 0.0;
  for (int q = 0; q < nq; q++) acc += S[cell * nq + q] * Tt[(long)q * nloc2 + ab];
  E[t] = (TE)acc;
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

// FP64 MFMA 16x16x4: one wavefront computes a 16-cell x 16-entry tile per
// N-step.  Operand layout of v_mfma_f64_16x16x4f64: lane l supplies
// A[l%16][l/16] and B[l/16][l%16]; it receives D[l/16 + 4v][l%16], v = 0..3
// (measured on gfx950: the four result rows of a lane are 4 apart).
__global__ __launch_bounds__(256) void k_elem_mfma_f64(long ncells, int nq, int nloc2,
                                                       const double *__restrict__ S,
                                                       const double *__restrict__ Tt, double *E) {
  long tile = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  long c0 = tile * 16;
  if (c0 >= ncells) return;
  int lane = threadIdx.x & 63;
  int i = lane & 15, kq = lane >> 4;
  long arow = c0 + i;
  for (int n0 = 0; n0 < nloc2; n0 += 16) {
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    int bcol = n0 + i;
    for (int k0 = 0; k0 < nq; k0 += 4) {
      int k = k0 + kq;
      double a = (arow < ncells && k < nq) ? S[arow * nq + k] : 0.0;
      double b = (k < nq && bcol < nloc2) ? Tt[(long)k * nloc2 + bcol] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    for (int v = 0; v < 4; v++) {
      long row = c0 + kq + 4 * v;
      if (row < ncells && bcol < nloc2) E[row * nloc2 + bcol] = acc[v];
    }
  }
}

__device__ __forceinline__ short f2bf(double x) {
  float f = (float)x;
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (short)(u >> 16);
}

__device__ __forceinline__ float bf2f(short h) {
  return __uint_as_float(((unsigned)(unsigned short)h) << 16);
}

// v = h + m + l + e with BF16 parts, each the BF16 rounding of what the previous ones leave
// (the remainders are exact in FP32), |e| <= 2^-25 |v|: the 24 bits of an FP32 value
__device__ __forceinline__ void bf16_split3(float v, short &h, short &m, short &l) {
  h = f2bf(v);
  const float r = v - bf2f(h);
  m = f2bf(r);
  l = f2bf(r - bf2f(m));
}

// acc += A B for operands split in three BF16 parts: the six products down to 2^-16 of the
// leading one, the smallest first (al bh, am bm, ah bl, am bh, ah bm, ah bh)
__device__ __forceinline__ v4f mfma_bf16x3(v4s ah, v4s am, v4s al, v4s bh, v4s bm, v4s bl,
                                           v4f acc) {
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, bm, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bm, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc, 0, 0, 0);
  return acc;
}

// BF16 MFMA 16x16x16 (FP32 accumulate): lane l supplies A[l%16][4(l/16)+v]
// and B[4(l/16)+v][l%16] (v = 0..3); it receives D[4(l/16)+v][l%16].
// Every operand is split into three BF16 parts (bf16_split3) and the six
// products down to 2^-16 of the leading one are accumulated (mfma_bf16x3):
// about 24 significant bits per product, as in FP32, instead of 8.
// The contraction is E = S_0 K + sum_{q>0} (S_q - S_0) T_q, K = sum_q T_q (row nq of the
// table; the same sum, regrouped): quadrature point 0 carries the cell's value against K and
// the others only its variation, so a cell with one coefficient value gets E = S_0 K with
// one rounding per entry, and the entries where K vanishes (the axis neighbours of 3D Q1)
// are exact zeros instead of the FP32 accumulator's cancellation error.
__global__ __launch_bounds__(256) void k_elem_mfma_bf16(long ncells, int nq, int nloc2,
                                                        const double *__restrict__ S,
                                                        const double *__restrict__ Tt, float *E) {
  long tile = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  long c0 = tile * 16;
  if (c0 >= ncells) return;
  int lane = threadIdx.x & 63;
  int i = lane & 15, kg = lane >> 4;
  long arow = c0 + i;
  for (int n0 = 0; n0 < nloc2; n0 += 16) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    int bcol = n0 + i;
    for (int k0 = 0; k0 < nq; k0 += 16) {
      v4s ah, am, al, bh, bm, bl;
      for (int v = 0; v < 4; v++) {
        int k = k0 + 4 * kg + v;
        const double s0 = arow < ncells ? S[arow * nq] : 0.0;
        float av = (float)((arow < ncells && k < nq) ? (k == 0 ? s0 : S[arow * nq + k] - s0)
                                                     : 0.0);
        float bv = (float)((k < nq && bcol < nloc2) ? Tt[(long)(k == 0 ? nq : k) * nloc2 + bcol]
                                                    : 0.0);
        short h, m, l;
        bf16_split3(av, h, m, l);
        ah[v] = h;
        am[v] = m;
        al[v] = l;
        bf16_split3(bv, h, m, l);
        bh[v] = h;
        bm[v] = m;
        bl[v] = l;
      }
      acc = mfma_bf16x3(ah, am, al, bh, bm, bl, acc);
    }
    for (int v = 0; v < 4; v++) {
      long row = c0 + 4 * kg + v;
      if (row < ncells && bcol < nloc2) E[row * nloc2 + bcol] = acc[v];
    }
  }
}

// ---- gather (assembly of the global matrix from element matrices) -------------------

// One thread per row: entry (P,Q) = sum over cells containing both DOFs, in
// ascending cell order, of E[cell][loc(P)][loc(Q)]; then Dirichlet
// elimination (identity on constrained rows, zeros in their columns) and the
// load vector rhs_P = f H^d sum_cells m[loc(P)].
template <class T, class TE>
__global__ void k_gather(long N, int dim, int order, int n, DevTable pat, const int *rowptr,
                         const TE *__restrict__ E, int nloc, T *vals, double rhs_scale,
                         const double *__restrict__ mload, T *rhs) {
  long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N) return;
  int M = pat.nr, nl1 = order + 1;
  int P[kMaxDim], lo[kMaxDim], hi[kMaxDim], e[kMaxDim];
  long q = r;
  bool rb = false;
  for (int k = dim - 1; k >= 0; k--) {
    P[k] = (int)(q % M);
    q /= M;
    lo[k] = pat.ptr[P[k]];
    hi[k] = pat.ptr[P[k] + 1];
    e[k] = lo[k];
    if (P[k] == 0 || P[k] == M - 1) rb = true;
  }
  long nloc2 = (long)nloc * nloc;
  long pos = rowptr[r];
  for (;;) {
    int Q[kMaxDim], clo[kMaxDim], chi[kMaxDim], c[kMaxDim];
    bool cb = false, diag = true;
    for (int k = 0; k < dim; k++) {
      Q[k] = pat.col[e[k]];
      if (Q[k] == 0 || Q[k] == M - 1) cb = true;
      if (Q[k] != P[k]) diag = false;
      int mx = max(P[k], Q[k]), mn = min(P[k], Q[k]);
      // cells c with order*c <= mn and mx <= order*c + order
      clo[k] = (mx <= order) ? 0 : (mx - 1) / order;
      chi[k] = min(mn / order, n - 1);
      c[k] = clo[k];
    }
    double acc = 0.0;
    if (!(rb || cb)) {
      for (;;) {
        long cell = 0;
        int la = 0, lb = 0;
        for (int k = 0; k < dim; k++) {
          cell = cell * n + c[k];
          la = la * nl1 + (P[k] - order * c[k]);
          lb = lb * nl1 + (Q[k] - order * c[k]);
        }
        acc += (double)E[cell * nloc2 + (long)la * nloc + lb];
        int k = dim - 1;
        while (k >= 0 && c[k] == chi[k]) {
          c[k] = clo[k];
          k--;
        }
        if (k < 0) break;
        c[k]++;
      }
    } else {
      acc = (rb && diag) ? 1.0 : 0.0;
    }
    vals[pos++] = (T)acc;
    int k = dim - 1;
    while (k >= 0 && e[k] == hi[k] - 1) {
      e[k] = lo[k];
      k--;
    }
    if (k < 0) break;
    e[k]++;
  }
  if (rhs) {
    double b = 0.0;
    if (!rb) {
      int clo[kMaxDim], chi[kMaxDim], c[kMaxDim];
      for (int k = 0; k < dim; k++) {
        clo[k] = (P[k] % order == 0) ? max(P[k] / order - 1, 0) : P[k] / order;
        chi[k] = (P[k] % order == 0) ? min(P[k] / order, n - 1) : P[k] / order;
        c[k] = clo[k];
      }
      for (;;) {
        int la = 0;
        for (int k = 0; k < dim; k++) la = la * nl1 + (P[k] - order * c[k]);
        b += mload[la];
        int k = dim - 1;
        while (k >= 0 && c[k] == chi[k]) {
          c[k] = clo[k];
          k--;
        }
        if (k < 0) break;
        c[k]++;
      }
      b *= rhs_scale;
    }
    rhs[r] = (T)b;
  }
}

// ---- 3D Q1 assembly in one pass, element matrices in LDS only ---------------------------
// A workgroup owns a 4 x 4 x 4 block of lattice nodes.  It samples the coefficient at the
// quadrature points of the 5 x 5 x 5 cells around the block (S, as k_coef), forms their element
// matrices E = S T in LDS with the contraction of the chosen mode (k_elem_fma's loop, or the
// 16 x 16 MFMA tiles of k_elem_mfma_f64 / k_elem_mfma_bf16 with the same operands in the same
// lanes), and then forms its nodes' CSR rows from LDS exactly as k_gather does: each entry is
// the sum, in ascending cell order and in FP64, of E over the cells holding both DOFs, with
// femlisp's Dirichlet elimination, and the load vector likewise.  So the values equal the
// two-pass assembly's bit for bit, while the element matrices never reach device memory.
constexpr int kAsmB = 4;                         // nodes per axis of a block
constexpr int kAsmC = kAsmB + 1;                 // cells per axis around it
constexpr int kAsmNC = kAsmC * kAsmC * kAsmC;    // 125 cells
constexpr int kAsmNCP = (kAsmNC + 15) / 16 * 16; // padded to MFMA tiles of 16 cells
constexpr int kAsmThreads = 256;                 // four threads per node of the block
constexpr int kAsmEsW = 65;  // row stride of the element matrices in LDS: one bank per cell

enum { ASM_FMA = 0, ASM_MFMA64 = 1, ASM_MFMA_BF16 = 2 };

template <class T, int MODE>
__global__ __launch_bounds__(kAsmThreads) void k_asm_q1(int n, const double *__restrict__ g,
                                                const double *__restrict__ kappa, int nk,
                                                double scale, const double *__restrict__ Tt,
                                                const int *__restrict__ rowptr, T *vals,
                                                double rhs_scale, const double *__restrict__ mload,
                                                T *rhs) {
  // E in the precision the two-pass path stores it in: FP64 for FP64 matrices, FP32 otherwise;
  // the rows of E are kAsmEsW apart, so that the nodes of a wave, whose cells differ, read
  // different banks.  S is kept in FP64 (FMA and FP64 MFMA) or as its three BF16 parts
  // (BF16 MFMA, bf16_split3), formed once per value.
  constexpr bool kBF = MODE == ASM_MFMA_BF16;
  constexpr int kSBytes = kBF ? kAsmNCP * 8 * 3 * (int)sizeof(short) : kAsmNCP * 8 * (int)sizeof(double);
  __shared__ __align__(16) unsigned char smem[kSBytes + kAsmNC * kAsmEsW * sizeof(T)];
  double(*Ss)[8] = (double(*)[8])smem;
  short(*Sh)[8] = (short(*)[8])smem;                       // BF16 parts: high
  short(*Sm)[8] = (short(*)[8])(smem + kSBytes / 3);       // middle
  short(*Sl)[8] = (short(*)[8])(smem + 2 * kSBytes / 3);   // low
  T(*Es)[kAsmEsW] = (T(*)[kAsmEsW])(smem + kSBytes);
  const int M = n + 1;
  const int nbpa = (M + kAsmB - 1) / kAsmB;
  int blk = blockIdx.x;
  const int bx = blk % nbpa;
  blk /= nbpa;
  const int by = blk % nbpa, bz = blk / nbpa;
  const int cz0 = bz * kAsmB - 1, cy0 = by * kAsmB - 1, cx0 = bx * kAsmB - 1;  // first cell
  const int tid = threadIdx.x;
  // S[c][q] = scale kappa(x_q) (k_coef); zero for cells outside the lattice (never read).
  // A thread fills kAsmFill entries; their coefficient loads are all issued before any is used.
  constexpr int kAsmFill = kAsmNCP * 8 / kAsmThreads;
  {
    const int r = nk / n;
    int gq[2] = {0, 0};  // the quadrature points' sub-cell offsets floor(g r)
    if (kappa) {
      gq[0] = (int)floor(g[0] * r);
      gq[1] = (int)floor(g[1] * r);
    }
    double kv[kAsmFill];
    bool in[kAsmFill];
#pragma unroll
    for (int u = 0; u < kAsmFill; u++) {
      const int e = tid + u * kAsmThreads, lc = e >> 3, q = e & 7;
      const int cz = cz0 + lc / (kAsmC * kAsmC), cy = cy0 + (lc / kAsmC) % kAsmC,
                cx = cx0 + lc % kAsmC;
      in[u] = lc < kAsmNC && cz >= 0 && cy >= 0 && cx >= 0 && cz < n && cy < n && cx < n;
      long idx = 0;  // cells outside the lattice load the first value, which is not used
      if (in[u]) {
        idx = (long)cz * r + gq[q >> 2];
        idx = idx * nk + (long)cy * r + gq[(q >> 1) & 1];
        idx = idx * nk + (long)cx * r + gq[q & 1];
      }
      kv[u] = kappa ? kappa[idx] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < kAsmFill; u++) {
      const int e = tid + u * kAsmThreads;
      const double sv = in[u] ? scale * kv[u] : 0.0;
      if (kBF) {
        // k_elem_mfma_bf16's operand: S_0 at q = 0, S_q - S_0 otherwise (the 8 values of a
        // cell are in 8 consecutive lanes, q = lane mod 8)
        const double s0 = __shfl(sv, (tid & 63) & ~7, 64);
        short h, m, l;
        bf16_split3((float)((e & 7) == 0 ? s0 : sv - s0), h, m, l);
        Sh[e >> 3][e & 7] = h;
        Sm[e >> 3][e & 7] = m;
        Sl[e >> 3][e & 7] = l;
      } else {
        Ss[e >> 3][e & 7] = sv;
      }
    }
  }
  __syncthreads();
  if (MODE == ASM_FMA) {
    // entry ab = tid mod 64 of every cell this thread forms: its 8 values of T stay in registers
    const int ab = tid & 63;
    double tq[8];
#pragma unroll
    for (int q = 0; q < 8; q++) tq[q] = Tt[q * 64 + ab];
    for (int lc = tid >> 6; lc < kAsmNC; lc += kAsmThreads / 64) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < 8; q++) acc += Ss[lc][q] * tq[q];
      Es[lc][ab] = (T)acc;
    }
  } else {
    // tiles of 16 cells x 16 entries; wave w forms the entries 16 w .. 16 w + 15 of every cell
    // (tile jobs w, w + 4, ...), so its operand from T is the same for all its tiles
    static_assert(kAsmThreads == 256, "four waves, one per 16-entry column of T");
    const int lane = tid & 63, w = tid >> 6;
    const int i = lane & 15, kg = lane >> 4;
    const int bcol = w * 16 + i;
    if (MODE == ASM_MFMA64) {
      const double b0 = Tt[kg * 64 + bcol], b1 = Tt[(4 + kg) * 64 + bcol];
      for (int c0 = 0; c0 < kAsmNCP; c0 += 16) {
        v4d acc = {0.0, 0.0, 0.0, 0.0};
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Ss[c0 + i][kg], b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Ss[c0 + i][4 + kg], b1, acc, 0, 0, 0);
        for (int v = 0; v < 4; v++)
          if (c0 + kg + 4 * v < kAsmNC) Es[c0 + kg + 4 * v][bcol] = (T)acc[v];
      }
    } else {
      v4s bh, bm, bl;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int k = 4 * kg + v;
        short h, m, l;
        bf16_split3((float)(k < 8 ? Tt[(k == 0 ? 8 : k) * 64 + bcol] : 0.0), h, m, l);
        bh[v] = h;
        bm[v] = m;
        bl[v] = l;
      }
      for (int c0 = 0; c0 < kAsmNCP; c0 += 16) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        const v4s zero = {0, 0, 0, 0};  // k >= 8: the padding of K to 16
        const v4s ah = kg < 2 ? *(const v4s *)&Sh[c0 + i][4 * kg] : zero;
        const v4s am = kg < 2 ? *(const v4s *)&Sm[c0 + i][4 * kg] : zero;
        const v4s al = kg < 2 ? *(const v4s *)&Sl[c0 + i][4 * kg] : zero;
        acc = mfma_bf16x3(ah, am, al, bh, bm, bl, acc);
        for (int v = 0; v < 4; v++)
          if (c0 + 4 * kg + v < kAsmNC) Es[c0 + 4 * kg + v][bcol] = (T)acc[v];
      }
    }
  }
  __syncthreads();
  // rows: four threads per node.  An interior row has all 27 neighbours (entry k at
  // rowptr + k) and all 8 cells around it, so thread p < 3 forms the 9 entries of z-offset
  // p - 1 with every cell and local index a constant offset from the node's (the loops over
  // the offsets and the cells are unrolled), and thread 3 the load vector.  A Dirichlet row
  // (on the lattice's boundary) is the identity in its own column and zero elsewhere, with
  // fewer neighbours: its threads take every fourth entry at its position among the present
  // ones.  Each entry is the sum, in ascending cell order and in FP64, of E over the cells
  // holding both DOFs, as k_gather forms it.
  const int node = tid >> 2, part = tid & 3;
  const int lz = node / (kAsmB * kAsmB), ly = (node / kAsmB) % kAsmB, lx = node % kAsmB;
  const int Pz = bz * kAsmB + lz, Py = by * kAsmB + ly, Px = bx * kAsmB + lx;
  if (Pz >= M || Py >= M || Px >= M) return;
  const long r = ((long)Pz * M + Py) * M + Px;
  const bool rb = Pz == 0 || Py == 0 || Px == 0 || Pz == M - 1 || Py == M - 1 || Px == M - 1;
  const int pos0 = rowptr[r];
  if (rb) {
    // the row's neighbours inside the lattice, in ascending column order: entry position =
    // the number of such neighbours before it
    unsigned live = 0;
    for (int k = 0; k < 27; k++) {
      const int Qz = Pz + k / 9 - 1, Qy = Py + (k / 3) % 3 - 1, Qx = Px + k % 3 - 1;
      if (Qz >= 0 && Qy >= 0 && Qx >= 0 && Qz < M && Qy < M && Qx < M) live |= 1u << k;
    }
    for (int k = part; k < 27; k += 4)
      if (live >> k & 1u) vals[pos0 + __popc(live & ((1u << k) - 1u))] = (T)(k == 13 ? 1.0 : 0.0);
    if (rhs && part == 0) rhs[r] = (T)0.0;
    return;
  }
  if (part < 3) {
    const int dz = part - 1;
    // the cell (P - 1) in the block's cell grid: cells start at (bz, by, bx) kAsmB - 1
    const int lbase = (lz * kAsmC + ly) * kAsmC + lx;
    const bool qz = (Pz + dz == 0) || (Pz + dz == M - 1);
#pragma unroll
    for (int e = 0; e < 9; e++) {
      const int dy = e / 3 - 1, dx = e % 3 - 1;
      const bool cb = qz || Py + dy == 0 || Py + dy == M - 1 || Px + dx == 0 || Px + dx == M - 1;
      double acc = 0.0;
      if (!cb) {
        // shared cells per axis: offset -1 -> the lower cell (a = 0), +1 -> the upper (a = 1),
        // 0 -> both; ascending (z slowest)
#pragma unroll
        for (int a = 0; a < 2; a++) {
          if ((dz < 0 && a == 1) || (dz > 0 && a == 0)) continue;
#pragma unroll
          for (int b = (dy > 0 ? 1 : 0); b <= (dy < 0 ? 0 : 1); b++)
#pragma unroll
            for (int c = (dx > 0 ? 1 : 0); c <= (dx < 0 ? 0 : 1); c++) {
              const int lc = lbase + (a * kAsmC + b) * kAsmC + c;
              const int la = ((1 - a) * 2 + (1 - b)) * 2 + (1 - c);
              const int lb = ((dz + 1 - a) * 2 + (dy + 1 - b)) * 2 + (dx + 1 - c);
              acc += (double)Es[lc][la * 8 + lb];
            }
        }
      }
      vals[pos0 + part * 9 + e] = (T)acc;
    }
  } else if (rhs) {
    double bsum = 0.0;
    for (int cz = Pz - 1; cz <= Pz; cz++)
      for (int cy = Py - 1; cy <= Py; cy++)
        for (int cx = Px - 1; cx <= Px; cx++)
          bsum += mload[((Pz - cz) * 2 + (Py - cy)) * 2 + (Px - cx)];
    rhs[r] = (T)(bsum * rhs_scale);
  }
}

// ---- drivers ---------------------------------------------------------------------------

struct LatticeBuild {
  RefElement R;
  Table1D pat1;
};

static void check_lattice_args(int dim, int order, int n) {
  FLGPU_REQUIRE(dim >= 1 && dim <= 3, "dim must be 1, 2 or 3");
  FLGPU_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
  FLGPU_REQUIRE(n >= 1, "n must be >= 1");
  long N = ipow((long)order * n + 1, dim);
  long nnz = ipow((long)(order == 1 ? 3 * n + 1 : 8 * n + 1), dim);
  FLGPU_REQUIRE(N < 0x7fffffffL && nnz < 0x7fffffffL, "lattice too large for int32 indices");
}

static void check_kappa(const flgpu_vec *kappa, int dim, int n, int nk) {
  if (!kappa) return;
  FLGPU_REQUIRE(kappa->prec == FLGPU_F64, "kappa must be an FP64 vector");
  FLGPU_REQUIRE(nk >= n && nk % n == 0, "nk must be a multiple of n");
  FLGPU_REQUIRE(kappa->n == ipow(nk, dim), "kappa must hold nk^dim values");
}

// element matrices of all cells of the level into E (TE = double or float)
static void element_matrices(flgpu_ctx *ctx, const RefElement &R, int n, const flgpu_vec *kappa,
                             int nk, int mode, int prec, void *E) {
  int dim = R.dim;
  long ncells = ipow(n, dim);
  double H = 1.0 / n;
  double scale = std::pow(H, dim - 2);
  double *dS = (double *)dev_alloc(sizeof(double) * ncells * R.nq);
  double *dT = (double *)dev_alloc(sizeof(double) * R.T.size());
  double *dg = (double *)dev_alloc(sizeof(double) * R.s);
  FLGPU_HIP(hipMemcpyAsync(dT, R.T.data(), sizeof(double) * R.T.size(), hipMemcpyHostToDevice,
                           ctx->stream));
  FLGPU_HIP(hipMemcpyAsync(dg, R.pts, sizeof(double) * R.s, hipMemcpyHostToDevice, ctx->stream));
  k_coef<<<grid_for(ncells * R.nq, 256), 256, 0, ctx->stream>>>(
      ncells, dim, n, R.s, R.nq, dg, kappa ? (const double *)kappa->d : nullptr, nk, scale, dS);
  FLGPU_HIP(hipGetLastError());
  if (mode == FLGPU_ASM_MFMA) {
    unsigned g = grid_for((ncells + 15) / 16, 4);
    if (prec == FLGPU_F64)
      k_elem_mfma_f64<<<g, 256, 0, ctx->stream>>>(ncells, R.nq, R.nloc2, dS, dT, (double *)E);
    else
      k_elem_mfma_bf16<<<g, 256, 0, ctx->stream>>>(ncells, R.nq, R.nloc2, dS, dT, (float *)E);
  } else {
    if (prec == FLGPU_F64)
      k_elem_fma<double><<<grid_for(ncells * R.nloc2, 256), 256, 0, ctx->stream>>>(
          ncells, R.nq, R.nloc2, dS, dT, (double *)E);
    else
      k_elem_fma<float><<<grid_for(ncells * R.nloc2, 256), 256, 0, ctx->stream>>>(
          ncells, R.nq, R.nloc2, dS, dT, (float *)E);
  }
  FLGPU_HIP(hipGetLastError());
  FLGPU_HIP(hipStreamSynchronize(ctx->stream));
  dev_free(dS);
  dev_free(dT);
  dev_free(dg);
}

// 3D Q1: the one-pass assembly (k_asm_q1)
static void assemble_values_q1(flgpu_mat *A, const flgpu_vec *kappa, int nk, int mode,
                               void *rhs_dev) {
  flgpu_ctx *ctx = A->ctx;
  const int n = A->lat_n;
  RefElement R = ref_element(3, 1);
  double *dT = (double *)dev_alloc(sizeof(double) * R.T.size());
  double *dg = nullptr, *dm = nullptr;
  try {
    dg = (double *)dev_alloc(sizeof(double) * R.s);
    dm = (double *)dev_alloc(sizeof(double) * R.nloc);
    FLGPU_HIP(hipMemcpyAsync(dT, R.T.data(), sizeof(double) * R.T.size(), hipMemcpyHostToDevice,
                             ctx->stream));
    FLGPU_HIP(hipMemcpyAsync(dg, R.pts, sizeof(double) * R.s, hipMemcpyHostToDevice,
                             ctx->stream));
    FLGPU_HIP(hipMemcpyAsync(dm, R.m.data(), sizeof(double) * R.nloc, hipMemcpyHostToDevice,
                             ctx->stream));
    const int nbpa = (n + 1 + kAsmB - 1) / kAsmB;
    const unsigned g = (unsigned)nbpa * nbpa * nbpa;
    const double scale = 1.0 / n;  // H^(d-2)
    const double rscale = A->lat_f * std::pow(1.0 / n, 3);
    const double *kd = kappa ? (const double *)kappa->d : nullptr;
    hipStream_t st = ctx->stream;
#define FLGPU_ASM(T, MODE)                                                                   \
  k_asm_q1<T, MODE><<<g, kAsmThreads, 0, st>>>(n, dg, kd, nk, scale, dT, A->rowptr, (T *)A->vals, \
                                       rscale, dm, (T *)rhs_dev)
    if (A->prec == FLGPU_F64) {
      if (mode == FLGPU_ASM_MFMA) FLGPU_ASM(double, ASM_MFMA64);
      else FLGPU_ASM(double, ASM_FMA);
    } else {
      if (mode == FLGPU_ASM_MFMA) FLGPU_ASM(float, ASM_MFMA_BF16);
      else FLGPU_ASM(float, ASM_FMA);
    }
#undef FLGPU_ASM
    FLGPU_HIP(hipGetLastError());
    FLGPU_HIP(hipStreamSynchronize(st));
  } catch (...) {
    dev_free(dT);
    dev_free(dg);
    dev_free(dm);
    throw;
  }
  dev_free(dT);
  dev_free(dg);
  dev_free(dm);
}

// values (and rhs) of a lattice matrix whose pattern is already built
static void assemble_values(flgpu_mat *A, const flgpu_vec *kappa, int nk, int mode,
                            void *rhs_dev) {
  flgpu_ctx *ctx = A->ctx;
  int dim = A->lat_dim, order = A->lat_order, n = A->lat_n;
  if (dim == 3 && order == 1) {
    assemble_values_q1(A, kappa, nk, mode, rhs_dev);
    return;
  }
  RefElement R = ref_element(dim, order);
  Table1D pat = table_pattern(order, n);
  long ncells = ipow(n, dim);
  // E in FP64 for the FP64 path, FP32 for the FP32 (BF16-MFMA) path
  int eprec = A->prec;
  void *E = dev_alloc(prec_bytes(eprec) * ncells * R.nloc2);
  element_matrices(ctx, R, n, kappa, nk, mode, eprec, E);
  DevTable dpat = upload_table(ctx, pat);
  double *dm = (double *)dev_alloc(sizeof(double) * R.nloc);
  FLGPU_HIP(hipMemcpyAsync(dm, R.m.data(), sizeof(double) * R.nloc, hipMemcpyHostToDevice,
                           ctx->stream));
  double rscale = A->lat_f * std::pow(1.0 / n, dim);
  long N = A->nrows;
  if (A->prec == FLGPU_F64)
    k_gather<double, double><<<grid_for(N, 128), 128, 0, ctx->stream>>>(
        N, dim, order, n, dpat, A->rowptr, (const double *)E, R.nloc, (double *)A->vals, rscale,
        dm, (double *)rhs_dev);
  else
    k_gather<float, float><<<grid_for(N, 128), 128, 0, ctx->stream>>>(
        N, dim, order, n, dpat, A->rowptr, (const float *)E, R.nloc, (float *)A->vals, rscale, dm,
        (float *)rhs_dev);
  FLGPU_HIP(hipGetLastError());
  FLGPU_HIP(hipStreamSynchronize(ctx->stream));
  free_table(dpat);
  dev_free(dm);
  dev_free(E);
}

template <class T>
__global__ void k_alpha(long ncells, double H, const double *__restrict__ kappa, T *alpha) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) alpha[c] = (T)(H * (kappa ? kappa[c] : 1.0));
}

// 3D Q1 with one coefficient value per cell (nk == n): the element matrix of
// cell c is H kappa_c Kref, Kref = sum_q T_q, and by the reflection symmetry
// of the cube Kref[a][b] = Kref[0][a xor b].  Stores alpha_c = H kappa_c and
// Kref[0][.] so that the kernels apply A without streaming its values.
static void lattice_set_alpha(flgpu_mat *A, const flgpu_vec *kappa, int nk) {
  int n = A->lat_n;
  if (A->lat_dim != 3 || A->lat_order != 1 || (kappa && nk != n)) {
    drop_alpha(A);
    return;
  }
  RefElement R = ref_element(3, 1);
  for (int b = 0; b < 8; b++) A->kt[b] = R.T[(size_t)R.nq * R.nloc2 + b];  // sum_q T_q
  long ncells = (long)n * n * n;
  if (!A->alpha) {
    A->alpha = dev_alloc(prec_bytes(A->prec) * ncells);
    sor_strips_alloc(A);
    layout_changed(A);
  }
  const double *kd = kappa ? (const double *)kappa->d : nullptr;
  if (A->prec == FLGPU_F64)
    k_alpha<double><<<grid_for(ncells, 256), 256, 0, A->ctx->stream>>>(ncells, 1.0 / n, kd,
                                                                       (double *)A->alpha);
  else
    k_alpha<float><<<grid_for(ncells, 256), 256, 0, A->ctx->stream>>>(ncells, 1.0 / n, kd,
                                                                      (float *)A->alpha);
  FLGPU_HIP(hipGetLastError());
}

// The columns of the 3D Q1 lattice pattern (every lattice neighbour, ascending), one lane per
// neighbour and 32 lanes per row: lane k writes neighbour k at its position among the present
// ones, so a row's columns are written side by side.  The same pattern as k_tp_fill's tensor
// product of the 1-D pattern {i - 1, i, i + 1}.
__global__ __launch_bounds__(256) void k_q1_pattern(int M, const int *__restrict__ rowptr,
                                                    int *colidx) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long N = (long)M * M * M;
  const long r = t >> 5;
  const int k = (int)(t & 31);
  if (r >= N || k >= 27) return;
  const int x = (int)(r % M), y = (int)((r / M) % M), z = (int)(r / ((long)M * M));
  unsigned live = 0;
#pragma unroll
  for (int j = 0; j < 27; j++) {
    const int qz = z + j / 9 - 1, qy = y + (j / 3) % 3 - 1, qx = x + j % 3 - 1;
    if (qz >= 0 && qy >= 0 && qx >= 0 && qz < M && qy < M && qx < M) live |= 1u << j;
  }
  if (!(live >> k & 1u)) return;
  const long c = r + (long)(k / 9 - 1) * M * M + (long)((k / 3) % 3 - 1) * M + (k % 3 - 1);
  colidx[rowptr[r] + __popc(live & ((1u << k) - 1u))] = (int)c;
}

flgpu_mat *lattice_pattern(flgpu_ctx *ctx, int dim, int order, int n, int prec) {
  check_lattice_args(dim, order, n);
  Table1D pat = table_pattern(order, n);
  long N = ipow(pat.nr, dim);
  long nnz = ipow((long)pat.col.size(), dim);
  flgpu_mat *A = mat_new(ctx, (int)N, (int)N, nnz, prec);
  A->lat_dim = dim;
  A->lat_order = order;
  A->lat_n = n;
  DevTable d = upload_table(ctx, pat);
  k_tp_rowptr<<<grid_for(N + 1, 256), 256, 0, ctx->stream>>>(N, dim, d, A->rowptr);
  if (dim == 3 && order == 1)
    k_q1_pattern<<<grid_for(32 * N, 256), 256, 0, ctx->stream>>>(n + 1, A->rowptr, A->colidx);
  else
    k_tp_fill<double><<<grid_for(N, 128), 128, 0, ctx->stream>>>(N, dim, d, A->rowptr,
                                                                 A->colidx, nullptr, 0);
  FLGPU_HIP(hipGetLastError());
  FLGPU_HIP(hipStreamSynchronize(ctx->stream));
  free_table(d);
  return A;
}

flgpu_mat *lattice_assemble(flgpu_ctx *ctx, int dim, int order, int n, const flgpu_vec *kappa,
                            int nk, double f, int prec, int mode, flgpu_vec **rhs) {
  check_lattice_args(dim, order, n);
  check_kappa(kappa, dim, n, nk);
  FLGPU_REQUIRE(mode == FLGPU_ASM_F64 || mode == FLGPU_ASM_MFMA, "unknown assembly mode");
  flgpu_mat *A = lattice_pattern(ctx, dim, order, n, prec);
  A->lat_f = f;
  flgpu_vec *b = nullptr;
  try {
    if (rhs) b = vec_new(ctx, A->nrows, prec);
    assemble_values(A, kappa, nk, mode, b ? b->d : nullptr);
    lattice_set_alpha(A, kappa, nk);
    mat_finish(A);
  } catch (...) {
    vec_free(b);
    mat_free(A);
    throw;
  }
  if (rhs) *rhs = b;
  return A;
}

// + identity on the diagonal of constrained (boundary) rows: m+! P after the
// elimination of the Dirichlet constraints (geomg.lisp:128-136)
template <class T>
__global__ void k_dirichlet_diag(long N, int dim, int M, const int *rowptr, const int *colidx,
                                 T *vals) {
  long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N) return;
  long q = r;
  bool b = false;
  for (int k = 0; k < dim; k++) {
    int c = (int)(q % M);
    q /= M;
    if (c == 0 || c == M - 1) b = true;
  }
  if (!b) return;
  for (int e = rowptr[r]; e < rowptr[r + 1]; e++)
    if (colidx[e] == r) vals[e] += (T)1;
}

void lattice_dirichlet_identity(flgpu_mat *A, int dim, int order, int n) {
  long N = A->nrows;
  int M = order * n + 1;
  if (A->prec == FLGPU_F64)
    k_dirichlet_diag<double><<<grid_for(N, 256), 256, 0, A->ctx->stream>>>(
        N, dim, M, A->rowptr, A->colidx, (double *)A->vals);
  else
    k_dirichlet_diag<float><<<grid_for(N, 256), 256, 0, A->ctx->stream>>>(
        N, dim, M, A->rowptr, A->colidx, (float *)A->vals);
  FLGPU_HIP(hipGetLastError());
}

// New values for a new kappa in the same pattern.  A matrix applied in the per-cell form gets
// the new cell factors with its values (or drops that form when kappa is no longer constant
// per cell: nk != n); a matrix whose form was capped below it stays capped.
void lattice_reassemble(flgpu_mat *A, const flgpu_vec *kappa, int nk, int mode) {
  check_kappa(kappa, A->lat_dim, A->lat_n, nk);
  assemble_values(A, kappa, nk, mode, nullptr);
  if (A->alpha) lattice_set_alpha(A, kappa, nk);
  mat_refresh_values(A);
}

// Coarse lattice matrix as the Galerkin product P^T A_fine P
// (geomg.lisp:122-123 with :galerkin-p t; amg.lisp:247-280)
flgpu_mat *lattice_galerkin(const flgpu_mat *P, const flgpu_mat *Afine, int dim, int order,
                            int nc) {
  flgpu_mat *Ac = lattice_pattern(Afine->ctx, dim, order, nc, Afine->prec);
  try {
    galerkin_product(P, Afine, Ac, true);
    long N = Ac->nrows;
    int M = order * nc + 1;
    if (Ac->prec == FLGPU_F64)
      k_dirichlet_diag<double><<<grid_for(N, 256), 256, 0, Ac->ctx->stream>>>(
          N, dim, M, Ac->rowptr, Ac->colidx, (double *)Ac->vals);
    else
      k_dirichlet_diag<float><<<grid_for(N, 256), 256, 0, Ac->ctx->stream>>>(
          N, dim, M, Ac->rowptr, Ac->colidx, (float *)Ac->vals);
    FLGPU_HIP(hipGetLastError());
    mat_finish(Ac);
  } catch (...) {
    mat_free(Ac);
    throw;
  }
  return Ac;
}

// ---- operator-dependent ("black-box") interpolation, 3D Q1 ------------------------------
// Not in femlisp: J. E. Dendy's black-box prolongation (J. Comput. Phys. 48 (1982)
// 366-386) for 27-point stencils.  W[b][i] is the weight of fine node i on corner b
// (bits z,y,x) of the coarse cell whose lowest corner is (z>>1, y>>1, x>>1).  A node with
// s odd coordinates solves its row, collapsed onto its odd axes, for itself:
//   e_i = sum_k (-a_k / a0) e_(i + k masked to the odd axes),
//   a0 = sum of the a_k whose offset vanishes on the odd axes,
// with the weights of the neighbours, which have fewer odd coordinates and come from an
// earlier pass.  Boundary nodes keep W = 0: fine Dirichlet rows and coarse Dirichlet
// columns of the interpolation are zero, as in femlisp's constrained interpolation
// (constraints.lisp:58-72).  An interior row of a lattice matrix holds its 27 neighbours
// in ascending order, so entry k of the row is offset k.
template <class T>
__global__ __launch_bounds__(256) void k_bbox_pass(int M, int pass, const int *__restrict__ rowptr,
                                                   const T *__restrict__ vals, double *W) {
  long N = (long)M * M * M;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int x = (int)(i % M), y = (int)((i / M) % M), z = (int)(i / ((long)M * M));
  int oz = z & 1, oy = y & 1, ox = x & 1;
  if (oz + oy + ox != pass) return;
  if (z == 0 || y == 0 || x == 0 || z == M - 1 || y == M - 1 || x == M - 1) return;
  if (pass == 0) {
    W[i] = 1.0;
    return;
  }
  const T *a = vals + rowptr[i];
  double a0 = 0.0;
#pragma unroll
  for (int k = 0; k < 27; k++) {
    int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
    if ((oz ? dz : 0) == 0 && (oy ? dy : 0) == 0 && (ox ? dx : 0) == 0) a0 += (double)a[k];
  }
  double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (a0 != 0.0) {
    for (int k = 0; k < 27; k++) {
      int kz = oz ? k / 9 - 1 : 0, ky = oy ? (k / 3) % 3 - 1 : 0, kx = ox ? k % 3 - 1 : 0;
      if (kz == 0 && ky == 0 && kx == 0) continue;
      double s = -(double)a[k] / a0;
      int nz = z + kz, ny = y + ky, nx = x + kx;
      long j = ((long)nz * M + ny) * M + nx;
      int sz = (nz >> 1) - (z >> 1), sy = (ny >> 1) - (y >> 1), sx = (nx >> 1) - (x >> 1);
#pragma unroll
      for (int b = 0; b < 8; b++) {
        int tz = sz + (b >> 2), ty = sy + ((b >> 1) & 1), tx = sx + (b & 1);
        if (tz > 1 || ty > 1 || tx > 1) continue;  // an even axis of j: no weight
        double w = W[(long)b * N + j];
        int tb = (tz << 2) | (ty << 1) | tx;
#pragma unroll
        for (int t = 0; t < 8; t++) acc[t] += (t == tb) ? s * w : 0.0;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 8; t++) W[(long)t * N + i] = acc[t];
}

// P's values from W: entry (i, c) takes the weight of i on coarse node c, 0 for coarse
// boundary nodes and for nodes outside the coarse cell around i
template <class T>
__global__ void k_bbox_write(int M, int Mc, const int *__restrict__ prowptr,
                             const int *__restrict__ pcolidx, const double *__restrict__ W,
                             T *pvals) {
  long N = (long)M * M * M;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int x = (int)(i % M), y = (int)((i / M) % M), z = (int)(i / ((long)M * M));
  for (int e = prowptr[i]; e < prowptr[i + 1]; e++) {
    int c = pcolidx[e];
    int cx = c % Mc, cy = (c / Mc) % Mc, cz = c / (Mc * Mc);
    int bz = cz - (z >> 1), by = cy - (y >> 1), bx = cx - (x >> 1);
    bool in = bz >= 0 && bz <= 1 && by >= 0 && by <= 1 && bx >= 0 && bx <= 1;
    bool cb = cz == 0 || cy == 0 || cx == 0 || cz == Mc - 1 || cy == Mc - 1 || cx == Mc - 1;
    pvals[e] = (in && !cb) ? (T)W[(long)(bz * 4 + by * 2 + bx) * N + i] : (T)0;
  }
}

void lattice_blackbox(const flgpu_mat *A, flgpu_mat *P) {
  FLGPU_REQUIRE(is_q1_lattice(A) && A->lat_n % 2 == 0,
                "black-box interpolation needs a 3D Q1 lattice matrix with an even n");
  int n = A->lat_n, M = n + 1, Mc = n / 2 + 1;
  long N = A->nrows;
  FLGPU_REQUIRE(A->nnz == ipow(3L * n + 1, 3), "the matrix does not hold the lattice pattern");
  FLGPU_REQUIRE(P->nrows == A->nrows && P->ncols == (long)Mc * Mc * Mc,
                "the interpolation does not map the coarse lattice to the matrix's lattice");
  flgpu_ctx *ctx = A->ctx;
  double *W = (double *)dev_alloc(sizeof(double) * 8 * (size_t)N);
  try {
    FLGPU_HIP(hipMemsetAsync(W, 0, sizeof(double) * 8 * (size_t)N, ctx->stream));
    for (int pass = 0; pass <= 3; pass++) {
      if (A->prec == FLGPU_F64)
        k_bbox_pass<double><<<grid_for(N, 256), 256, 0, ctx->stream>>>(
            M, pass, A->rowptr, (const double *)A->vals, W);
      else
        k_bbox_pass<float><<<grid_for(N, 256), 256, 0, ctx->stream>>>(
            M, pass, A->rowptr, (const float *)A->vals, W);
      FLGPU_HIP(hipGetLastError());
    }
    if (P->prec == FLGPU_F64)
      k_bbox_write<double><<<grid_for(N, 256), 256, 0, ctx->stream>>>(
          M, Mc, P->rowptr, P->colidx, W, (double *)P->vals);
    else
      k_bbox_write<float><<<grid_for(N, 256), 256, 0, ctx->stream>>>(
          M, Mc, P->rowptr, P->colidx, W, (float *)P->vals);
    FLGPU_HIP(hipGetLastError());
    FLGPU_HIP(hipStreamSynchronize(ctx->stream));
  } catch (...) {
    dev_free(W);
    throw;
  }
  dev_free(W);
  mat_refresh_values(P);
}

// lattice colouring: colour = per-axis (coordinate mod (order+1)) digits
void lattice_colouring(flgpu_mat *A, int dim, int order, int n) {
  int M = order * n + 1, m = order + 1;
  int ncol = (int)ipow(m, dim);
  std::vector<int> colour(A->nrows);
  for (long r = 0; r < A->nrows; r++) {
    long q = r;
    int cc = 0, mul = 1;
    for (int k = dim - 1; k >= 0; k--) {
      cc += (int)((q % M) % m) * mul;
      mul *= m;
      q /= M;
    }
    colour[r] = cc;
  }
  std::vector<std::vector<int>> rows(ncol);
  for (long r = 0; r < A->nrows; r++) rows[colour[r]].push_back((int)r);
  // rows of one lattice colour never share a cell, so the sweep over a
  // colour is order independent; no coupling check needed
  for (auto &s : A->colour_sell) sell_free(s);
  A->colour_sell.clear();
  for (int c = 0; c < ncol; c++) A->colour_sell.push_back(sell_build_rows(A, rows[c]));
  A->ncolours = ncol;
}

// FAS projection, nc -> 2 nc cells per axis: coarse row c has one entry 1 at the fine DOF of
// the same point, which doubles every lattice coordinate (the injection of the dof
// functionals, fetransfer.lisp:201-213)
template <class T>
__global__ void k_projection_fill(long Nc, int dim, int Mc, int Mf, int *rowptr, int *colidx,
                                  T *vals) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > Nc) return;
  rowptr[c] = (int)c;
  if (c == Nc) return;
  long q = c, f = 0, stride = 1;
  for (int k = 0; k < dim; k++) {
    f += 2 * (q % Mc) * stride;
    q /= Mc;
    stride *= Mf;
  }
  colidx[c] = (int)f;
  vals[c] = (T)1;
}

flgpu_mat *lattice_projection(flgpu_ctx *ctx, int dim, int order, int nc, int prec) {
  check_lattice_args(dim, order, 2 * nc);
  const int Mc = order * nc + 1, Mf = order * 2 * nc + 1;
  const long Nc = ipow(Mc, dim), Nf = ipow(Mf, dim);
  flgpu_mat *R = mat_new(ctx, (int)Nc, (int)Nf, Nc, prec);
  try {
    if (prec == FLGPU_F64)
      k_projection_fill<double><<<grid_for(Nc + 1, 256), 256, 0, ctx->stream>>>(
          Nc, dim, Mc, Mf, R->rowptr, R->colidx, (double *)R->vals);
    else
      k_projection_fill<float><<<grid_for(Nc + 1, 256), 256, 0, ctx->stream>>>(
          Nc, dim, Mc, Mf, R->rowptr, R->colidx, (float *)R->vals);
    FLGPU_HIP(hipGetLastError());
    mat_finish(R);
    FLGPU_HIP(hipStreamSynchronize(ctx->stream));
  } catch (...) {
    mat_free(R);
    throw;
  }
  return R;
}

flgpu_mat *lattice_interpolation(flgpu_ctx *ctx, int dim, int order, int nc, int prec) {
  check_lattice_args(dim, order, 2 * nc);
  Table1D t = table_interp(order, nc);
  Table1D tt = table_transpose(t);
  long Nf = ipow(t.nr, dim), Nc = ipow(t.nc, dim);
  long nnz = ipow((long)t.col.size(), dim);
  flgpu_mat *P = mat_new(ctx, (int)Nf, (int)Nc, nnz, prec);
  flgpu_mat *PT = nullptr;
  try {
    DevTable d = upload_table(ctx, t);
    k_tp_rowptr<<<grid_for(Nf + 1, 256), 256, 0, ctx->stream>>>(Nf, dim, d, P->rowptr);
    if (prec == FLGPU_F64)
      k_tp_fill<double><<<grid_for(Nf, 128), 128, 0, ctx->stream>>>(Nf, dim, d, P->rowptr,
                                                                    P->colidx, (double *)P->vals, 1);
    else
      k_tp_fill<float><<<grid_for(Nf, 128), 128, 0, ctx->stream>>>(Nf, dim, d, P->rowptr, P->colidx,
                                                                   (float *)P->vals, 1);
    FLGPU_HIP(hipGetLastError());
    // the transpose directly from the transposed 1-D table: rows = coarse
    // DOFs, columns = fine DOFs; the zero rule is symmetric in rows/columns
    PT = mat_new(ctx, (int)Nc, (int)Nf, nnz, prec);
    DevTable dt = upload_table(ctx, tt);
    k_tp_rowptr<<<grid_for(Nc + 1, 256), 256, 0, ctx->stream>>>(Nc, dim, dt, PT->rowptr);
    if (prec == FLGPU_F64)
      k_tp_fill<double><<<grid_for(Nc, 128), 128, 0, ctx->stream>>>(
          Nc, dim, dt, PT->rowptr, PT->colidx, (double *)PT->vals, 1);
    else
      k_tp_fill<float><<<grid_for(Nc, 128), 128, 0, ctx->stream>>>(Nc, dim, dt, PT->rowptr,
                                                                   PT->colidx, (float *)PT->vals, 1);
    FLGPU_HIP(hipGetLastError());
    FLGPU_HIP(hipStreamSynchronize(ctx->stream));
    free_table(d);
    free_table(dt);
    mat_finish(P);
    mat_finish(PT);
    P->transpose = PT;
    if (dim == 3 && order == 1) {  // the trilinear pattern: products with implied columns
      P->tp_kind = 1;
      PT->tp_kind = 2;
      P->tp_nc = PT->tp_nc = nc;
      tp_prepare(P);
      tp_prepare(PT);
    }
  } catch (...) {
    mat_free(P);
    mat_free(PT);
    throw;
  }
  return P;
}

}  // namespace flgpu

using namespace flgpu;

extern "C" int flgpu_lattice_assemble(flgpu_ctx *ctx, int dim, int order, int n,
                                      const flgpu_vec *kappa, int nk, double f, int prec, int mode,
                                      flgpu_mat **A, flgpu_vec **rhs) {
  return guard([&] {
    FLGPU_REQUIRE(ctx && A, "ctx and A must not be NULL");
    FLGPU_REQUIRE(!kappa || kappa->ctx == ctx, "kappa belongs to another context");
    CTX_LOCK(ctx);
    *A = lattice_assemble(ctx, dim, order, n, kappa, nk, f, prec, mode, rhs);
    return 0;
  });
}

extern "C" int flgpu_lattice_reassemble(flgpu_mat *A, const flgpu_vec *kappa, int nk, int mode) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    FLGPU_REQUIRE(A->lat_dim > 0, "matrix was not assembled on a lattice");
    FLGPU_REQUIRE(!kappa || kappa->ctx == A->ctx, "kappa belongs to another context");
    CTX_LOCK(A->ctx);
    lattice_reassemble(A, kappa, nk, mode);
    FLGPU_HIP(hipStreamSynchronize(A->ctx->stream));
    return 0;
  });
}

extern "C" int flgpu_lattice_interpolation(flgpu_ctx *ctx, int dim, int order, int nc, int prec,
                                           flgpu_mat **P) {
  return guard([&] {
    FLGPU_REQUIRE(ctx && P, "ctx and P must not be NULL");
    CTX_LOCK(ctx);
    *P = lattice_interpolation(ctx, dim, order, nc, prec);
    return 0;
  });
}

extern "C" int flgpu_lattice_projection(flgpu_ctx *ctx, int dim, int order, int nc, int prec,
                                        flgpu_mat **R) {
  return guard([&] {
    FLGPU_REQUIRE(ctx && R, "ctx and R must not be NULL");
    CTX_LOCK(ctx);
    *R = lattice_projection(ctx, dim, order, nc, prec);
    return 0;
  });
}

extern "C" int flgpu_lattice_blackbox_interpolation(const flgpu_mat *A, flgpu_mat *P) {
  return guard([&] {
    FLGPU_REQUIRE(A && P, "arguments must not be NULL");
    FLGPU_REQUIRE(A->ctx == P->ctx, "A and P belong to different contexts");
    CTX_LOCK(A->ctx);
    lattice_blackbox(A, P);
    return 0;
  });
}

extern "C" int flgpu_lattice_colouring(flgpu_mat *A, int dim, int order, int n) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    check_lattice_args(dim, order, n);
    FLGPU_REQUIRE(A->nrows == ipow((long)order * n + 1, dim) && A->ncols == A->nrows,
                  "matrix size does not match the lattice");
    FLGPU_REQUIRE(A->lat_dim == 0 ||
                      (A->lat_dim == dim && A->lat_order == order && A->lat_n == n),
                  "the matrix belongs to another lattice");
    CTX_LOCK(A->ctx);
    lattice_colouring(A, dim, order, n);
    // a caller's matrix in lattice numbering: the block smoothers find their blocks from this
    A->lat_dim = dim;
    A->lat_order = order;
    A->lat_n = n;
    return 0;
  });
}

extern "C" int flgpu_lattice_element_matrices(flgpu_ctx *ctx, int dim, int order, int n,
                                              const flgpu_vec *kappa, int nk, int prec, int mode,
                                              long first_cell, long ncells, double *out) {
  return guard([&] {
    FLGPU_REQUIRE(ctx && out, "ctx and out must not be NULL");
    check_lattice_args(dim, order, n);
    check_kappa(kappa, dim, n, nk);
    FLGPU_REQUIRE(prec == FLGPU_F64 || prec == FLGPU_F32, "precision must be F64 or F32");
    long total = ipow(n, dim);
    FLGPU_REQUIRE(first_cell >= 0 && ncells >= 0 && first_cell + ncells <= total,
                  "cell range out of bounds");
    CTX_LOCK(ctx);
    RefElement R = ref_element(dim, order);
    void *E = dev_alloc(prec_bytes(prec) * total * R.nloc2);
    try {
      element_matrices(ctx, R, n, kappa, nk, mode, prec, E);
      long cnt = ncells * R.nloc2;
      if (prec == FLGPU_F64) {
        FLGPU_HIP(hipMemcpy(out, (double *)E + first_cell * R.nloc2, sizeof(double) * cnt,
                            hipMemcpyDeviceToHost));
      } else {
        std::vector<float> tmp(cnt);
        FLGPU_HIP(hipMemcpy(tmp.data(), (float *)E + first_cell * R.nloc2, sizeof(float) * cnt,
                            hipMemcpyDeviceToHost));
        for (long i = 0; i < cnt; i++) out[i] = tmp[i];
      }
    } catch (...) {
      dev_free(E);
      throw;
    }
    dev_free(E);
    return 0;
  });
}
