// Sparse matrices of libflgpu.so: CSR as handed over by femlisp, streamed by
// the kernels in SELL-64 form (one wavefront per 64-row slice, column-major
// entries so that each load instruction of a wavefront reads 64 consecutive
// values and 64 consecutive column indices).
//
// Restates femlisp's <sparse-matrix> operations for scalar (1x1-block) DOFs:
//   gemm-nn!/gemm-tn!      src/matlisp/sparse-matrix.lisp:378-442
//   compute-residual       src/iteration/linit.lisp:73-80
//   <jacobi> iterate       src/iteration/linit.lisp:220-250
//   <sor> iterate          src/iteration/linit.lisp:288-323
// Row sums accumulate like femlisp's template: start from beta*x_i and add
// (alpha*y_j)*A_ij in ascending column order (femlisp's order is hash order).
#include <algorithm>
#include <cstring>
#include <numeric>

#include "flgpu_internal.h"

namespace flgpu {

// ---- SELL construction ----------------------------------------------------------

__global__ void k_slice_width(int nsub, const int *rows, const int *rowptr, int *width) {
  int li = blockIdx.x * blockDim.x + threadIdx.x;
  int len = 0;
  if (li < nsub) {
    int r = rows ? rows[li] : li;
    len = rowptr[r + 1] - rowptr[r];
  }
  for (int off = 32; off > 0; off >>= 1) len = max(len, __shfl_down(len, off, 64));
  if ((threadIdx.x & 63) == 0 && (li - (int)(threadIdx.x & 63)) < nsub)
    width[li >> 6] = len;
}

template <class T>
__global__ void k_sell_fill(int nsub, const int *rows, const int *rowptr, const int *colidx,
                            const T *vals, const long *slice_ptr, int *scol, T *sval,
                            int fill_cols) {
  int li = blockIdx.x * blockDim.x + threadIdx.x;
  int s = li >> 6, lane = li & 63;
  int nslices = (nsub + 63) >> 6;
  if (s >= nslices) return;
  long base = slice_ptr[s];
  int w = (int)((slice_ptr[s + 1] - base) >> 6);
  int r = -1, b = 0, len = 0;
  if (li < nsub) {
    r = rows ? rows[li] : li;
    b = rowptr[r];
    len = rowptr[r + 1] - b;
  }
  int padc = (len > 0) ? colidx[b + len - 1] : 0;
  for (int k = 0; k < w; k++) {
    long p = base + (long)k * 64 + lane;
    if (k < len) {
      if (fill_cols) scol[p] = colidx[b + k];
      sval[p] = vals[b + k];
    } else {
      if (fill_cols) scol[p] = padc;
      sval[p] = (T)0;
    }
  }
}

// The same for all rows (rows == nullptr) and rows of at most kSellLdsW entries: a block takes
// one slice, reads the slice's CSR entries (contiguous) into LDS with coalesced loads, and
// writes the slice column by column.
constexpr int kSellLdsW = 32;

template <class T>
__global__ __launch_bounds__(64) void k_sell_fill_lds(int nrows, const int *__restrict__ rowptr,
                                                      const int *__restrict__ colidx,
                                                      const T *__restrict__ vals,
                                                      const long *__restrict__ slice_ptr,
                                                      int *scol, T *sval, int fill_cols) {
  __shared__ T sv[64 * kSellLdsW];
  __shared__ int sc[64 * kSellLdsW];
  const int s = blockIdx.x, lane = threadIdx.x;
  const int r0 = s * 64, r = r0 + lane;
  const int e0 = rowptr[r0], e1 = rowptr[min(r0 + 64, nrows)];
  for (int e = e0 + lane; e < e1; e += 64) {
    sv[e - e0] = vals[e];
    sc[e - e0] = colidx[e];
  }
  int b = 0, len = 0;
  if (r < nrows) {
    b = rowptr[r] - e0;
    len = rowptr[r + 1] - rowptr[r];
  }
  __syncthreads();
  const long base = slice_ptr[s];
  const int w = (int)((slice_ptr[s + 1] - base) >> 6);
  const int padc = len > 0 ? sc[b + len - 1] : 0;
  for (int k = 0; k < w; k++) {
    const long p = base + (long)k * 64 + lane;
    if (k < len) {
      if (fill_cols) scol[p] = sc[b + k];
      sval[p] = sv[b + k];
    } else {
      if (fill_cols) scol[p] = padc;
      sval[p] = (T)0;
    }
  }
}

void sell_free(Sell &s) {
  dev_free(s.rows);
  dev_free(s.slice_ptr);
  dev_free(s.col);
  dev_free(s.val);
  s = Sell();
}

template <class T>
static void sell_fill(flgpu_ctx *ctx, const flgpu_mat *A, Sell &s, int fill_cols) {
  if (s.nrows == 0) return;
  if (!s.rows && s.maxw <= kSellLdsW) {
    k_sell_fill_lds<T><<<s.nslices, 64, 0, ctx->stream>>>(s.nrows, A->rowptr, A->colidx,
                                                          (const T *)A->vals, s.slice_ptr, s.col,
                                                          (T *)s.val, fill_cols);
    FLGPU_HIP(hipGetLastError());
    return;
  }
  int nthreads = s.nslices * 64;
  k_sell_fill<T><<<grid_for(nthreads, 256), 256, 0, ctx->stream>>>(
      s.nrows, s.rows, A->rowptr, A->colidx, (const T *)A->vals, s.slice_ptr, s.col, (T *)s.val,
      fill_cols);
  FLGPU_HIP(hipGetLastError());
}

// rows_host: row ids of the subset (nullptr: all rows)
static Sell sell_build(flgpu_mat *A, const std::vector<int> *rows_host) {
  flgpu_ctx *ctx = A->ctx;
  Sell s;
  s.nrows = rows_host ? (int)rows_host->size() : A->nrows;
  s.nslices = (s.nrows + 63) / 64;
  if (rows_host && s.nrows > 0) {
    s.rows = (int *)dev_alloc(sizeof(int) * s.nrows);
    FLGPU_HIP(hipMemcpyAsync(s.rows, rows_host->data(), sizeof(int) * s.nrows,
                             hipMemcpyHostToDevice, ctx->stream));
  }
  std::vector<int> w(s.nslices);
  std::vector<long> sp(s.nslices + 1, 0);
  if (s.nslices > 0) {
    int *dw = (int *)dev_alloc(sizeof(int) * s.nslices);
    k_slice_width<<<grid_for((long)s.nslices * 64, 256), 256, 0, ctx->stream>>>(
        s.nrows, s.rows, A->rowptr, dw);
    FLGPU_HIP(hipGetLastError());
    FLGPU_HIP(hipMemcpyAsync(w.data(), dw, sizeof(int) * s.nslices, hipMemcpyDeviceToHost,
                             ctx->stream));
    FLGPU_HIP(hipStreamSynchronize(ctx->stream));
    dev_free(dw);
  }
  for (int i = 0; i < s.nslices; i++) {
    sp[i + 1] = sp[i] + 64L * w[i];
    s.maxw = std::max(s.maxw, w[i]);
  }
  s.nstored = sp[s.nslices];
  s.slice_ptr = (long *)dev_alloc(sizeof(long) * (s.nslices + 1));
  FLGPU_HIP(hipMemcpyAsync(s.slice_ptr, sp.data(), sizeof(long) * (s.nslices + 1),
                           hipMemcpyHostToDevice, ctx->stream));
  s.col = (int *)dev_alloc(sizeof(int) * s.nstored);
  s.val = dev_alloc(prec_bytes(A->prec) * s.nstored);
  if (A->prec == FLGPU_F64) sell_fill<double>(ctx, A, s, 1);
  else sell_fill<float>(ctx, A, s, 1);
  FLGPU_HIP(hipStreamSynchronize(ctx->stream));
  return s;
}

Sell sell_build_rows(flgpu_mat *A, const std::vector<int> &rows) { return sell_build(A, &rows); }

// ---- diagonal -----------------------------------------------------------------------

// pos[r] = the position of A(r, r) in the CSR (the last one if repeated), -1 if absent: 32
// lanes per row read the row's columns side by side (coalesced)
__global__ __launch_bounds__(256) void k_diagpos32(int n, const int *__restrict__ rowptr,
                                                   const int *__restrict__ colidx, int *pos) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(t >> 5), k = (int)(t & 31);
  if (r >= n) return;  // the 32 lanes of a row leave together
  int p = -1;
  for (int e = rowptr[r] + k; e < rowptr[r + 1]; e += 32)
    if (colidx[e] == r) p = e;
  for (int off = 16; off > 0; off >>= 1) p = max(p, __shfl_xor(p, off, 32));
  if (k == 0) pos[r] = p;
}

template <class T>
__global__ void k_diag(int n, const int *pos, const T *vals, T *diag) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int p = pos[r];
  diag[r] = p >= 0 ? vals[p] : (T)0;
}

// T = A^T with T's pattern already built: T_cf = A_fc (row c of T, column f)
template <class T>
__global__ void k_transpose_gather(int nrows_t, const int *__restrict__ trowptr,
                                   const int *__restrict__ tcolidx, T *tvals,
                                   const int *__restrict__ rowptr, const int *__restrict__ colidx,
                                   const T *__restrict__ vals) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrows_t) return;
  for (int p = trowptr[c]; p < trowptr[c + 1]; p++) {
    int f = tcolidx[p];
    T v = (T)0;
    for (int e = rowptr[f]; e < rowptr[f + 1]; e++)
      if (colidx[e] == c) {
        v = vals[e];
        break;
      }
    tvals[p] = v;
  }
}

static void diag_refresh(flgpu_mat *A) {
  int n = std::min(A->nrows, A->ncols);
  if (n == 0) return;
  if (A->prec == FLGPU_F64)
    k_diag<double><<<grid_for(n, 256), 256, 0, A->ctx->stream>>>(n, A->diagpos, (const double *)A->vals,
                                                                 (double *)A->diag);
  else
    k_diag<float><<<grid_for(n, 256), 256, 0, A->ctx->stream>>>(n, A->diagpos, (const float *)A->vals,
                                                                (float *)A->diag);
  FLGPU_HIP(hipGetLastError());
}

flgpu_mat *mat_new(flgpu_ctx *ctx, int nrows, int ncols, long nnz, int prec) {
  FLGPU_REQUIRE(nrows >= 0 && ncols >= 0 && nnz >= 0, "negative matrix dimensions");
  FLGPU_REQUIRE(nnz <= 0x7fffffffL, "nnz exceeds the int32 index range");
  FLGPU_REQUIRE(prec == FLGPU_F64 || prec == FLGPU_F32, "precision must be FLGPU_F64 or FLGPU_F32");
  flgpu_mat *A = new flgpu_mat;
  A->ctx = ctx;
  A->nrows = nrows;
  A->ncols = ncols;
  A->nnz = nnz;
  A->prec = prec;
  A->rowptr = (int *)dev_alloc(sizeof(int) * (nrows + 1));
  A->colidx = (int *)dev_alloc(sizeof(int) * nnz);
  A->vals = dev_alloc(prec_bytes(prec) * nnz);
  return A;
}

void mat_finish(flgpu_mat *A) {
  flgpu_ctx *ctx = A->ctx;
  int n = std::min(A->nrows, A->ncols);
  A->diagpos = (int *)dev_alloc(sizeof(int) * std::max(n, 1));
  A->diag = dev_alloc(prec_bytes(A->prec) * std::max(n, 1));
  if (n > 0) {
    k_diagpos32<<<grid_for(32L * n, 256), 256, 0, ctx->stream>>>(n, A->rowptr, A->colidx,
                                                                 A->diagpos);
    FLGPU_HIP(hipGetLastError());
    diag_refresh(A);
  }
  A->sell = sell_build(A, nullptr);
  if (is_q1_lattice(A)) {
    A->val27 = dev_alloc(prec_bytes(A->prec) * 27 * (size_t)A->nrows);
    stencil_refresh(A);
    if (A->sor_strips == 0) sor_strips_alloc(A);
  }
}

void mat_free(flgpu_mat *A) {
  if (!A) return;
  dev_free(A->rowptr);
  dev_free(A->colidx);
  dev_free(A->vals);
  dev_free(A->diagpos);
  dev_free(A->diag);
  sell_free(A->sell);
  for (auto &s : A->colour_sell) sell_free(s);
  if (A->transpose) mat_free(A->transpose);
  dev_free(A->alpha);
  dev_free(A->sor_ghost);
  dev_free(A->val27);
  dev_free(A->tp27);
  dev_free(A->blk_inv);
  dev_free(A->dot_part);
  delete A;
}

// the strip <sor> sweeps' boundary copies, once neither lattice form is left
static void drop_sor_strips(flgpu_mat *A) {
  if (A->alpha || A->val27) return;
  dev_free(A->sor_ghost);
  A->sor_ghost = nullptr;
  A->sor_strips = A->sor_h = 0;
}

// values no longer follow the per-cell form (new values or another ordering)
void drop_alpha(flgpu_mat *A) {
  if (A->alpha) layout_changed(A);
  dev_free(A->alpha);
  A->alpha = nullptr;
  drop_sor_strips(A);
}

void mat_refresh_values(flgpu_mat *A) {
  diag_refresh(A);
  stencil_refresh(A);
  if (A->tp_kind) tp_prepare(A);
  if (A->blk_inv) blocks_prepare(A);
  if (A->prec == FLGPU_F64) {
    sell_fill<double>(A->ctx, A, A->sell, 0);
    for (auto &s : A->colour_sell) sell_fill<double>(A->ctx, A, s, 0);
  } else {
    sell_fill<float>(A->ctx, A, A->sell, 0);
    for (auto &s : A->colour_sell) sell_fill<float>(A->ctx, A, s, 0);
  }
  if (A->transpose) {
    // same pattern: the transpose's values are gathered from the new ones
    flgpu_mat *T = A->transpose;
    if (T->nrows > 0) {
      if (A->prec == FLGPU_F64)
        k_transpose_gather<double><<<grid_for(T->nrows, 256), 256, 0, A->ctx->stream>>>(
            T->nrows, T->rowptr, T->colidx, (double *)T->vals, A->rowptr, A->colidx,
            (const double *)A->vals);
      else
        k_transpose_gather<float><<<grid_for(T->nrows, 256), 256, 0, A->ctx->stream>>>(
            T->nrows, T->rowptr, T->colidx, (float *)T->vals, A->rowptr, A->colidx,
            (const float *)A->vals);
      FLGPU_HIP(hipGetLastError());
    }
    mat_refresh_values(T);
  }
}

static void download_csr(const flgpu_mat *A, std::vector<int> &rp, std::vector<int> &ci,
                         std::vector<double> *v) {
  flgpu_ctx *ctx = A->ctx;
  rp.resize(A->nrows + 1);
  ci.resize(A->nnz);
  FLGPU_HIP(hipMemcpyAsync(rp.data(), A->rowptr, sizeof(int) * (A->nrows + 1),
                           hipMemcpyDeviceToHost, ctx->stream));
  FLGPU_HIP(hipMemcpyAsync(ci.data(), A->colidx, sizeof(int) * A->nnz, hipMemcpyDeviceToHost,
                           ctx->stream));
  if (v) {
    v->resize(A->nnz);
    if (A->prec == FLGPU_F64) {
      FLGPU_HIP(hipMemcpyAsync(v->data(), A->vals, sizeof(double) * A->nnz,
                               hipMemcpyDeviceToHost, ctx->stream));
    } else {
      double *tmp = (double *)dev_alloc(sizeof(double) * A->nnz);
      vcopy(ctx, A->nnz, FLGPU_F32, A->vals, FLGPU_F64, tmp);
      FLGPU_HIP(hipMemcpyAsync(v->data(), tmp, sizeof(double) * A->nnz, hipMemcpyDeviceToHost,
                               ctx->stream));
      FLGPU_HIP(hipStreamSynchronize(ctx->stream));
      dev_free(tmp);
    }
  }
  FLGPU_HIP(hipStreamSynchronize(ctx->stream));
}

static void upload_vals(flgpu_mat *A, const double *vals) {
  flgpu_ctx *ctx = A->ctx;
  if (A->nnz == 0) return;
  if (A->prec == FLGPU_F64) {
    FLGPU_HIP(hipMemcpyAsync(A->vals, vals, sizeof(double) * A->nnz, hipMemcpyHostToDevice,
                             ctx->stream));
    FLGPU_HIP(hipStreamSynchronize(ctx->stream));
  } else {
    double *tmp = (double *)dev_alloc(sizeof(double) * A->nnz);
    FLGPU_HIP(hipMemcpyAsync(tmp, vals, sizeof(double) * A->nnz, hipMemcpyHostToDevice,
                             ctx->stream));
    vcopy(ctx, A->nnz, FLGPU_F64, tmp, FLGPU_F32, A->vals);
    FLGPU_HIP(hipStreamSynchronize(ctx->stream));
    dev_free(tmp);
  }
}

// Transpose via the host (setup only): columns of A become rows, row
// indices ascending within each new row.
void mat_build_transpose(const flgpu_mat *A) {
  if (A->transpose) return;
  std::vector<int> rp, ci, trp(A->ncols + 1, 0), tci(A->nnz);
  std::vector<double> v, tv(A->nnz);
  download_csr(A, rp, ci, &v);
  for (long e = 0; e < A->nnz; e++) trp[ci[e] + 1]++;
  for (int c = 0; c < A->ncols; c++) trp[c + 1] += trp[c];
  std::vector<int> fill(trp.begin(), trp.end() - 1);
  for (int r = 0; r < A->nrows; r++)
    for (int e = rp[r]; e < rp[r + 1]; e++) {
      int p = fill[ci[e]]++;
      tci[p] = r;
      tv[p] = v[e];
    }
  flgpu_mat *T = mat_new(A->ctx, A->ncols, A->nrows, A->nnz, A->prec);
  FLGPU_HIP(hipMemcpyAsync(T->rowptr, trp.data(), sizeof(int) * (A->ncols + 1),
                           hipMemcpyHostToDevice, A->ctx->stream));
  FLGPU_HIP(hipMemcpyAsync(T->colidx, tci.data(), sizeof(int) * A->nnz, hipMemcpyHostToDevice,
                           A->ctx->stream));
  upload_vals(T, tv.data());
  mat_finish(T);
  const_cast<flgpu_mat *>(A)->transpose = T;
  layout_changed(const_cast<flgpu_mat *>(A));
}

flgpu_mat *mat_convert(const flgpu_mat *A, int prec) {
  flgpu_ctx *ctx = A->ctx;
  flgpu_mat *B = mat_new(ctx, A->nrows, A->ncols, A->nnz, prec);
  FLGPU_HIP(hipMemcpyAsync(B->rowptr, A->rowptr, sizeof(int) * (A->nrows + 1),
                           hipMemcpyDeviceToDevice, ctx->stream));
  if (A->nnz > 0) {
    FLGPU_HIP(hipMemcpyAsync(B->colidx, A->colidx, sizeof(int) * A->nnz,
                             hipMemcpyDeviceToDevice, ctx->stream));
    vcopy(ctx, A->nnz, A->prec, A->vals, prec, B->vals);
  }
  B->lat_dim = A->lat_dim;
  B->lat_order = A->lat_order;
  B->lat_n = A->lat_n;
  B->lat_f = A->lat_f;
  B->tp_kind = A->tp_kind;
  B->tp_nc = A->tp_nc;
  if (A->alpha) {
    long nc = (long)A->lat_n * A->lat_n * A->lat_n;
    B->alpha = dev_alloc(prec_bytes(prec) * nc);
    vcopy(ctx, nc, A->prec, A->alpha, prec, B->alpha);
    for (int j = 0; j < 8; j++) B->kt[j] = A->kt[j];
    sor_strips_alloc(B);
  }
  mat_finish(B);
  if (A->transpose) B->transpose = mat_convert(A->transpose, prec);
  if (B->tp_kind) tp_prepare(B);
  if (A->ncolours > 0) {
    // rebuild the colour subsets from A's row lists
    std::vector<int> colour(A->nrows, 0);
    for (int c = 0; c < A->ncolours; c++) {
      const Sell &s = A->colour_sell[c];
      std::vector<int> rows(s.nrows);
      if (s.nrows)
        FLGPU_HIP(hipMemcpy(rows.data(), s.rows, sizeof(int) * s.nrows, hipMemcpyDeviceToHost));
      for (int r : rows) colour[r] = c;
    }
    set_colouring(B, A->ncolours, colour.data());
  }
  return B;
}

void set_colouring(flgpu_mat *A, int ncolours, const int *colour) {
  FLGPU_REQUIRE(A->nrows == A->ncols, "colouring needs a square matrix");
  FLGPU_REQUIRE(ncolours >= 1, "ncolours must be >= 1");
  std::vector<std::vector<int>> rows(ncolours);
  for (int r = 0; r < A->nrows; r++) {
    FLGPU_REQUIRE(colour[r] >= 0 && colour[r] < ncolours, "colour out of range");
    rows[colour[r]].push_back(r);
  }
  // same-colour rows must not couple (otherwise the sweep would not be the
  // sequential <sor> sweep in (colour, index) order)
  std::vector<int> rp, ci;
  download_csr(A, rp, ci, nullptr);
  for (int r = 0; r < A->nrows; r++)
    for (int e = rp[r]; e < rp[r + 1]; e++)
      FLGPU_REQUIRE(ci[e] == r || colour[ci[e]] != colour[r],
                    "colouring couples two rows of one colour");
  for (auto &s : A->colour_sell) sell_free(s);
  A->colour_sell.clear();
  layout_changed(A);
  for (int c = 0; c < ncolours; c++) A->colour_sell.push_back(sell_build(A, &rows[c]));
  A->ncolours = ncolours;
}

// ---- kernels ----------------------------------------------------------------------

template <class T>
__device__ __forceinline__ T ld_stream(const T *p) {
  return __builtin_nontemporal_load(p);
}

// acc + sum_k (alpha x[col_k]) val_k over the w entries of one SELL row (stride 64), in
// order.  Chunks of 8: the 8 values and columns, then the 8 gathers, then the 8
// products, so a chunk costs two memory latencies rather than two per entry.
template <class T>
__device__ __forceinline__ T sell_row(const T *__restrict__ vp, const int *__restrict__ cp, int w,
                                      T alpha, const T *__restrict__ x, T acc) {
  constexpr int C = 8;
  int k = 0;
  for (; k + C <= w; k += C) {
    T a[C], xv[C];
    int c[C];
#pragma unroll
    for (int j = 0; j < C; j++) {
      a[j] = ld_stream(vp + (long)(k + j) * 64);
      c[j] = ld_stream(cp + (long)(k + j) * 64);
    }
#pragma unroll
    for (int j = 0; j < C; j++) xv[j] = x[c[j]];
#pragma unroll
    for (int j = 0; j < C; j++) acc += (alpha * xv[j]) * a[j];
  }
  int rem = w - k;  // uniform over the slice
  if (rem > 0) {
    T a[C], xv[C];
    int c[C];
#pragma unroll
    for (int j = 0; j < C; j++) {
      a[j] = j < rem ? ld_stream(vp + (long)(k + j) * 64) : (T)0;
      c[j] = j < rem ? ld_stream(cp + (long)(k + j) * 64) : 0;
    }
#pragma unroll
    for (int j = 0; j < C; j++) xv[j] = j < rem ? x[c[j]] : (T)0;
#pragma unroll
    for (int j = 0; j < C; j++)
      if (j < rem) acc += (alpha * xv[j]) * a[j];
  }
  return acc;
}

// out[row] = beta*z[row] + sum_k (alpha*x[col]) * val  over one SELL structure
template <class T>
__global__ __launch_bounds__(256) void k_sell_spmv(int nsub, int nslices, const int *__restrict__ rows,
                                                   const long *__restrict__ slice_ptr,
                                                   const int *__restrict__ col,
                                                   const T *__restrict__ val, T alpha,
                                                   const T *__restrict__ x, T beta, const T *z,
                                                   T *out) {
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  int s = (int)lb * 4 + (threadIdx.x >> 6);
  if (s >= nslices) return;
  int lane = threadIdx.x & 63;
  int li = s * 64 + lane;
  long base = slice_ptr[s];
  int w = (int)((slice_ptr[s + 1] - base) >> 6);
  if (li >= nsub) return;
  int row = rows ? rows[li] : li;
  T acc = (beta == (T)0) ? (T)0 : beta * z[row];
  acc = sell_row<T>(val + base + lane, col + base + lane, w, alpha, x, acc);
  out[row] = acc;
}

template <class T>
static void launch_spmv(flgpu_ctx *ctx, const Sell &s, double alpha, const void *x, double beta,
                        const void *z, void *out) {
  if (s.nrows == 0) return;
  unsigned g = xcd_grid(grid_for(s.nslices, 4));
  k_sell_spmv<T><<<g, 256, 0, ctx->stream>>>(s.nrows, s.nslices, s.rows, s.slice_ptr, s.col,
                                             (const T *)s.val, (T)alpha, (const T *)x, (T)beta,
                                             (const T *)z, (T *)out);
  FLGPU_HIP(hipGetLastError());
}

// ---- 3D Q1 lattice with a per-cell coefficient: A applied from alpha ---------------

struct Q1K {
  double k[8];
};

// The neighbours along one axis (one nonzero offset): the Q1 stiffness of a cube couples the
// nodes of a common edge with exactly 0, and kt holds the FP64 Gauss sum's residue, about
// 1e-17 of kt[0]: such an entry is 1e-17 of the row's diagonal, far below the FP32 resolution
// (6e-8), so the FP32 per-cell passes skip those 6 terms; FP64 keeps them, as the assembled
// values do.
template <class T>
__device__ __forceinline__ constexpr bool q1_axis(int dz, int dy, int dx) {
  return sizeof(T) == 4 && (dz != 0) + (dy != 0) + (dx != 0) == 1;
}

// acc + sum over interior neighbours j of row i (ascending j) of (coef x_j) A_ij
// with A_ij = kt[|dz|*4 + |dy|*2 + |dx|] * (sum of alpha over the cells shared
// by i and j); *dii receives A_ii.  Row i = (z,y,x) must be interior.  All
// loads are issued before the first product: a Dirichlet column (eliminated,
// A_ij = 0) reads x_i and adds a zero product, so the sum keeps its order and
// the row waits for one memory latency instead of one per neighbour.
template <class T>
__device__ __forceinline__ T q1_row(int M, int n, long i, int z, int y, int x,
                                    const T *__restrict__ alpha, const Q1K &kt, T acc, T coef,
                                    const T *__restrict__ xv, T *dii) {
  T a[2][2][2];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int s = 0; s < 2; s++)
        a[p][q][s] = alpha[((long)(z - 1 + p) * n + (y - 1 + q)) * n + (x - 1 + s)];
  long M2 = (long)M * M;
  bool lz[3] = {z - 1 != 0, true, z + 1 != M - 1}, ly[3] = {y - 1 != 0, true, y + 1 != M - 1},
       lx[3] = {x - 1 != 0, true, x + 1 != M - 1};
  T xn[27];
#pragma unroll
  for (int k = 0; k < 27; k++) {
    int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
    bool live = lz[dz + 1] && ly[dy + 1] && lx[dx + 1];
    xn[k] = xv[live ? i + dz * M2 + dy * M + dx : i];
  }
#pragma unroll
  for (int k = 0; k < 27; k++) {
    int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
    if (q1_axis<T>(dz, dy, dx)) continue;
    int pz0 = dz > 0 ? 1 : 0, pz1 = dz < 0 ? 0 : 1;
    int py0 = dy > 0 ? 1 : 0, py1 = dy < 0 ? 0 : 1;
    int px0 = dx > 0 ? 1 : 0, px1 = dx < 0 ? 0 : 1;
    T s = (T)0;
#pragma unroll
    for (int p = pz0; p <= pz1; p++)
#pragma unroll
      for (int q = py0; q <= py1; q++)
#pragma unroll
        for (int r = px0; r <= px1; r++) s += a[p][q][r];
    T aij = (T)kt.k[(dz != 0) * 4 + (dy != 0) * 2 + (dx != 0)] * s;
    if (k == 13) *dii = aij;
    bool live = lz[dz + 1] && ly[dy + 1] && lx[dx + 1];
    acc += (coef * xn[k]) * (live ? aij : (T)0);
  }
  return acc;
}

// ---- whole-lattice passes in the per-cell form, blocked along z ---------------------------
// A thread owns QZB consecutive rows along z at one (y, x).  It loads its own x column of the
// 3 lines y-1..y+1 over the QZB + 2 planes and its own cell-factor column of the lines y-1, y
// over QZB + 1 cell planes (18 + 10 loads for 4 rows), and takes the x-1 and x+1 columns from
// the neighbouring lanes of its wave (DPP whole-wave shifts).  Each wave computes 62
// consecutive positions; its two edge lanes only load the columns its first and last position
// need.  Dirichlet values are zeroed as they are loaded, which is what the explicit zeros of
// the eliminated columns contribute.  Each row is formed with exactly q1_row's arithmetic (same
// terms, same order; the cell sums start from their first term, as alpha > 0).
enum { Q1_SPMV = 0, Q1_JACOBI = 1 };

template <class T> struct Q1KT {
  T k[8];
};

// value of lane - 1 (DPP wave_shr:1) and of lane + 1 (DPP wave_shl:1); 0 on the end lanes
__device__ __forceinline__ int lane_up_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int lane_down_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ float lane_up(float v) { return __int_as_float(lane_up_i(__float_as_int(v))); }
__device__ __forceinline__ float lane_down(float v) {
  return __int_as_float(lane_down_i(__float_as_int(v)));
}
__device__ __forceinline__ double lane_up(double v) {
  return __hiloint2double(lane_up_i(__double2hiint(v)), lane_up_i(__double2loint(v)));
}
__device__ __forceinline__ double lane_down(double v) {
  return __hiloint2double(lane_down_i(__double2hiint(v)), lane_down_i(__double2loint(v)));
}

template <class T, class I>
__device__ __forceinline__ T ld_at(const T *__restrict__ p, I byte_offset) {
  return *(const T *)((const char *)p + byte_offset);
}

// OP = Q1_SPMV:   out_i = beta z_i + sum_j (coef x_j) A_ij          (v = z)
// OP = Q1_JACOBI: out_i = x_i + coef (b_i - sum_j A_ij x_j) / A_ii  (v = b, coef = damp;
//                 A_ii from the factors, 1 on Dirichlet rows)
// I is the offset type: unsigned when the byte size of a lattice vector fits 32 bits and the
// position count 31 bits.  QZB rows per thread.
// W2: every result is also stored to out2 in FP64 (the last step of a cycle run in FP32 for
// an FP64 caller).
// DOT (Q1_SPMV): the block's share of sum_i out_i x_i goes to dpart[blockIdx.x] (each lane its
// rows in order, then the wave, then the block's waves in order).
template <class T, int OP, class I, int QZB, bool W2 = false, bool DOT = false>
__global__ __launch_bounds__(256) void k_q1_zblock(int M, int n, const T *__restrict__ alpha,
                                                   Q1KT<T> kt, T coef, T beta,
                                                   const T *__restrict__ v,
                                                   const T *__restrict__ x,
                                                   T *__restrict__ out,
                                                   double *__restrict__ out2 = nullptr,
                                                   double *__restrict__ dpart = nullptr) {
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x & 63;
  const I Mi = (I)M, M2 = Mi * Mi;
  const I total = M2 * (I)((M + QZB - 1) / QZB);
  const I tw = ((I)lb * (blockDim.x / 64) + (threadIdx.x >> 6)) * 62 + (I)lane;  // position + 1
  const bool compute = lane >= 1 && lane <= 62 && tw <= total;  // lanes 0 and 63 are halo
  const I t = tw == 0 ? (I)0 : (tw - 1 < total ? tw - 1 : total - 1);  // every lane loads
  const I line = t / Mi, zb = line / Mi;
  const int gx = (int)(t - line * Mi), gy = (int)(line - zb * Mi), z0 = (int)zb * QZB;
  const bool xbd = gx == 0 || gx == M - 1;
  // byte offsets (the narrow I keeps them in 32 bits: loads then take one VGPR offset from
  // the SGPR base)
  constexpr I B = (I)sizeof(T);
  const I M2B = M2 * B;
  I yo[3];
  bool ybd[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    int y = min(max(gy + d - 1, 0), M - 1);
    yo[d] = (I)(y * M + gx) * B;
    ybd[d] = y == 0 || y == M - 1;
  }
  // plane k (z = z0 - 1 + k, clamped to the lattice) starts at zpl[k]; the stride is uniform,
  // so only plane 0 needs a product (zbase wraps around for z0 = 0, where plane 0 clamps to 0)
  I zpl[QZB + 2];
  const I zbase = (I)z0 * M2B - M2B, ztop = (I)(M - 1) * M2B;
#pragma unroll
  for (int k = 0; k < QZB + 2; k++)
    zpl[k] = k == 0 ? (z0 == 0 ? (I)0 : zbase) : min(zbase + (I)k * M2B, ztop);
  const I nnB = (I)n * (I)n * B;
  I co[2];
#pragma unroll
  for (int q = 0; q < 2; q++) co[q] = (I)(min(max(gy - 1 + q, 0), n - 1) * n + min(gx, n - 1)) * B;
  const I cb = (I)z0 * nnB - nnB, ctop = (I)(n - 1) * nnB;

  // loads, issued in the order the rows consume them (row r needs x planes r..r+2, cell
  // planes r, r+1 and v at row r): the in-order memory counter then lets row r start while
  // the later planes are in flight
  T xv[QZB + 2][3][3], xi[QZB], av[QZB + 1][2][2], vr[QZB];
  auto load_x = [&](int k) {
#pragma unroll
    for (int dy = 0; dy < 3; dy++) xv[k][dy][1] = ld_at(x, zpl[k] + yo[dy]);
  };
  auto load_a = [&](int k) {
    I zo = k == 0 ? (z0 == 0 ? (I)0 : cb) : min(cb + (I)k * nnB, ctop);
#pragma unroll
    for (int q = 0; q < 2; q++) av[k][q][1] = ld_at(alpha, zo + co[q]);
  };
  load_x(0);
  load_x(1);
  load_a(0);
#pragma unroll
  for (int r = 0; r < QZB; r++) {
    load_x(r + 2);
    load_a(r + 1);
    if (OP == Q1_SPMV) vr[r] = (beta == (T)0) ? (T)0 : ld_at(v, zpl[r + 1] + yo[1]);
    else vr[r] = ld_at(v, zpl[r + 1] + yo[1]);
  }

  // Dirichlet values zeroed (the row's own value kept in xi); the x-1 / x+1 columns and the x-1
  // cell column from the neighbouring lanes: the same line and z-block for every interior row
  // (a line ends only at a Dirichlet row, which reads none).  Every lane runs to the end (the
  // shifts read all lanes); only computing lanes store.
  auto prep_x = [&](int k) {
    int z = min(max(z0 - 1 + k, 0), M - 1);
    bool zbd = z == 0 || z == M - 1;
#pragma unroll
    for (int dy = 0; dy < 3; dy++) {
      T val = xv[k][dy][1];
      if (dy == 1 && k >= 1 && k <= QZB) xi[k - 1] = val;
      val = (zbd || ybd[dy] || xbd) ? (T)0 : val;
      xv[k][dy][1] = val;
      xv[k][dy][0] = lane_up(val);
      xv[k][dy][2] = lane_down(val);
    }
  };
  // cs[k][dy][dx]: the ordered sum over cell plane k of the cells an entry with in-plane
  // offset (dy, dx) shares; pl = kt * cs is the entry itself when the offset crosses planes
  // (dz = -1 of row k, dz = +1 of row k - 1).  An entry within a plane (dz = 0) continues
  // row r's plane-r sum with the plane-(r + 1) cells, which is q1_row's order.
  T cs[QZB + 1][3][3], pl[QZB + 1][3][3];
  auto prep_a = [&](int k) {
#pragma unroll
    for (int q = 0; q < 2; q++) av[k][q][0] = lane_up(av[k][q][1]);
#pragma unroll
    for (int e = 0; e < 9; e++) {
      int dy = e / 3 - 1, dx = e % 3 - 1;
      T sum = (T)-0.0;
#pragma unroll
      for (int q = (dy > 0); q <= (dy >= 0); q++)
#pragma unroll
        for (int c = (dx > 0); c <= (dx >= 0); c++) sum += av[k][q][c];
      cs[k][dy + 1][dx + 1] = sum;
      pl[k][dy + 1][dx + 1] = kt.k[4 + (dy != 0) * 2 + (dx != 0)] * sum;
    }
  };
  const bool side = xbd || gy == 0 || gy == M - 1;
  const T cf = (OP == Q1_SPMV) ? coef : (T)-1;
  double dsum = 0.0;
  prep_x(0);
  prep_x(1);
  prep_a(0);
#pragma unroll
  for (int r = 0; r < QZB; r++) {
    prep_x(r + 2);
    prep_a(r + 1);
    const int zz = z0 + r;
    const I i = zpl[r + 1] + yo[1];  // row zz, clamped to the lattice
    const T acc0 = vr[r];
    T acc = acc0, d = (T)1;
#pragma unroll
    for (int k = 0; k < 27; k++) {
      int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
      if (q1_axis<T>(dz, dy, dx)) continue;
      T aij;
      if (dz != 0) {
        aij = pl[r + (dz > 0)][dy + 1][dx + 1];
      } else {
        T sum = cs[r][dy + 1][dx + 1];
#pragma unroll
        for (int q = (dy > 0); q <= (dy >= 0); q++)
#pragma unroll
          for (int c = (dx > 0); c <= (dx >= 0); c++) sum += av[r + 1][q][c];
        aij = kt.k[(dy != 0) * 2 + (dx != 0)] * sum;
      }
      if (k == 13) d = aij;
      acc += (cf * xv[r + 1 + dz][dy + 1][dx + 1]) * aij;
    }
    const bool dir = side || zz == 0 || zz == M - 1;
    T res;
    if (OP == Q1_SPMV) {
      res = dir ? acc0 + (cf * xi[r]) * (T)1 : acc;
    } else {
      T accd = acc0 + (cf * xi[r]) * (T)1;
      res = xi[r] + coef * (dir ? accd / (T)1 : acc / d);
    }
    if (compute && zz < M) {
      *(T *)((char *)out + i) = res;
      if (W2) *(double *)((char *)out2 + i * (I)(sizeof(double) / sizeof(T))) = (double)res;
      if (DOT) dsum += (double)res * (double)xi[r];
    }
  }
  if (DOT) {  // every lane gets here: no lane leaves the kernel early
    __shared__ double ws[4];
    for (int off = 32; off > 0; off >>= 1) dsum += __shfl_down(dsum, off, 64);
    if (lane == 0) ws[threadIdx.x >> 6] = dsum;
    __syncthreads();
    if (threadIdx.x == 0) dpart[blockIdx.x] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
  }
}

// rows per thread of the per-cell passes (measured on 129^3 lattices: FP32 5 rows; FP64 3 for
// the product, 4 for the <jacobi> step)
#ifndef FLGPU_QZB_F32
#define FLGPU_QZB_F32 5
#endif
#ifndef FLGPU_QZB_F64_SPMV
#define FLGPU_QZB_F64_SPMV 3
#endif
#ifndef FLGPU_QZB_F64_JACOBI
#define FLGPU_QZB_F64_JACOBI 4
#endif

template <class T, int OP>
static void launch_q1_pass(const flgpu_mat *A, T coef, T beta, const void *v, const void *x,
                           void *out, double *out2 = nullptr, double *dpart = nullptr) {
  constexpr int QZB = sizeof(T) == 4 ? FLGPU_QZB_F32
                      : (OP == Q1_SPMV ? FLGPU_QZB_F64_SPMV : FLGPU_QZB_F64_JACOBI);
  int M = A->lat_n + 1;
  long total = (long)M * M * ((M + QZB - 1) / QZB);
  long waves = (total + 61) / 62;
  unsigned g = xcd_grid(grid_for(waves * 64, 256));
  Q1KT<T> kt;
  for (int b = 0; b < 8; b++) kt.k[b] = (T)A->kt[b];
  bool narrow = (long)M * M * M * (long)sizeof(T) < (1L << 32) && (long)g * 256 < (1L << 31);
  if constexpr (OP == Q1_JACOBI && sizeof(T) == 4) {
    if (out2) {  // the FP64 byte offsets must fit too
      if (narrow && (long)M * M * M * (long)sizeof(double) < (1L << 32))
        k_q1_zblock<T, OP, unsigned, QZB, true><<<g, 256, 0, A->ctx->stream>>>(
            M, A->lat_n, (const T *)A->alpha, kt, coef, beta, (const T *)v, (const T *)x,
            (T *)out, out2);
      else
        k_q1_zblock<T, OP, unsigned long, QZB, true><<<g, 256, 0, A->ctx->stream>>>(
            M, A->lat_n, (const T *)A->alpha, kt, coef, beta, (const T *)v, (const T *)x,
            (T *)out, out2);
      FLGPU_HIP(hipGetLastError());
      return;
    }
  }
  FLGPU_REQUIRE(out2 == nullptr, "an FP64 copy is written only by the FP32 <jacobi> pass");
  if constexpr (OP == Q1_SPMV) {
    if (dpart) {  // the product and its dot with x; 4 waves per block, as launched below
      if (narrow)
        k_q1_zblock<T, OP, unsigned, QZB, false, true><<<g, 256, 0, A->ctx->stream>>>(
            M, A->lat_n, (const T *)A->alpha, kt, coef, beta, (const T *)v, (const T *)x,
            (T *)out, nullptr, dpart);
      else
        k_q1_zblock<T, OP, unsigned long, QZB, false, true><<<g, 256, 0, A->ctx->stream>>>(
            M, A->lat_n, (const T *)A->alpha, kt, coef, beta, (const T *)v, (const T *)x,
            (T *)out, nullptr, dpart);
      FLGPU_HIP(hipGetLastError());
      return;
    }
  }
  FLGPU_REQUIRE(dpart == nullptr, "the dot is formed only by the per-cell product");
  if (narrow)
    k_q1_zblock<T, OP, unsigned, QZB><<<g, 256, 0, A->ctx->stream>>>(
        M, A->lat_n, (const T *)A->alpha, kt, coef, beta, (const T *)v, (const T *)x, (T *)out);
  else
    k_q1_zblock<T, OP, unsigned long, QZB><<<g, 256, 0, A->ctx->stream>>>(
        M, A->lat_n, (const T *)A->alpha, kt, coef, beta, (const T *)v, (const T *)x, (T *)out);
  FLGPU_HIP(hipGetLastError());
}

// one colour (cz, cy, cx) of a <sor> sweep from alpha
template <class T>
__global__ __launch_bounds__(256) void k_q1_sor(int M, int n, int cz, int cy, int cx,
                                                const T *__restrict__ alpha, Q1K kt, T omega, T *x,
                                                const T *b, T *r) {
  int nx = (M - cx + 1) / 2, ny = (M - cy + 1) / 2, nz = (M - cz + 1) / 2;
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  long t = (long)lb * blockDim.x + threadIdx.x;
  if (t >= (long)nx * ny * nz) return;
  int xx = 2 * (int)(t % nx) + cx, yy = 2 * (int)((t / nx) % ny) + cy,
      zz = 2 * (int)(t / ((long)nx * ny)) + cz;
  long i = ((long)zz * M + yy) * M + xx;
  bool bd = xx == 0 || yy == 0 || zz == 0 || xx == M - 1 || yy == M - 1 || zz == M - 1;
  T corr = b[i];
  T d = (T)1;
  if (bd) corr += (-x[i]) * (T)1;
  else corr = q1_row<T>(M, n, i, zz, yy, xx, alpha, kt, corr, (T)-1, x, &d);
  corr = corr / d;
  r[i] = corr;
  x[i] += omega * corr;
}

// ---- the per-cell <sor> sweep in plane strips ----------------------------------------------
// Colours 0-3 lie on the even z planes, 4-7 on the odd ones, and a row couples to the other z
// parity only through the planes z - 1 and z + 1, which no row of its own parity changes.  So
// one launch sweeps the four colours of one parity.  A block owns the lines [y0, y1) (whole
// lines in x) of one interior plane z.  It stages the plane's x over the lines
// [y0 - 4, y1 + 4), b plus the terms of the planes z - 1 and z + 1 for every row of
// [y0 - 3, y1 + 3), and the cell factors of the cell planes z - 1 and z summed per cell column.
// Then it runs the four in-plane colours with a barrier between them.  Colour j of the four
// needs the in-plane values of the colours before it one line away, so the block computes it
// on 3 - j lines beyond its own on each side.  It does so from the same values with the same
// arithmetic as the block owning those lines, so the values it uses equal the owner's bit for
// bit.  The owner overwrites those lines in the same launch, so a block reads them from a copy
// that k_q1_sor_prep takes before the sweep (gx, gb: for every plane z and strip boundary
// sb = 1..S-1 the lines sb H - 4 .. sb H + 3).  k_q1_sor_prep also updates the Dirichlet
// planes z = 0 and z = M - 1, whose rows couple to nothing.  The row sums b, then the
// neighbours of planes z - 1 and z + 1, then the in-plane ones, and divide by A_ii.  That is
// q1_row's arithmetic (linit.lisp:309-318) with the in-plane cell sums taken over cell columns.
// A row updated here gets r_i = corr and x_i += omega corr, as in k_q1_sor.  b may be r: every
// row's b is read before its r is written, by the owner or from the copy.
template <class T>
__global__ __launch_bounds__(256) void k_q1_sor_prep(int M, int H, int S, FastDiv fM,
                                                     FastDiv fS1, T omega, T *x, const T *b,
                                                     T *r, T *gx, T *gb) {
  const unsigned uM = (unsigned)M, M2 = uM * uM;
  const unsigned nb = (unsigned)(M - 2) * (unsigned)(S - 1) * 8u * uM;  // boundary lines
  const unsigned nt = nb + 2u * M2;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    if (t < nb) {
      unsigned u = fdiv(t, fM);
      const unsigned xc = t - u * uM, k = u & 7u;
      u >>= 3;  // (z - 1) (S - 1) + sb - 1
      const unsigned zq = S > 2 ? fdiv(u, fS1) : u;
      const int sb = (int)(u - zq * (unsigned)(S - 1)) + 1, z = (int)zq + 1;
      const int y = sb * H - 4 + (int)k;
      if (y < 0 || y >= M) continue;
      // copy index ((z (S - 1) + sb - 1) 8 + k) M + xc, u = (z - 1) (S - 1) + sb - 1
      const unsigned g = ((u + (unsigned)(S - 1)) * 8u + k) * uM + xc;
      const unsigned i = ((unsigned)z * uM + (unsigned)y) * uM + xc;
      gx[g] = x[i];
      gb[g] = b[i];
    } else {
      const unsigned e = t - nb;
      const unsigned i = (e < M2 ? 0u : (uM - 1u) * M2) + (e < M2 ? e : e - M2);
      T corr = b[i];
      corr += (-x[i]) * (T)1;
      corr = corr / (T)1;
      r[i] = corr;
      x[i] += omega * corr;
    }
  }
}

constexpr int kSorBatch = 6;  // staged elements per thread and load batch

// Index arithmetic is 32-bit (the lattice has fewer than 2^31 nodes) and divides by the
// lattice widths through FastDiv (fM: M, fn: n, fx0 / fx1: the colour columns (M + 1) / 2 and
// M / 2, 0 when that width is 1): a runtime integer division costs tens of VALU instructions,
// and the kernel is bound by VALU issue.
template <class T, int CAP>
__global__ __launch_bounds__(1024) void k_q1_sor_strips(int M, int n, int pz, int H, int S,
                                                        int backwards, FastDiv fM, FastDiv fn,
                                                        FastDiv fx0, FastDiv fx1,
                                                        const T *__restrict__ alpha, Q1K kt,
                                                        T omega, T *__restrict__ x, const T *b,
                                                        T *r, const T *__restrict__ gx,
                                                        const T *__restrict__ gb) {
  __shared__ T lds[CAP];
  T *xs = lds;                   // x of plane z, line y0 - 4 + l
  T *ps = xs + (H + 8) * M;      // b plus the off-plane terms, line y0 - 3 + l
  T *bs = ps + (H + 6) * M;      // alpha(z - 1) + alpha(z) per cell column, cell line y0 - 4 + l
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int s = (int)(lb % S);
  const int z = (pz == 0 ? 2 : 1) + 2 * (int)(lb / S);
  if (z > M - 2) return;  // the grid's padding (whole blocks)
  const int y0 = s * H, y1 = min(M, y0 + H);
  const unsigned uM = (unsigned)M, un = (unsigned)n, M2 = uM * uM;
  const unsigned zb = (unsigned)z * M2;                    // plane z
  const unsigned cb0 = (unsigned)(z - 1) * un * un, cb1 = (unsigned)z * un * un;  // cell planes
  // the copies of boundary s (lines y0 - 4 ..) and s + 1 (lines y1 - 4 ..); only read where
  // that boundary exists
  const unsigned ol = (unsigned)(z * (S - 1) + s - 1) * 8u * uM, ou = ol + 8u * uM;
  // Staging is a chain of global loads; every thread issues its loads of x and of the cell
  // columns for kSorBatch elements before it stores any, and the off-plane terms of two rows
  // at a time, so that a block waits for a few memory latencies instead of one per element.
  const int nxs = (H + 8) * M, nbs = (H + 7) * n;
  for (int e0 = threadIdx.x; e0 < max(nxs, nbs); e0 += kSorBatch * blockDim.x) {
    T xv[kSorBatch], bv[kSorBatch];
#pragma unroll
    for (int u = 0; u < kSorBatch; u++) {
      const int e = e0 + u * blockDim.x;
      xv[u] = (T)0;
      bv[u] = (T)0;
      if (e < nxs) {
        const int l = (int)fdiv((unsigned)e, fM), xc = e - l * M, y = y0 - 4 + l;
        if (y >= 0 && y < M) {
          if (y < y0) xv[u] = gx[ol + (unsigned)((y - y0 + 4) * M + xc)];
          else if (y >= y1) xv[u] = gx[ou + (unsigned)((y - y1 + 4) * M + xc)];
          else xv[u] = x[zb + (unsigned)(y * M + xc)];
        }
      }
      if (e < nbs) {
        const int l = (int)fdiv((unsigned)e, fn), xc = e - l * n, yc = y0 - 4 + l;
        if (yc >= 0 && yc < n) {
          const unsigned o = (unsigned)(yc * n + xc);
          bv[u] = alpha[cb0 + o] + alpha[cb1 + o];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kSorBatch; u++) {
      const int e = e0 + u * blockDim.x;
      if (e < nxs) xs[e] = xv[u];
      if (e < nbs) bs[e] = bv[u];
    }
  }
  // The off-plane terms: a wave takes 62 consecutive positions of the lines [y0 - 3, y1 + 3)
  // (lanes 1..62; lanes 0 and 63 only load), each lane loading the x column and the cell
  // column of its own position and taking the x - 1 / x + 1 columns from its neighbouring
  // lanes (DPP shifts): 11 loads per row instead of 27.  A row's x - 1 and x + 1 are on its
  // own line (a line ends only at a Dirichlet row, which reads none).  Every lane of a wave
  // runs every iteration (the shifts read all lanes); loads are clamped into the lattice.
  const bool lz[2] = {z - 1 != 0, z + 1 != M - 1};
  const int nps = (H + 6) * M;
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int c0 = (threadIdx.x >> 6) * 62; c0 < nps; c0 += nw * 62) {
    const int e = c0 + lane - 1;
    const int ec = min(max(e, 0), nps - 1);
    const int l = (int)fdiv((unsigned)ec, fM), xc = ec - l * M, y = y0 - 3 + l;
    const int yl = min(max(y, 0), M - 1), xcc = min(xc, n - 1);
    T xcol[2][3], acol[2][2];
#pragma unroll
    for (int p = 0; p < 2; p++) {
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const int yy = min(max(yl + d - 1, 0), M - 1);
        xcol[p][d] = x[(p ? zb + M2 : zb - M2) + (unsigned)(yy * M + xc)];
      }
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const int yc = min(max(yl - 1 + q, 0), n - 1);
        acol[p][q] = alpha[(p ? cb1 : cb0) + (unsigned)(yc * n + xcc)];
      }
    }
    const bool row = lane >= 1 && lane <= 62 && e < nps && y >= 0 && y < M;
    T acc = (T)0;
    if (row) {
      const unsigned i = zb + (unsigned)(y * M + xc);
      if (y < y0) acc = gb[ol + (unsigned)((y - y0 + 4) * M + xc)];
      else if (y >= y1) acc = gb[ou + (unsigned)((y - y1 + 4) * M + xc)];
      else acc = b[i];
    }
    T xn[2][3][3], a[2][2][2];
#pragma unroll
    for (int p = 0; p < 2; p++) {
#pragma unroll
      for (int d = 0; d < 3; d++) {
        xn[p][d][1] = xcol[p][d];
        xn[p][d][0] = lane_up(xcol[p][d]);
        xn[p][d][2] = lane_down(xcol[p][d]);
      }
#pragma unroll
      for (int q = 0; q < 2; q++) {
        a[p][q][1] = acol[p][q];
        a[p][q][0] = lane_up(acol[p][q]);
      }
    }
    if (row && y != 0 && y != M - 1 && xc != 0 && xc != M - 1) {
      const bool ly[3] = {y - 1 != 0, true, y + 1 != M - 1};
      const bool lx[3] = {xc - 1 != 0, true, xc + 1 != M - 1};
#pragma unroll
      for (int p = 0; p < 2; p++)
#pragma unroll
        for (int k = 0; k < 9; k++) {
          const int dy = k / 3 - 1, dx = k % 3 - 1;
          const int q0 = dy > 0 ? 1 : 0, q1 = dy < 0 ? 0 : 1;
          const int c0c = dx > 0 ? 1 : 0, c1c = dx < 0 ? 0 : 1;
          T sum = (T)0;
#pragma unroll
          for (int q = q0; q <= q1; q++)
#pragma unroll
            for (int c = c0c; c <= c1c; c++) sum += a[p][q][c];
          if (q1_axis<T>(1, dy, dx)) continue;
          const T aij = (T)kt.k[4 + (dy != 0) * 2 + (dx != 0)] * sum;
          const bool live = lz[p] && ly[dy + 1] && lx[dx + 1];
          acc += ((T)-1 * xn[p][dy + 1][dx + 1]) * (live ? aij : (T)0);
        }
    }
    if (row) ps[e] = acc;
  }
  __syncthreads();
  for (int j = 0; j < 4; j++) {
    const int c = backwards ? 3 - j : j;
    const int cy = c >> 1, cx = c & 1;
    int lo = max(0, y0 - (3 - j));
    const int hi = min(M, y1 + (3 - j));
    lo += (lo - cy) & 1;  // the first line of parity cy
    const int nyl = hi > lo ? (hi - lo + 1) / 2 : 0, nxl = (M - cx + 1) / 2;
    const FastDiv fx = cx ? fx1 : fx0;
    for (int e = threadIdx.x; e < nyl * nxl; e += blockDim.x) {
      const int ey = nxl > 1 ? (int)fdiv((unsigned)e, fx) : e;
      const int y = lo + 2 * ey, xc = cx + 2 * (e - ey * nxl);
      const int li = (y - y0 + 4) * M + xc;
      const T xi = xs[li];
      T acc = ps[(y - y0 + 3) * M + xc];
      T d = (T)1;
      if (y == 0 || y == M - 1 || xc == 0 || xc == M - 1) {
        acc += (-xi) * (T)1;
      } else {
        const T *bc = bs + (y - y0 + 3) * n + (xc - 1);  // cell column (y - 1, xc - 1)
        const T be[2][2] = {{bc[0], bc[1]}, {bc[n], bc[n + 1]}};
        const bool ly[3] = {y - 1 != 0, true, y + 1 != M - 1};
        const bool lx[3] = {xc - 1 != 0, true, xc + 1 != M - 1};
#pragma unroll
        for (int k = 0; k < 9; k++) {
          const int dy = k / 3 - 1, dx = k % 3 - 1;
          const int q0 = dy > 0 ? 1 : 0, q1 = dy < 0 ? 0 : 1;
          const int c0 = dx > 0 ? 1 : 0, c1 = dx < 0 ? 0 : 1;
          T sum = (T)0;
#pragma unroll
          for (int q = q0; q <= q1; q++)
#pragma unroll
            for (int cc = c0; cc <= c1; cc++) sum += be[q][cc];
          if (q1_axis<T>(0, dy, dx)) continue;
          const T aij = (T)kt.k[(dy != 0) * 2 + (dx != 0)] * sum;
          if (k == 4) d = aij;
          const bool live = ly[dy + 1] && lx[dx + 1];
          acc += ((T)-1 * xs[live ? li + dy * M + dx : li]) * (live ? aij : (T)0);
        }
      }
      const T corr = acc / d;
      const T xnew = xi + omega * corr;
      xs[li] = xnew;
      if (y >= y0 && y < y1) {
        const unsigned i = zb + (unsigned)(y * M + xc);
        r[i] = corr;
        x[i] = xnew;
      }
    }
    __syncthreads();
  }
}

// ---- 3D Q1 lattice matrices in 27-point stencil storage ------------------------------

// val27[k][i] = A(i, i + off_k) for the CSR entries of row i, 0 elsewhere
template <class T>
__global__ void k_st27_fill(int M, const int *__restrict__ rowptr, const int *__restrict__ colidx,
                            const T *__restrict__ vals, T *val27) {
  long N = (long)M * M * M;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int k = 0; k < 27; k++) val27[k * N + i] = (T)0;
  int x = (int)(i % M), y = (int)((i / M) % M), z = (int)(i / ((long)M * M));
  for (int e = rowptr[i]; e < rowptr[i + 1]; e++) {
    long j = colidx[e];
    int jx = (int)(j % M), jy = (int)((j / M) % M), jz = (int)(j / ((long)M * M));
    int ax = jx - x + 1, ay = jy - y + 1, az = jz - z + 1;
    // lattice patterns only hold neighbours; anything else is never stored
    if ((unsigned)ax > 2u || (unsigned)ay > 2u || (unsigned)az > 2u) continue;
    val27[(long)(az * 9 + ay * 3 + ax) * N + i] = vals[e];
  }
}

// The same for rows of at most 27 entries on lattices of M >= 3: a block takes 128 rows, reads
// their CSR entries (contiguous) into LDS with coalesced loads, and every thread walks its row
// once along the 27 neighbour columns in ascending order (columns that are no lattice neighbour
// are passed over), writing every plane with coalesced stores.
constexpr int kSt27Rows = 128;

template <class T>
__global__ __launch_bounds__(kSt27Rows) void k_st27_fill_lds(int M,
                                                             const int *__restrict__ rowptr,
                                                             const int *__restrict__ colidx,
                                                             const T *__restrict__ vals,
                                                             T *val27) {
  __shared__ T sv[kSt27Rows * 27];
  __shared__ int sc[kSt27Rows * 27];
  const long N = (long)M * M * M;
  const long r0 = (long)blockIdx.x * kSt27Rows, i = r0 + threadIdx.x;
  const int e0 = rowptr[r0], e1 = rowptr[min(r0 + kSt27Rows, N)];
  for (int e = e0 + (int)threadIdx.x; e < e1; e += kSt27Rows) {
    sv[e - e0] = vals[e];
    sc[e - e0] = colidx[e];
  }
  __syncthreads();
  if (i >= N) return;
  const long M2 = (long)M * M;
  const int x = (int)(i % M), y = (int)((i / M) % M), z = (int)(i / M2);
  int e = rowptr[i] - e0;
  const int eend = rowptr[i + 1] - e0;
#pragma unroll
  for (int k = 0; k < 27; k++) {
    const int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
    T v = (T)0;
    if ((unsigned)(x + dx) < (unsigned)M && (unsigned)(y + dy) < (unsigned)M &&
        (unsigned)(z + dz) < (unsigned)M) {
      const long j = i + dz * M2 + dy * M + dx;
      while (e < eend && sc[e] < j) e++;
      if (e < eend && sc[e] == j) v = sv[e++];
    }
    val27[k * N + i] = v;
  }
}

// acc + sum over the lattice neighbours j of row i (ascending j) of
// (coef x_j) A_ij; the same entries in the same order as the CSR row.  The 27
// values and neighbour x are loaded first (a neighbour outside the lattice has
// a stored 0 and reads x_i), then summed in order: one memory latency per row.
// coefficient k of row i, whose neighbour i + off is in the lattice when live.  SYM: the values
// are exactly symmetric (A(i, i + d) == A(i + d, i), flgpu_mat::st_sym), so the lower half
// (k < 13) is read from the upper plane 26 - k of the neighbour row: 14 planes are streamed
// instead of 27, and the values read are the same
template <class T, bool SYM>
__device__ __forceinline__ T st27_coef(const T *__restrict__ val27, long N, int k, long i, long off,
                                       bool live) {
  if (SYM && k < 13) return live ? val27[(26 - k) * N + i + off] : (T)0;
  return val27[k * N + i];
}

template <class T, bool SYM = false>
__device__ __forceinline__ T st27_row(int M, long N, long i, int z, int y, int x,
                                      const T *__restrict__ val27, T acc, T coef,
                                      const T *__restrict__ xv, T *dii) {
  long M2 = (long)M * M;
  bool lz[3] = {z > 0, true, z < M - 1}, ly[3] = {y > 0, true, y < M - 1},
       lx[3] = {x > 0, true, x < M - 1};
  T av[27], xn[27];
#pragma unroll
  for (int k = 0; k < 27; k++) {
    int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
    bool live = lz[dz + 1] && ly[dy + 1] && lx[dx + 1];
    long off = dz * M2 + dy * M + dx;
    av[k] = st27_coef<T, SYM>(val27, N, k, i, off, live);
    xn[k] = xv[live ? i + off : i];
  }
  *dii = av[13];
#pragma unroll
  for (int k = 0; k < 27; k++) acc += (coef * xn[k]) * av[k];
  return acc;
}

template <class T, bool SYM>
__global__ __launch_bounds__(256) void k_st27_spmv(int M, const T *__restrict__ val27, T coef,
                                                   const T *__restrict__ x, T beta, const T *z,
                                                   T *out) {
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  long i = (long)lb * blockDim.x + threadIdx.x;
  long N = (long)M * M * M;
  if (i >= N) return;
  int xx = (int)(i % M), yy = (int)((i / M) % M), zz = (int)(i / ((long)M * M));
  T acc = (beta == (T)0) ? (T)0 : beta * z[i];
  T d;
  out[i] = st27_row<T, SYM>(M, N, i, zz, yy, xx, val27, acc, coef, x, &d);
}

// <sor> update of the t-th row of colour (cz, cy, cx): corr = (b_i - sum_j A_ij x_j) / A_ii,
// r_i = corr, x_i += omega corr
template <class T, bool SYM>
__device__ __forceinline__ void st27_sor_row(int M, int cz, int cy, int cx, int nx, int ny,
                                             long t, const T *__restrict__ val27, T omega, T *x,
                                             const T *b, T *r) {
  int xx = 2 * (int)(t % nx) + cx, yy = 2 * (int)((t / nx) % ny) + cy,
      zz = 2 * (int)(t / ((long)nx * ny)) + cz;
  long N = (long)M * M * M;
  long i = ((long)zz * M + yy) * M + xx;
  T d = (T)1;
  T corr = st27_row<T, SYM>(M, N, i, zz, yy, xx, val27, b[i], (T)-1, x, &d);
  corr = corr / d;
  r[i] = corr;
  x[i] += omega * corr;
}

template <class T, bool SYM>
__global__ __launch_bounds__(256) void k_st27_sor(int M, int cz, int cy, int cx,
                                                  const T *__restrict__ val27, T omega, T *x,
                                                  const T *b, T *r) {
  int nx = (M - cx + 1) / 2, ny = (M - cy + 1) / 2, nz = (M - cz + 1) / 2;
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  long t = (long)lb * blockDim.x + threadIdx.x;
  if (t >= (long)nx * ny * nz) return;
  st27_sor_row<T, SYM>(M, cz, cy, cx, nx, ny, t, val27, omega, x, b, r);
}

// ---- the stencil-form <sor> sweep in plane strips --------------------------------------------
// The decomposition of k_q1_sor_strips for matrices in 27-point stencil storage: one launch per
// z parity sweeps the four in-plane colours of its planes (all of them, the boundary planes
// included, since a stencil row is swept like any other), a block owns the whole lines
// [y0, y1) of one plane and recomputes colour j on 3 - j lines beyond them, and the lines of a
// neighbouring strip are read from the copies k_st27_sor_prep takes before the sweep.  Each row
// is st27_row's sum in its order: b, then the terms k = 0..8 (plane z - 1), 9..17 (plane z),
// 18..26 (plane z + 1), then the division by the diagonal entry.  The planes z - 1 and z + 1
// have the other parity and do not change during the launch, so the first nine terms are
// summed for every staged row before the colour phases, and the last nine are loaded from
// device memory in the row's colour phase.  A term whose neighbour lies outside the lattice
// takes the row's own x, as st27_row does, which is the staged value: it still holds x_i
// before the row's update.
template <class T>
__global__ __launch_bounds__(256) void k_st27_sor_prep(int M, int H, int S, FastDiv fM,
                                                       FastDiv fS1, const T *x, const T *b,
                                                       T *gx, T *gb) {
  const unsigned uM = (unsigned)M;
  const unsigned nb = uM * (unsigned)(S - 1) * 8u * uM;  // boundary lines of every plane
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < nb; t += gridDim.x * blockDim.x) {
    unsigned u = fdiv(t, fM);
    const unsigned xc = t - u * uM, k = u & 7u;
    u >>= 3;  // z (S - 1) + sb - 1
    const unsigned z = S > 2 ? fdiv(u, fS1) : u;
    const int sb = (int)(u - z * (unsigned)(S - 1)) + 1;
    const int y = sb * H - 4 + (int)k;
    if (y < 0 || y >= M) continue;
    const unsigned g = (u * 8u + k) * uM + xc;  // the layout k_st27_sor_strips reads
    const unsigned i = (z * uM + (unsigned)y) * uM + xc;
    gx[g] = x[i];
    gb[g] = b[i];
  }
}

// rows per thread and load batch of the plane z - 1 terms
template <class T> constexpr int sor_batch_p() { return sizeof(T) == 8 ? 2 : 4; }

// st27_coef with 32-bit indices (27 N < 2^32) and every load issued (the neighbour's index
// replaced by the row's own when the neighbour is outside the lattice), the value selected
// afterwards
template <class T, bool SYM>
__device__ __forceinline__ T st27_coef_sel(const T *__restrict__ val27, unsigned N, int k,
                                           unsigned i, int off, bool live) {
  if (SYM && k < 13) {
    const T v = val27[(unsigned)(26 - k) * N + (live ? i + (unsigned)off : i)];
    return live ? v : (T)0;
  }
  return val27[(unsigned)k * N + i];
}

template <class T, bool SYM, int CAP>
__global__ __launch_bounds__(1024) void k_st27_sor_strips(int M, int pz, int H, int S,
                                                          int backwards, FastDiv fM, FastDiv fx0,
                                                          FastDiv fx1,
                                                          const T *__restrict__ val27, T omega,
                                                          T *__restrict__ x, const T *b, T *r,
                                                          const T *__restrict__ gx,
                                                          const T *__restrict__ gb) {
  __shared__ T lds[CAP];
  T *xs = lds;               // x of plane z, line y0 - 4 + l
  T *ps = xs + (H + 8) * M;  // b plus the plane z - 1 terms, line y0 - 3 + l
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int s = (int)(lb % S);
  const int z = pz + 2 * (int)(lb / S);
  if (z > M - 1) return;  // the grid's padding (whole blocks)
  const int y0 = s * H, y1 = min(M, y0 + H);
  const int M2 = M * M;
  const unsigned N = (unsigned)M2 * (unsigned)M, zb = (unsigned)z * (unsigned)M2;
  // the copies of boundary s (lines y0 - 4 ..) and s + 1 (lines y1 - 4 ..); only read where
  // that boundary exists
  const unsigned ol = (unsigned)(z * (S - 1) + s - 1) * 8u * (unsigned)M, ou = ol + 8u * M;
  const int nxs = (H + 8) * M;
  for (int e0 = threadIdx.x; e0 < nxs; e0 += kSorBatch * blockDim.x) {
    T xv[kSorBatch];
#pragma unroll
    for (int u = 0; u < kSorBatch; u++) {
      const int e = e0 + u * blockDim.x;
      xv[u] = (T)0;
      if (e < nxs) {
        const int l = (int)fdiv((unsigned)e, fM), xc = e - l * M, y = y0 - 4 + l;
        if (y >= 0 && y < M) {
          if (y < y0) xv[u] = gx[ol + (unsigned)((y - y0 + 4) * M + xc)];
          else if (y >= y1) xv[u] = gx[ou + (unsigned)((y - y1 + 4) * M + xc)];
          else xv[u] = x[zb + (unsigned)(y * M + xc)];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kSorBatch; u++) {
      const int e = e0 + u * blockDim.x;
      if (e < nxs) xs[e] = xv[u];
    }
  }
  __syncthreads();
  // b plus the terms k = 0..8 for the rows of the lines [y0 - 3, y1 + 3)
  const int nps = (H + 6) * M;
  constexpr int BP = sor_batch_p<T>();
  for (int e0 = threadIdx.x; e0 < nps; e0 += BP * blockDim.x) {
    T acc[BP], av[BP][9], xn[BP][9];
    int yq[BP], xq[BP];
#pragma unroll
    for (int u = 0; u < BP; u++) {
      const int e = e0 + u * blockDim.x;
      const int l = (int)fdiv((unsigned)min(e, nps - 1), fM);
      xq[u] = min(e, nps - 1) - l * M;
      const int y = y0 - 3 + l;
      yq[u] = e < nps && y >= 0 && y < M ? y : -1;  // -1: no row
      const int yc = min(max(y, 0), M - 1), xc = xq[u];
      const unsigned i = zb + (unsigned)(yc * M + xc);
      acc[u] = (T)0;
      if (yq[u] >= 0) {
        if (y < y0) acc[u] = gb[ol + (unsigned)((y - y0 + 4) * M + xc)];
        else if (y >= y1) acc[u] = gb[ou + (unsigned)((y - y1 + 4) * M + xc)];
        else acc[u] = b[i];
      }
      const bool ly[3] = {yc > 0, true, yc < M - 1}, lx[3] = {xc > 0, true, xc < M - 1};
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const int dy = k / 3 - 1, dx = k % 3 - 1;
        const bool live = z > 0 && ly[dy + 1] && lx[dx + 1];
        const int off = -M2 + dy * M + dx;
        av[u][k] = st27_coef_sel<T, SYM>(val27, N, k, i, off, live);
        xn[u][k] = x[live ? i + (unsigned)off : i];  // not used when the neighbour is outside
      }
    }
#pragma unroll
    for (int u = 0; u < BP; u++) {
      const int y = yq[u], xc = xq[u];
      if (y < 0) continue;
      const bool ly[3] = {y > 0, true, y < M - 1}, lx[3] = {xc > 0, true, xc < M - 1};
      const T xi = xs[(y - y0 + 4) * M + xc];
      T a = acc[u];
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const int dy = k / 3 - 1, dx = k % 3 - 1;
        const bool live = z > 0 && ly[dy + 1] && lx[dx + 1];
        a += ((T)-1 * (live ? xn[u][k] : xi)) * av[u][k];
      }
      ps[e0 + u * blockDim.x] = a;
    }
  }
  __syncthreads();
  for (int j = 0; j < 4; j++) {
    const int c = backwards ? 3 - j : j;
    const int cy = c >> 1, cx = c & 1;
    int lo = max(0, y0 - (3 - j));
    const int hi = min(M, y1 + (3 - j));
    lo += (lo - cy) & 1;  // the first line of parity cy
    const int nyl = hi > lo ? (hi - lo + 1) / 2 : 0, nxl = (M - cx + 1) / 2;
    const FastDiv fx = cx ? fx1 : fx0;
    for (int e = threadIdx.x; e < nyl * nxl; e += blockDim.x) {
      const int ey = nxl > 1 ? (int)fdiv((unsigned)e, fx) : e;
      const int y = lo + 2 * ey, xc = cx + 2 * (e - ey * nxl);
      const int li = (y - y0 + 4) * M + xc;
      const unsigned i = zb + (unsigned)(y * M + xc);
      const bool ly[3] = {y > 0, true, y < M - 1}, lx[3] = {xc > 0, true, xc < M - 1};
      T av[18], xp[9];
#pragma unroll
      for (int k = 9; k < 27; k++) {
        const int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
        const bool live = (dz == 0 || z < M - 1) && ly[dy + 1] && lx[dx + 1];
        const int off = dz * M2 + dy * M + dx;
        av[k - 9] = st27_coef_sel<T, SYM>(val27, N, k, i, off, live);
        if (dz > 0) xp[k - 18] = x[live ? i + (unsigned)off : i];  // not used when outside
      }
      const T xi = xs[li];
      T acc = ps[(y - y0 + 3) * M + xc];
#pragma unroll
      for (int k = 9; k < 27; k++) {
        const int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
        const bool live = (dz == 0 || z < M - 1) && ly[dy + 1] && lx[dx + 1];
        const T xk = dz == 0 ? xs[live ? li + dy * M + dx : li] : (live ? xp[k - 18] : xi);
        acc += ((T)-1 * xk) * av[k - 9];
      }
      const T corr = acc / av[13 - 9];
      const T xnew = xi + omega * corr;
      xs[li] = xnew;
      if (y >= y0 && y < y1) {
        r[i] = corr;
        x[i] = xnew;
      }
    }
    __syncthreads();
  }
}

bool is_q1_lattice(const flgpu_mat *A) {
  return A->lat_dim == 3 && A->lat_order == 1 &&
         (long)A->nrows == (long)(A->lat_n + 1) * (A->lat_n + 1) * (A->lat_n + 1) &&
         A->nrows == A->ncols;
}

// asym[0] = 1 unless every stored A(i, i + d), d in the lower half, equals A(i + d, i) and the
// coefficients of neighbours outside the lattice are zero
template <class T>
__global__ void k_st27_symcheck(int M, const T *__restrict__ val27, int *asym) {
  const long N = (long)M * M * M, M2 = (long)M * M;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int x = (int)(i % M), y = (int)((i / M) % M), z = (int)(i / M2);
  const bool lz[3] = {z > 0, true, z < M - 1}, ly[3] = {y > 0, true, y < M - 1},
             lx[3] = {x > 0, true, x < M - 1};
  bool bad = false;
  for (int k = 0; k < 13; k++) {
    int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
    T a = val27[k * N + i];
    if (lz[dz + 1] && ly[dy + 1] && lx[dx + 1])
      bad |= !(a == val27[(26 - k) * N + i + dz * M2 + dy * M + dx]);
    else
      bad |= !(a == (T)0);
  }
  if (bad) asym[0] = 1;
}

void stencil_refresh(flgpu_mat *A) {
  if (!A->val27) return;
  int M = A->lat_n + 1;
  long N = A->nrows;
  int *asym = (int *)dev_alloc(sizeof(int));
  FLGPU_HIP(hipMemsetAsync(asym, 0, sizeof(int), A->ctx->stream));
  const bool lds = M >= 3 && A->sell.maxw <= 27;
  const unsigned gl = (unsigned)((N + kSt27Rows - 1) / kSt27Rows);
  if (A->prec == FLGPU_F64) {
    if (lds)
      k_st27_fill_lds<double><<<gl, kSt27Rows, 0, A->ctx->stream>>>(
          M, A->rowptr, A->colidx, (const double *)A->vals, (double *)A->val27);
    else
      k_st27_fill<double><<<grid_for(N, 256), 256, 0, A->ctx->stream>>>(
          M, A->rowptr, A->colidx, (const double *)A->vals, (double *)A->val27);
    k_st27_symcheck<double><<<grid_for(N, 256), 256, 0, A->ctx->stream>>>(
        M, (const double *)A->val27, asym);
  } else {
    if (lds)
      k_st27_fill_lds<float><<<gl, kSt27Rows, 0, A->ctx->stream>>>(
          M, A->rowptr, A->colidx, (const float *)A->vals, (float *)A->val27);
    else
      k_st27_fill<float><<<grid_for(N, 256), 256, 0, A->ctx->stream>>>(
          M, A->rowptr, A->colidx, (const float *)A->vals, (float *)A->val27);
    k_st27_symcheck<float><<<grid_for(N, 256), 256, 0, A->ctx->stream>>>(
        M, (const float *)A->val27, asym);
  }
  FLGPU_HIP(hipGetLastError());
  int h = 1;
  FLGPU_HIP(hipMemcpyAsync(&h, asym, sizeof(int), hipMemcpyDeviceToHost, A->ctx->stream));
  FLGPU_HIP(hipStreamSynchronize(A->ctx->stream));
  dev_free(asym);
  if (A->st_sym != (h == 0)) layout_changed(A);
  A->st_sym = (h == 0);
}

void drop_stencil(flgpu_mat *A) {
  if (A->val27) layout_changed(A);
  dev_free(A->val27);
  A->val27 = nullptr;
  A->st_sym = false;
  drop_sor_strips(A);
}

static Q1K q1k(const flgpu_mat *A) {
  Q1K k;
  for (int j = 0; j < 8; j++) k.k[j] = A->kt[j];
  return k;
}

// ---- 3D order-1 lattice interpolation with the pattern implied ------------------------------
// 1-D: fine node i (0..2nc) has the coarse neighbours i/2 (even i) or (i-1)/2, (i+1)/2 (odd
// i), c(i) = 1 + (i & 1) of them, and its entries start at o(i) = i + i/2 of the S = 3nc + 1
// entries of the 1-D table.  Coarse node c has the fine neighbours 2c-1 (c > 0), 2c, 2c+1
// (c < nc), ct(c) of them, starting at ot(c) = 3c - 1 (c > 0), 0.  A 3-D row (z, y, x) of the
// tensor product starts at o(z) S^2 + c(z) (o(y) S + c(y) o(x)) and lists its entries in
// ascending column order, z-major.  The products below form every row exactly as the SELL
// product does: acc = beta z_i, then acc += (alpha x_j) a_ij in ascending j.

// start of fine row (fz, fy, fx) in the interpolation's values
__device__ __forceinline__ long tp_row_start(int nc, int fz, int fy, int fx) {
  const long S = 3L * nc + 1;
  return (long)(fz + fz / 2) * S * S +
         (long)(1 + (fz & 1)) * ((long)(fy + fy / 2) * S + (long)(1 + (fy & 1)) * (fx + fx / 2));
}

// (z, y, x) of lattice index i < M^3 < 2^31, M = dM.d
__device__ __forceinline__ void lat_decode(unsigned i, FastDiv dM, int &z, int &y, int &x) {
  const unsigned line = fdiv(i, dM), zz = fdiv(line, dM);
  x = (int)(i - line * dM.d);
  y = (int)(line - zz * dM.d);
  z = (int)zz;
}

// The interpolation's values as 27 planes over the coarse nodes c, for the fine nodes
// 2c + (a, b, e) of c's block, a, b, e in {0, 1} (parity classes in lexicographic order): class
// (a, b, e) owns 2^(a+b+e) planes from tp_class_base, one per coarse corner c + (pz, py, px),
// pz <= a, py <= b, px <= e, in the row's own (ascending column) order.  Classes whose fine
// node lies past the lattice (a = 1 at cz = nc, ...) hold zeros there.
__host__ __device__ constexpr int tp_class_base(int cls) {
  return cls == 0 ? 0 : cls == 1 ? 1 : cls == 2 ? 3 : cls == 3 ? 5 : cls == 4 ? 9
       : cls == 5 ? 11 : cls == 6 ? 15 : 19;
}

template <class T>
__global__ void k_tp8_fill(int nc, const T *__restrict__ vals, T *__restrict__ p27) {
  const int Mc = nc + 1, Mf = 2 * nc + 1;
  const long Nc = (long)Mc * Mc * Mc;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Nc) return;
  const int cx = (int)(c % Mc), cy = (int)((c / Mc) % Mc), cz = (int)(c / ((long)Mc * Mc));
  for (int cls = 0; cls < 8; cls++) {
    const int a = cls >> 2, b = (cls >> 1) & 1, e = cls & 1;
    const int fz = 2 * cz + a, fy = 2 * cy + b, fx = 2 * cx + e;
    const bool in = fz < Mf && fy < Mf && fx < Mf;
    const long rs = in ? tp_row_start(nc, fz, fy, fx) : 0;
    for (int k = 0; k < (1 << (a + b + e)); k++)
      p27[(long)(tp_class_base(cls) + k) * Nc + c] = in ? vals[rs + k] : (T)0;
  }
}

// out = beta z + alpha P x: one thread per coarse node c writes the fine nodes of c's block
// from the 8 coarse values of the cell [c, c + 1]^3 and the 27 planes
template <class T>
__global__ __launch_bounds__(256) void k_tp_prolong(int nc, FastDiv dMc,
                                                    const T *__restrict__ p27, T alpha,
                                                    const T *__restrict__ x, T beta,
                                                    const T *__restrict__ z, T *__restrict__ out) {
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const unsigned Mc = dMc.d, Mf = 2 * nc + 1, Nc = Mc * Mc * Mc;
  const unsigned c = lb * blockDim.x + threadIdx.x;
  if (c >= Nc) return;
  int cz, cy, cx;
  lat_decode(c, dMc, cz, cy, cx);
  const bool uz = cz < nc, uy = cy < nc, ux = cx < nc;  // the upper neighbours exist
  T xc[8], w[27], zf[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int pz = q >> 2, py = (q >> 1) & 1, px = q & 1;
    if ((!pz || uz) && (!py || uy) && (!px || ux))
      xc[q] = x[c + ((unsigned)pz * Mc + (unsigned)py) * Mc + (unsigned)px];
  }
#pragma unroll
  for (int k = 0; k < 27; k++) w[k] = p27[(unsigned long)k * Nc + c];
  const unsigned f0 = ((unsigned)(2 * cz) * Mf + (unsigned)(2 * cy)) * Mf + (unsigned)(2 * cx);
#pragma unroll
  for (int cls = 0; cls < 8; cls++) {
    const int a = cls >> 2, b = (cls >> 1) & 1, e = cls & 1;
    if ((!a || uz) && (!b || uy) && (!e || ux) && beta != (T)0)
      zf[cls] = z[f0 + ((unsigned)a * Mf + (unsigned)b) * Mf + (unsigned)e];
  }
#pragma unroll
  for (int cls = 0; cls < 8; cls++) {
    const int a = cls >> 2, b = (cls >> 1) & 1, e = cls & 1;
    if ((!a || uz) && (!b || uy) && (!e || ux)) {
      T acc = (beta == (T)0) ? (T)0 : beta * zf[cls];
#pragma unroll
      for (int pz = 0; pz <= a; pz++)
#pragma unroll
        for (int py = 0; py <= b; py++)
#pragma unroll
          for (int px = 0; px <= e; px++)
            acc += (alpha * xc[(pz * 2 + py) * 2 + px]) *
                   w[tp_class_base(cls) + (pz * (b + 1) + py) * (e + 1) + px];
      out[f0 + ((unsigned)a * Mf + (unsigned)b) * Mf + (unsigned)e] = acc;
    }
  }
}

// coarse row c of the transpose: its fine neighbours per axis and where offset 0 sits
struct TpRow {
  int tz, ty, tx, lz, ly, lx;
  long off;
};
__device__ __forceinline__ TpRow tp_trow(int nc, int cz, int cy, int cx) {
  const long S = 3L * nc + 1;
  auto cnt = [&](int q) { return 3 - (q == 0) - (q == nc); };
  auto ot = [&](int q) { return q == 0 ? 0 : 3 * q - 1; };
  TpRow t;
  t.tz = cnt(cz);
  t.ty = cnt(cy);
  t.tx = cnt(cx);
  t.lz = cz > 0;
  t.ly = cy > 0;
  t.lx = cx > 0;
  t.off = (long)ot(cz) * S * S + (long)t.tz * ((long)ot(cy) * S + (long)t.ty * ot(cx));
  return t;
}

// tp27[k][c] from the transpose's CSR values
template <class T>
__global__ void k_tp27_fill(int nc, const T *__restrict__ tvals, T *__restrict__ t27) {
  const unsigned Mc = nc + 1;
  const unsigned Nc = Mc * Mc * Mc;
  const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Nc) return;
  const unsigned line = c / Mc;
  const int cx = (int)(c - line * Mc), cy = (int)(line % Mc), cz = (int)(line / Mc);
  const TpRow t = tp_trow(nc, cz, cy, cx);
  for (int e = 0; e < 27; e++) {
    const int iz = e / 9 - 1 + t.lz, iy = (e / 3) % 3 - 1 + t.ly, ix = e % 3 - 1 + t.lx;
    const bool live = iz >= 0 && iz < t.tz && iy >= 0 && iy < t.ty && ix >= 0 && ix < t.tx;
    t27[(long)e * Nc + c] = live ? tvals[t.off + (iz * t.ty + iy) * t.tx + ix] : (T)0;
  }
}

// out = beta z + alpha P^T r (coarse rows; t27 = the transpose's values as 27 planes).  FIRST:
// also the first <jacobi> step from zero on the coarse level, xo = damp (out / diag), as
// k_jacobi_zero forms it
template <class T, bool FIRST = false>
__global__ __launch_bounds__(256) void k_tp_restrict(int nc, FastDiv dMc,
                                                     const T *__restrict__ t27, T alpha,
                                                     const T *__restrict__ r, T beta,
                                                     const T *__restrict__ z,
                                                     T *__restrict__ out,
                                                     const T *__restrict__ diag = nullptr,
                                                     T damp = (T)0, T *__restrict__ xo = nullptr) {
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const unsigned Mf = 2 * nc + 1, Mc = dMc.d;
  const unsigned Nc = Mc * Mc * Mc;
  const unsigned c = lb * blockDim.x + threadIdx.x;
  if (c >= Nc) return;
  int cz, cy, cx;
  lat_decode(c, dMc, cz, cy, cx);
  const bool vz[3] = {cz > 0, true, cz < nc}, vy[3] = {cy > 0, true, cy < nc},
             vx[3] = {cx > 0, true, cx < nc};
  const unsigned f0 = ((unsigned)(2 * cz) * Mf + (unsigned)(2 * cy)) * Mf + (unsigned)(2 * cx);
  T a[27], rv[27];
#pragma unroll
  for (int e = 0; e < 27; e++) {
    const int dz = e / 9 - 1, dy = (e / 3) % 3 - 1, dx = e % 3 - 1;
    const bool live = vz[dz + 1] && vy[dy + 1] && vx[dx + 1];
    a[e] = t27[(long)e * Nc + c];
    if (live) rv[e] = r[f0 + (unsigned)((dz * (int)Mf + dy) * (int)Mf + dx)];
  }
  T acc = (beta == (T)0) ? (T)0 : beta * z[c];
#pragma unroll
  for (int e = 0; e < 27; e++) {
    const int dz = e / 9 - 1, dy = (e / 3) % 3 - 1, dx = e % 3 - 1;
    if (vz[dz + 1] && vy[dy + 1] && vx[dx + 1]) acc += (alpha * rv[e]) * a[e];
  }
  out[c] = acc;
  if (FIRST) xo[c] = damp * (acc / diag[c]);
}

void tp_prepare(flgpu_mat *T) {
  if (T->tp_kind == 0) return;
  const int nc = T->tp_nc;
  const long Nc = (long)(nc + 1) * (nc + 1) * (nc + 1);
  if (!T->tp27) {
    T->tp27 = dev_alloc(27 * Nc * prec_bytes(T->prec));
    layout_changed(T);
  }
  const unsigned g = grid_for(Nc, 256);
  if (T->tp_kind == 1) {
    if (T->prec == FLGPU_F64)
      k_tp8_fill<double><<<g, 256, 0, T->ctx->stream>>>(nc, (const double *)T->vals,
                                                        (double *)T->tp27);
    else
      k_tp8_fill<float><<<g, 256, 0, T->ctx->stream>>>(nc, (const float *)T->vals,
                                                       (float *)T->tp27);
  } else {
    if (T->prec == FLGPU_F64)
      k_tp27_fill<double><<<g, 256, 0, T->ctx->stream>>>(nc, (const double *)T->vals,
                                                         (double *)T->tp27);
    else
      k_tp27_fill<float><<<g, 256, 0, T->ctx->stream>>>(nc, (const float *)T->vals,
                                                        (float *)T->tp27);
  }
  FLGPU_HIP(hipGetLastError());
}

void tp_drop(flgpu_mat *A) {
  if (A->tp_kind == 1 && A->transpose) tp_drop(A->transpose);
  if (A->tp_kind || A->tp27) layout_changed(A);
  dev_free(A->tp27);
  A->tp27 = nullptr;
  A->tp_kind = 0;
}

template <class T>
static void launch_tp(const flgpu_mat *A, T alpha, const void *x, T beta, const void *z,
                      void *out) {
  const int nc = A->tp_nc;
  const long Nc = (long)(nc + 1) * (nc + 1) * (nc + 1);
  unsigned g = xcd_grid(grid_for(Nc, 256));
  if (A->tp_kind == 1)
    k_tp_prolong<T><<<g, 256, 0, A->ctx->stream>>>(nc, fast_div(nc + 1), (const T *)A->tp27,
                                                   alpha, (const T *)x, beta, (const T *)z,
                                                   (T *)out);
  else
    k_tp_restrict<T><<<g, 256, 0, A->ctx->stream>>>(nc, fast_div(nc + 1), (const T *)A->tp27,
                                                    alpha, (const T *)x, beta, (const T *)z,
                                                    (T *)out);
  FLGPU_HIP(hipGetLastError());
}

template <class T>
static void launch_st27_spmv(const flgpu_mat *A, T alpha, const T *x, T beta, const T *z, T *out) {
  int M = A->lat_n + 1;
  unsigned g = xcd_grid(grid_for(A->nrows, 256));
  if (A->st_sym)
    k_st27_spmv<T, true><<<g, 256, 0, A->ctx->stream>>>(M, (const T *)A->val27, alpha, x, beta,
                                                        z, out);
  else
    k_st27_spmv<T, false><<<g, 256, 0, A->ctx->stream>>>(M, (const T *)A->val27, alpha, x, beta,
                                                         z, out);
  FLGPU_HIP(hipGetLastError());
}

bool restrict_first_step(const flgpu_mat *PT, const void *r, void *rhs, const flgpu_mat *Ac,
                         double damp, void *xo) {
  if (PT->tp_kind != 2 || !Ac->diag || Ac->nrows != PT->nrows) return false;
  const int nc = PT->tp_nc;
  const long Nc = (long)(nc + 1) * (nc + 1) * (nc + 1);
  unsigned g = xcd_grid(grid_for(Nc, 256));
  if (PT->prec == FLGPU_F64)
    k_tp_restrict<double, true><<<g, 256, 0, PT->ctx->stream>>>(
        nc, fast_div(nc + 1), (const double *)PT->tp27, 1.0, (const double *)r, 0.0,
        (const double *)rhs, (double *)rhs, (const double *)Ac->diag, damp, (double *)xo);
  else
    k_tp_restrict<float, true><<<g, 256, 0, PT->ctx->stream>>>(
        nc, fast_div(nc + 1), (const float *)PT->tp27, 1.0f, (const float *)r, 0.0f,
        (const float *)rhs, (float *)rhs, (const float *)Ac->diag, (float)damp, (float *)xo);
  FLGPU_HIP(hipGetLastError());
  return true;
}

template <class S, class T>
__global__ void k_stage_first(long n, const S *__restrict__ r, T *__restrict__ b,
                              const T *__restrict__ diag, T damp, T *__restrict__ xo) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T bi = (T)r[i];
  b[i] = bi;
  xo[i] = damp * (bi / diag[i]);
}

void stage_first_step(const flgpu_mat *A, int rprec, const void *r, void *b, double damp,
                      void *xo) {
  long n = A->nrows;
  unsigned g = grid_for(n, 256);
  if (rprec == FLGPU_F64 && A->prec == FLGPU_F32)
    k_stage_first<double, float><<<g, 256, 0, A->ctx->stream>>>(
        n, (const double *)r, (float *)b, (const float *)A->diag, (float)damp, (float *)xo);
  else if (rprec == FLGPU_F32 && A->prec == FLGPU_F64)
    k_stage_first<float, double><<<g, 256, 0, A->ctx->stream>>>(
        n, (const float *)r, (double *)b, (const double *)A->diag, damp, (double *)xo);
  else if (A->prec == FLGPU_F64)
    k_stage_first<double, double><<<g, 256, 0, A->ctx->stream>>>(
        n, (const double *)r, (double *)b, (const double *)A->diag, damp, (double *)xo);
  else
    k_stage_first<float, float><<<g, 256, 0, A->ctx->stream>>>(
        n, (const float *)r, (float *)b, (const float *)A->diag, (float)damp, (float *)xo);
  FLGPU_HIP(hipGetLastError());
}

void spmv(const flgpu_mat *A, double alpha, const void *x, double beta, const void *z,
          void *out) {
  if (A->tp_kind) {
    if (A->prec == FLGPU_F64) launch_tp<double>(A, alpha, x, beta, z, out);
    else launch_tp<float>(A, (float)alpha, x, (float)beta, z, out);
    return;
  }
  if (A->alpha) {
    if (A->prec == FLGPU_F64) launch_q1_pass<double, Q1_SPMV>(A, alpha, beta, z, x, out);
    else launch_q1_pass<float, Q1_SPMV>(A, (float)alpha, (float)beta, z, x, out);
    return;
  }
  if (A->val27) {
    if (A->prec == FLGPU_F64)
      launch_st27_spmv<double>(A, alpha, (const double *)x, beta, (const double *)z,
                               (double *)out);
    else
      launch_st27_spmv<float>(A, (float)alpha, (const float *)x, (float)beta, (const float *)z,
                              (float *)out);
    return;
  }
  if (A->prec == FLGPU_F64) launch_spmv<double>(A->ctx, A->sell, alpha, x, beta, z, out);
  else launch_spmv<float>(A->ctx, A->sell, alpha, x, beta, z, out);
}

// blocks of a per-cell pass over A (launch_q1_pass's grid)
static long q1_blocks(const flgpu_mat *A, int qzb) {
  long M = A->lat_n + 1;
  long total = M * M * ((M + qzb - 1) / qzb);
  return (long)xcd_grid(grid_for((total + 61) / 62 * 64, 256));
}

bool spmv_dot(const flgpu_mat *A, const void *x, void *out, int slot) {
  if (!A->alpha || A->tp_kind) return false;
  long nb = q1_blocks(A, A->prec == FLGPU_F64 ? FLGPU_QZB_F64_SPMV : FLGPU_QZB_F32);
  flgpu_mat *M = const_cast<flgpu_mat *>(A);
  if (M->dot_np < nb) {
    dev_free(M->dot_part);
    M->dot_part = (double *)dev_alloc(sizeof(double) * nb);
    M->dot_np = nb;
  }
  if (A->prec == FLGPU_F64)
    launch_q1_pass<double, Q1_SPMV>(A, 1.0, 0.0, out, x, out, nullptr, M->dot_part);
  else
    launch_q1_pass<float, Q1_SPMV>(A, 1.0f, 0.0f, out, x, out, nullptr, M->dot_part);
  dot_final(A->ctx, M->dot_part, (int)nb, slot);
  return true;
}

void spmv_t(const flgpu_mat *A, double alpha, const void *x, double beta, void *out) {
  mat_build_transpose(A);
  spmv(A->transpose, alpha, x, beta, out, out);
}

template <class T>
__global__ void k_jacobi(int n, T damp, const T *__restrict__ diag, T *x, T *r) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T ri = r[i] / diag[i];
  r[i] = ri;
  x[i] += damp * ri;
}

// One <jacobi> step of a product-iterator (linit.lisp:94-120,220-250) with the residual it
// needs recomputed in the same pass, out of place: r_i = b_i - (A x)_i (the sum of spmv with
// alpha = -1, beta = 1), r_i /= A_ii, xo_i = x_i + damp r_i.  The per-cell form runs
// k_q1_zblock (A_ii from the cell factors); the stencil and SELL forms read A_ii from diag:
template <class T, bool SYM>
__global__ __launch_bounds__(256) void k_st27_jacobi(int M, const T *__restrict__ val27,
                                                     const T *__restrict__ diag, T damp,
                                                     const T *__restrict__ x,
                                                     const T *__restrict__ b, T *__restrict__ xo) {
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  long i = (long)lb * blockDim.x + threadIdx.x;
  long N = (long)M * M * M;
  if (i >= N) return;
  int xx = (int)(i % M), yy = (int)((i / M) % M), zz = (int)(i / ((long)M * M));
  T d;
  T acc = st27_row<T, SYM>(M, N, i, zz, yy, xx, val27, b[i], (T)-1, x, &d);
  xo[i] = x[i] + damp * (acc / diag[i]);
}

template <class T>
__global__ __launch_bounds__(256) void k_sell_jacobi(int nsub, int nslices,
                                                     const long *__restrict__ slice_ptr,
                                                     const int *__restrict__ col,
                                                     const T *__restrict__ val,
                                                     const T *__restrict__ diag, T damp,
                                                     const T *__restrict__ x,
                                                     const T *__restrict__ b, T *__restrict__ xo) {
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  int s = (int)lb * 4 + (threadIdx.x >> 6);
  if (s >= nslices) return;
  int lane = threadIdx.x & 63;
  int row = s * 64 + lane;
  long base = slice_ptr[s];
  int w = (int)((slice_ptr[s + 1] - base) >> 6);
  if (row >= nsub) return;
  T acc = sell_row<T>(val + base + lane, col + base + lane, w, (T)-1, x, b[row]);
  xo[row] = x[row] + damp * (acc / diag[row]);
}

// S <jacobi> steps of a small stencil-form level in one pass.  A block owns a region of R^3
// rows: a tile of TE^3 rows with a halo of S - 1 (R = TE + 2 (S - 1)).  x is staged in LDS over
// the region and one more layer; step s (0-based) updates the rows of the region shrunk by s,
// from the previous step's values in LDS, so the last step has exact values on the tile, which
// alone is stored.  Each row keeps its 27 coefficients in registers across the steps and is
// formed as k_st27_jacobi forms it (st27_row's order; a neighbour outside the lattice reads the
// row's own value against its zero coefficient).  All global loads of a thread are issued
// before its first LDS store.  Halo rows are recomputed by neighbouring blocks; this pays only
// where a level is small enough that a launch costs more than its work.
constexpr int MJ_MAXLOAD = 8;  // x values a thread stages: (R + 2)^3 <= MJ_MAXLOAD * block

template <class T, int S, bool SYM>
__global__ __launch_bounds__(1024) void k_st27_jacobi_multi(int M, int TE, const T *__restrict__ val27,
                                                           const T *__restrict__ diag, T damp,
                                                           const T *__restrict__ x,
                                                           const T *__restrict__ b,
                                                           T *__restrict__ xo) {
  extern __shared__ unsigned char mj_lds[];
  const int R = TE + 2 * (S - 1), L = R + 2, L3 = L * L * L;
  T *xs0 = (T *)mj_lds, *xs1 = xs0 + L3;
  const int nt = (M + TE - 1) / TE;
  const int tb = blockIdx.x;
  const int oz = (tb / (nt * nt)) * TE - (S - 1), oy = ((tb / nt) % nt) * TE - (S - 1),
            ox = (tb % nt) * TE - (S - 1);  // region origin
  const long M2 = (long)M * M, N = M2 * M;
  const int t = threadIdx.x;
  const int rz = t / (R * R), ry = (t / R) % R, rx = t % R;
  const int gz = oz + rz, gy = oy + ry, gx = ox + rx;
  const bool row = t < R * R * R && gz >= 0 && gz < M && gy >= 0 && gy < M && gx >= 0 && gx < M;
  const long i = row ? gz * M2 + (long)gy * M + gx : 0;
  const bool lz[3] = {gz > 0, true, gz < M - 1}, ly[3] = {gy > 0, true, gy < M - 1},
             lx[3] = {gx > 0, true, gx < M - 1};
  T av[27], bi = (T)0, di = (T)1;
  if (row) {
#pragma unroll
    for (int k = 0; k < 27; k++) {
      int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
      av[k] = st27_coef<T, SYM>(val27, N, k, i, dz * M2 + dy * M + dx,
                                lz[dz + 1] && ly[dy + 1] && lx[dx + 1]);
    }
    bi = b[i];
    di = diag[i];
  }
  // x over the region and one more layer (lattice values only; the rest is never read)
  T xl[MJ_MAXLOAD];
#pragma unroll
  for (int j = 0; j < MJ_MAXLOAD; j++) {
    const int k = t + j * blockDim.x;
    const int gz2 = oz - 1 + k / (L * L), gy2 = oy - 1 + (k / L) % L, gx2 = ox - 1 + k % L;
    const bool in = k < L3 && gz2 >= 0 && gz2 < M && gy2 >= 0 && gy2 < M && gx2 >= 0 && gx2 < M;
    xl[j] = in ? x[gz2 * M2 + (long)gy2 * M + gx2] : (T)0;
  }
#pragma unroll
  for (int j = 0; j < MJ_MAXLOAD; j++) {
    const int k = t + j * blockDim.x;
    if (k < L3) xs0[k] = xl[j];
  }
  const int own = ((rz + 1) * L + (ry + 1)) * L + (rx + 1);
  __syncthreads();
#pragma unroll
  for (int st = 0; st < S; st++) {
    const T *cur = (st & 1) ? xs1 : xs0;
    T *nxt = (st & 1) ? xs0 : xs1;
    const bool act = row && rz >= st && rz < R - st && ry >= st && ry < R - st && rx >= st &&
                     rx < R - st;
    if (act) {
      T acc = bi;
#pragma unroll
      for (int k = 0; k < 27; k++) {
        int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
        bool live = lz[dz + 1] && ly[dy + 1] && lx[dx + 1];
        T xn = cur[live ? own + (dz * L + dy) * L + dx : own];
        acc += ((T)-1 * xn) * av[k];
      }
      T xn = cur[own] + damp * (acc / di);
      if (st == S - 1) xo[i] = xn;  // the tile (the region shrunk by S - 1)
      else nxt[own] = xn;
    }
    if (st < S - 1) __syncthreads();
  }
}

// tile edge of a multi-step pass over a lattice of size M: about 200 blocks, a region of at
// most 1024 rows and an x box a block can stage; 0 when S steps do not fit
static int mj_tile(int M, int S) {
  int te = std::max(2, (M + 5) / 6);
  for (; te >= 2; te--) {
    int R = te + 2 * (S - 1), L = R + 2;
    int threads = (R * R * R + 63) / 64 * 64;
    if (threads <= 1024 && L * L * L <= MJ_MAXLOAD * threads) return te;
  }
  return 0;
}

template <class T>
static void launch_jacobi_multi(const flgpu_mat *A, int S, int te, T damp, const T *x,
                                const T *b, T *xo) {
  const int M = A->lat_n + 1;
  const int nt = (M + te - 1) / te, R = te + 2 * (S - 1), L = R + 2;
  const unsigned g = (unsigned)nt * nt * nt;
  const unsigned threads = (R * R * R + 63) / 64 * 64;
  const size_t lds = 2 * sizeof(T) * (size_t)L * L * L;
  hipStream_t st = A->ctx->stream;
  const T *v = (const T *)A->val27, *d = (const T *)A->diag;
  FLGPU_REQUIRE(S >= 2 && S <= 4, "2 to 4 steps per pass");
  if (A->st_sym) {
    if (S == 2) k_st27_jacobi_multi<T, 2, true><<<g, threads, lds, st>>>(M, te, v, d, damp, x, b, xo);
    if (S == 3) k_st27_jacobi_multi<T, 3, true><<<g, threads, lds, st>>>(M, te, v, d, damp, x, b, xo);
    if (S == 4) k_st27_jacobi_multi<T, 4, true><<<g, threads, lds, st>>>(M, te, v, d, damp, x, b, xo);
  } else {
    if (S == 2) k_st27_jacobi_multi<T, 2, false><<<g, threads, lds, st>>>(M, te, v, d, damp, x, b, xo);
    if (S == 3) k_st27_jacobi_multi<T, 3, false><<<g, threads, lds, st>>>(M, te, v, d, damp, x, b, xo);
    if (S == 4) k_st27_jacobi_multi<T, 4, false><<<g, threads, lds, st>>>(M, te, v, d, damp, x, b, xo);
  }
  FLGPU_HIP(hipGetLastError());
}

// stencil-form levels up to this lattice size take several <jacobi> steps per launch
// (measured, FP32: 17^3 lattice, 3 steps 7.7 us against 3 x 3.7 us; 33^3, 3 steps 13.6 us
// against 3 x 3.85 us, where the recomputed halos cost more than the launches)
#ifndef FLGPU_MULTI_MAX_M
#define FLGPU_MULTI_MAX_M 20
#endif

template <class T>
static void jacobi_step(const flgpu_mat *A, T damp, const T *x, const T *b, T *xo) {
  hipStream_t st = A->ctx->stream;
  if (A->alpha) {
    launch_q1_pass<T, Q1_JACOBI>(A, damp, (T)0, b, x, xo);
    return;
  } else if (A->val27) {
    unsigned g = xcd_grid(grid_for(A->nrows, 256));
    if (A->st_sym)
      k_st27_jacobi<T, true><<<g, 256, 0, st>>>(A->lat_n + 1, (const T *)A->val27,
                                                (const T *)A->diag, damp, x, b, xo);
    else
      k_st27_jacobi<T, false><<<g, 256, 0, st>>>(A->lat_n + 1, (const T *)A->val27,
                                                 (const T *)A->diag, damp, x, b, xo);
  } else {
    const Sell &s = A->sell;
    if (s.nrows == 0) return;
    unsigned g = xcd_grid(grid_for(s.nslices, 4));
    k_sell_jacobi<T><<<g, 256, 0, st>>>(s.nrows, s.nslices, s.slice_ptr, s.col, (const T *)s.val,
                                        (const T *)A->diag, damp, x, b, xo);
  }
  FLGPU_HIP(hipGetLastError());
}

// the first step from x = 0: b - A 0 = b exactly, so xo = damp b / A_ii without reading A
template <class T>
__global__ void k_jacobi_zero(long n, const T *__restrict__ diag, T damp, const T *__restrict__ b,
                              T *__restrict__ xo) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xo[i] = damp * (b[i] / diag[i]);
}

// `steps` <jacobi> steps, each from a freshly computed residual, x ping-ponged through tmp
// (n values); the residual vector of the iteration is not touched (left stale).  x_zero:
// x holds no values yet and stands for zero.  done = 1: the first step from zero has been
// taken already and its result is in tmp.
bool jacobi_steps(const flgpu_mat *A, double damp, int steps, void *x, const void *b, void *tmp,
                  bool x_zero, int done, double *x64) {
  if (steps <= 0 || A->nrows == 0) return false;
  FLGPU_REQUIRE(A->nrows == A->ncols, "<jacobi> needs a square matrix");
  FLGPU_REQUIRE(done == 0 || (done == 1 && x_zero), "only the first step from zero comes done");
  // x64: the last step also stores x in FP64 when it is a per-cell FP32 pass that ends in x
  const bool w2 = x64 && A->alpha && A->prec == FLGPU_F32 && steps % 2 == 0 && steps - 1 >= done;
  // small stencil-form levels: several steps per launch (each launch writes the buffer it
  // does not read, so the remaining steps are split to end in x)
  const bool multi = A->val27 && !A->alpha && A->lat_n + 1 <= FLGPU_MULTI_MAX_M;
  void *cur = done ? tmp : x;  // where the iterate is (x stands for zero when x_zero, !done)
  int k = done;
  while (k < steps) {
    void *dst = (cur == x) ? tmp : x;
    const int rem = steps - k;
    if (k == 0 && x_zero) {
      unsigned g = grid_for(A->nrows, 256);
      if (A->prec == FLGPU_F64)
        k_jacobi_zero<double><<<g, 256, 0, A->ctx->stream>>>(
            A->nrows, (const double *)A->diag, damp, (const double *)b, (double *)dst);
      else
        k_jacobi_zero<float><<<g, 256, 0, A->ctx->stream>>>(
            A->nrows, (const float *)A->diag, (float)damp, (const float *)b, (float *)dst);
      FLGPU_HIP(hipGetLastError());
      k++;
    } else if (multi && rem >= 2) {
      // from tmp one launch lands in x; from x two are needed
      int S = (cur == tmp) ? rem : rem - rem / 2;
      S = S < 4 ? S : 4;
      int te = 0;
      while (S >= 2 && (te = mj_tile(A->lat_n + 1, S)) == 0) S--;
      if (S >= 2) {
        if (A->prec == FLGPU_F64)
          launch_jacobi_multi<double>(A, S, te, damp, (const double *)cur, (const double *)b,
                                      (double *)dst);
        else
          launch_jacobi_multi<float>(A, S, te, (float)damp, (const float *)cur,
                                     (const float *)b, (float *)dst);
      } else {
        S = 1;
        if (A->prec == FLGPU_F64)
          jacobi_step<double>(A, damp, (const double *)cur, (const double *)b, (double *)dst);
        else
          jacobi_step<float>(A, (float)damp, (const float *)cur, (const float *)b, (float *)dst);
      }
      k += S;
    } else if (w2 && k == steps - 1) {
      launch_q1_pass<float, Q1_JACOBI>(A, (float)damp, 0.0f, b, cur, dst, x64);
      k++;
    } else {
      if (A->prec == FLGPU_F64)
        jacobi_step<double>(A, damp, (const double *)cur, (const double *)b, (double *)dst);
      else
        jacobi_step<float>(A, (float)damp, (const float *)cur, (const float *)b, (float *)dst);
      k++;
    }
    cur = dst;
  }
  if (cur != x) vcopy(A->ctx, A->nrows, A->prec, cur, A->prec, x);
  return w2 && cur == x;
}

void jacobi_iterate(const flgpu_mat *A, double damp, void *x, void *r) {
  int n = A->nrows;
  if (n == 0) return;
  if (A->prec == FLGPU_F64)
    k_jacobi<double><<<grid_for(n, 256), 256, 0, A->ctx->stream>>>(
        n, damp, (const double *)A->diag, (double *)x, (double *)r);
  else
    k_jacobi<float><<<grid_for(n, 256), 256, 0, A->ctx->stream>>>(
        n, (float)damp, (const float *)A->diag, (float *)x, (float *)r);
  FLGPU_HIP(hipGetLastError());
}

// One colour of a <sor> sweep: rows of one colour do not couple, so the
// sequential sweep over them equals this parallel update.
template <class T>
__global__ __launch_bounds__(256) void k_sor_colour(int nsub, int nslices, const int *__restrict__ rows,
                                                    const long *__restrict__ slice_ptr,
                                                    const int *__restrict__ col,
                                                    const T *__restrict__ val,
                                                    const T *__restrict__ diag, T omega, T *x,
                                                    const T *b, T *r) {
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  int s = (int)lb * 4 + (threadIdx.x >> 6);
  if (s >= nslices) return;
  int lane = threadIdx.x & 63;
  int li = s * 64 + lane;
  long base = slice_ptr[s];
  int w = (int)((slice_ptr[s + 1] - base) >> 6);
  if (li >= nsub) return;
  int row = rows[li];
  T corr = sell_row<T>(val + base + lane, col + base + lane, w, (T)-1, x, b[row]);
  corr = corr / diag[row];
  r[row] = corr;
  x[row] += omega * corr;
}

// LDS of the strip sweep (values): 128 KB in FP64, 64 KB in FP32
constexpr int kSorCapF64 = 16384, kSorCapF32 = 16384;

// values the strip sweep stages for strips of H lines on an M^3 lattice
static long sor_strip_lds(int M, int H) {
  return (long)(H + 8) * M + (long)(H + 6) * M + (long)(H + 7) * (M - 1);
}

void sor_strips_alloc(flgpu_mat *A) {
  dev_free(A->sor_ghost);
  A->sor_ghost = nullptr;
  A->sor_strips = A->sor_h = 0;
  const int M = A->lat_n + 1;
  if (A->lat_dim != 3 || A->lat_order != 1 || M < 3) return;
  const long cap = A->prec == FLGPU_F64 ? kSorCapF64 : kSorCapF32;
  // about one block per CU and launch, strips of at least 8 lines, within the LDS (two blocks
  // per CU in FP32 with strips half as high measured slower: 63 against 48 us at 129^3)
  const int planes = (M - 1) / 2;
  int S = std::max(1, std::min((256 + planes - 1) / planes, (M + 7) / 8));
  while (S <= M && sor_strip_lds(M, (M + S - 1) / S) > cap) S++;
  if (S > M) return;  // lattices this wide keep the colour launches
  const int H = (M + S - 1) / S;
  S = (M + H - 1) / H;  // no empty strip
  A->sor_strips = S;
  A->sor_h = H;
  if (S > 1)
    A->sor_ghost = dev_alloc(prec_bytes(A->prec) * 2 * (size_t)M * (S - 1) * 8 * M);
}

template <class T, int CAP>
static void q1_sor_strips(const flgpu_mat *A, T omega, int backwards, T *x, const T *b, T *r) {
  const int M = A->lat_n + 1, S = A->sor_strips, H = A->sor_h;
  FLGPU_REQUIRE(sor_strip_lds(M, H) <= CAP, "strip sweep: LDS size");
  T *gx = (T *)A->sor_ghost, *gb = gx ? gx + (long)M * (S - 1) * 8 * M : nullptr;
  const long prep = (long)(M - 2) * (S - 1) * 8 * M + 2L * M * M;
  const FastDiv fS1 = S > 2 ? fast_div((unsigned)(S - 1)) : FastDiv{1, 0, 0};
  k_q1_sor_prep<T><<<grid_for(prep, 256), 256, 0, A->ctx->stream>>>(
      M, H, S, fast_div((unsigned)M), fS1, omega, x, b, r, gx, gb);
  FLGPU_HIP(hipGetLastError());
  for (int k = 0; k < 2; k++) {
    const int pz = backwards ? 1 - k : k;
    const int z0 = pz == 0 ? 2 : 1;
    if (z0 > M - 2) continue;
    const int planes = (M - 2 - z0) / 2 + 1;
    const unsigned g = xcd_grid((unsigned)(planes * S));
    const FastDiv fx1 = M / 2 > 1 ? fast_div((unsigned)(M / 2)) : FastDiv{1, 0, 0};
    k_q1_sor_strips<T, CAP><<<g, 1024, 0, A->ctx->stream>>>(
        M, A->lat_n, pz, H, S, backwards, fast_div((unsigned)M), fast_div((unsigned)A->lat_n),
        fast_div((unsigned)((M + 1) / 2)), fx1, (const T *)A->alpha, q1k(A), omega, x, b, r, gx,
        gb);
    FLGPU_HIP(hipGetLastError());
  }
}

// the stencil-form sweep in strips: the copies of the strip boundaries of every plane, then one
// launch per z parity (the same strips as the per-cell sweep, whose LDS need is larger)
template <class T, int CAP>
static void st27_sor_strips(const flgpu_mat *A, T omega, int backwards, T *x, const T *b, T *r) {
  const int M = A->lat_n + 1, S = A->sor_strips, H = A->sor_h;
  FLGPU_REQUIRE((long)(2 * H + 14) * M <= CAP, "strip sweep: LDS size");
  FLGPU_REQUIRE(27L * M * M * M < (1L << 32), "strip sweep: lattice too large");
  T *gx = (T *)A->sor_ghost, *gb = gx ? gx + (long)M * (S - 1) * 8 * M : nullptr;
  const FastDiv fM = fast_div((unsigned)M);
  if (S > 1) {
    const long prep = (long)M * (S - 1) * 8 * M;
    const FastDiv fS1 = S > 2 ? fast_div((unsigned)(S - 1)) : FastDiv{1, 0, 0};
    k_st27_sor_prep<T><<<grid_for(prep, 256), 256, 0, A->ctx->stream>>>(M, H, S, fM, fS1, x, b,
                                                                        gx, gb);
    FLGPU_HIP(hipGetLastError());
  }
  const FastDiv fx0 = fast_div((unsigned)((M + 1) / 2));
  const FastDiv fx1 = M / 2 > 1 ? fast_div((unsigned)(M / 2)) : FastDiv{1, 0, 0};
  for (int k = 0; k < 2; k++) {
    const int pz = backwards ? 1 - k : k;
    const int planes = (M - 1 - pz) / 2 + 1;
    const unsigned g = xcd_grid((unsigned)(planes * S));
    if (A->st_sym)
      k_st27_sor_strips<T, true, CAP><<<g, 1024, 0, A->ctx->stream>>>(
          M, pz, H, S, backwards, fM, fx0, fx1, (const T *)A->val27, omega, x, b, r, gx, gb);
    else
      k_st27_sor_strips<T, false, CAP><<<g, 1024, 0, A->ctx->stream>>>(
          M, pz, H, S, backwards, fM, fx0, fx1, (const T *)A->val27, omega, x, b, r, gx, gb);
    FLGPU_HIP(hipGetLastError());
  }
}

void sor_sweep(const flgpu_mat *A, double omega, int backwards, void *x, const void *b, void *r) {
  if (A->alpha && A->sor_strips > 0) {
    if (A->prec == FLGPU_F64)
      q1_sor_strips<double, kSorCapF64>(A, omega, backwards, (double *)x, (const double *)b,
                                        (double *)r);
    else
      q1_sor_strips<float, kSorCapF32>(A, (float)omega, backwards, (float *)x,
                                       (const float *)b, (float *)r);
    return;
  }
  if (!A->alpha && A->val27 && A->sor_strips > 0 && 27L * A->nrows < (1L << 32)) {
    if (A->prec == FLGPU_F64)
      st27_sor_strips<double, kSorCapF64>(A, omega, backwards, (double *)x, (const double *)b,
                                          (double *)r);
    else
      st27_sor_strips<float, kSorCapF32>(A, (float)omega, backwards, (float *)x,
                                         (const float *)b, (float *)r);
    return;
  }
  if (A->alpha || A->val27) {
    // lattice colours in the order of the lattice colouring: cz*4 + cy*2 + cx
    int M = A->lat_n + 1;
    for (int k = 0; k < 8; k++) {
      int c = backwards ? 7 - k : k;
      int cz = c >> 2, cy = (c >> 1) & 1, cx = c & 1;
      long cnt = (long)((M - cx + 1) / 2) * ((M - cy + 1) / 2) * ((M - cz + 1) / 2);
      unsigned g = xcd_grid(grid_for(cnt, 256));
      if (!A->alpha) {
        if (A->prec == FLGPU_F64 && A->st_sym)
          k_st27_sor<double, true><<<g, 256, 0, A->ctx->stream>>>(
              M, cz, cy, cx, (const double *)A->val27, omega, (double *)x, (const double *)b,
              (double *)r);
        else if (A->prec == FLGPU_F64)
          k_st27_sor<double, false><<<g, 256, 0, A->ctx->stream>>>(
              M, cz, cy, cx, (const double *)A->val27, omega, (double *)x, (const double *)b,
              (double *)r);
        else if (A->st_sym)
          k_st27_sor<float, true><<<g, 256, 0, A->ctx->stream>>>(
              M, cz, cy, cx, (const float *)A->val27, (float)omega, (float *)x, (const float *)b,
              (float *)r);
        else
          k_st27_sor<float, false><<<g, 256, 0, A->ctx->stream>>>(
              M, cz, cy, cx, (const float *)A->val27, (float)omega, (float *)x, (const float *)b,
              (float *)r);
      } else if (A->prec == FLGPU_F64)
        k_q1_sor<double><<<g, 256, 0, A->ctx->stream>>>(M, A->lat_n, cz, cy, cx,
                                                        (const double *)A->alpha, q1k(A), omega,
                                                        (double *)x, (const double *)b,
                                                        (double *)r);
      else
        k_q1_sor<float><<<g, 256, 0, A->ctx->stream>>>(M, A->lat_n, cz, cy, cx,
                                                       (const float *)A->alpha, q1k(A),
                                                       (float)omega, (float *)x, (const float *)b,
                                                       (float *)r);
      FLGPU_HIP(hipGetLastError());
    }
    return;
  }
  FLGPU_REQUIRE(A->ncolours > 0, "<sor> needs a colouring of the matrix");
  for (int k = 0; k < A->ncolours; k++) {
    const Sell &s = A->colour_sell[backwards ? A->ncolours - 1 - k : k];
    if (s.nrows == 0) continue;
    unsigned g = xcd_grid(grid_for(s.nslices, 4));
    if (A->prec == FLGPU_F64)
      k_sor_colour<double><<<g, 256, 0, A->ctx->stream>>>(
          s.nrows, s.nslices, s.rows, s.slice_ptr, s.col, (const double *)s.val,
          (const double *)A->diag, omega, (double *)x, (const double *)b, (double *)r);
    else
      k_sor_colour<float><<<g, 256, 0, A->ctx->stream>>>(
          s.nrows, s.nslices, s.rows, s.slice_ptr, s.col, (const float *)s.val,
          (const float *)A->diag, (float)omega, (float *)x, (const float *)b, (float *)r);
    FLGPU_HIP(hipGetLastError());
  }
}

// `steps` <sor> sweeps (the product-iterator of a smoother, linit.lisp:94-120)
void sor_sweeps(const flgpu_mat *A, double omega, int steps, int backwards, void *x, const void *b,
                void *r) {
  for (int k = 0; k < steps; k++) sor_sweep(A, omega, backwards, x, b, r);
}

// ---- Galerkin product P^T A P (src/iteration/amg.lisp:247-280) ---------------------

// One thread per coarse row k: for fine l in column k of P (row k of P^T,
// ascending), entries A_lm of row l, and P's row m: acc[k][n] += (P_lk A_lm)
// P_mn, accumulated in FP64 into the caller's coarse pattern.  For a fixed
// coarse entry the contributions arrive in (l, m) order, as in femlisp's
// loop nest.  A product entry outside the pattern sets *bad.
template <class T>
__global__ void k_rap(int nc, const int *__restrict__ ptrp, const int *__restrict__ ptci,
                      const T *__restrict__ ptv, const int *__restrict__ arp,
                      const int *__restrict__ aci, const T *__restrict__ av,
                      const int *__restrict__ prp, const int *__restrict__ pci,
                      const T *__restrict__ pv, const int *__restrict__ crp,
                      const int *__restrict__ cci, double *acc, int *bad, int upper) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nc) return;
  int c0 = crp[k], c1 = crp[k + 1];
  for (int e = c0; e < c1; e++) acc[e] = 0.0;
  for (int a = ptrp[k]; a < ptrp[k + 1]; a++) {
    int l = ptci[a];
    double plk = (double)ptv[a];
    for (int b = arp[l]; b < arp[l + 1]; b++) {
      int m = aci[b];
      double w = plk * (double)av[b];
      for (int c = prp[m]; c < prp[m + 1]; c++) {
        int n = pci[c];
        if (upper && n < k) continue;  // the lower triangle is mirrored afterwards
        int lo = c0, hi = c1 - 1;
        while (lo < hi) {
          int mid = (lo + hi) >> 1;
          if (cci[mid] < n) lo = mid + 1;
          else hi = mid;
        }
        if (lo >= c1 || cci[lo] != n) {
          *bad = 1;
          continue;
        }
        acc[lo] += w * (double)pv[c];
      }
    }
  }
}

// The same with the coarse row's columns and accumulators in LDS (coarse rows of at most
// kRapW entries): the column searches and the FP64 sums stay on the CU instead of reaching
// device memory for every contribution.  Same loops, same order, same results.
constexpr int kRapW = 27, kRapB = 128;

template <class T>
__global__ __launch_bounds__(kRapB) void k_rap_lds(
    int nc, const int *__restrict__ ptrp, const int *__restrict__ ptci, const T *__restrict__ ptv,
    const int *__restrict__ arp, const int *__restrict__ aci, const T *__restrict__ av,
    const int *__restrict__ prp, const int *__restrict__ pci, const T *__restrict__ pv,
    const int *__restrict__ crp, const int *__restrict__ cci, double *acc, int *bad, int upper) {
  __shared__ double sacc[kRapW * kRapB];  // entry j of this thread's row at j * kRapB + tid
  __shared__ int scol[kRapW * kRapB];
  const int tid = threadIdx.x;
  const int k = blockIdx.x * kRapB + tid;
  if (k >= nc) return;
  const int c0 = crp[k], w = crp[k + 1] - c0;
  for (int j = 0; j < w; j++) {
    scol[j * kRapB + tid] = cci[c0 + j];
    sacc[j * kRapB + tid] = 0.0;
  }
  for (int a = ptrp[k]; a < ptrp[k + 1]; a++) {
    const int l = ptci[a];
    const double plk = (double)ptv[a];
    for (int b = arp[l]; b < arp[l + 1]; b++) {
      const int m = aci[b];
      const double wv = plk * (double)av[b];
      for (int c = prp[m]; c < prp[m + 1]; c++) {
        const int n = pci[c];
        if (upper && n < k) continue;  // the lower triangle is mirrored afterwards
        int lo = 0, hi = w - 1;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (scol[mid * kRapB + tid] < n) lo = mid + 1;
          else hi = mid;
        }
        if (lo >= w || scol[lo * kRapB + tid] != n) {
          *bad = 1;
          continue;
        }
        sacc[lo * kRapB + tid] += wv * (double)pv[c];
      }
    }
  }
  for (int j = 0; j < w; j++) acc[c0 + j] = sacc[j * kRapB + tid];
}

// The same product on 3D Q1 lattices, one thread per coarse entry (k, n): the fine matrix holds
// the lattice pattern (every lattice neighbour, ascending), P the pattern of the lattice
// interpolation (fine node m takes the coarse nodes floor(m / 2) and ceil(m / 2) per axis,
// ascending) and the coarse matrix the coarse lattice pattern, so every position is computed
// instead of searched, the row starts included.  The entry sums the contributions
// (P_lk A_lm) P_mn of the fine l around 2k (ascending) and the m next to l and to 2n
// (ascending): exactly the contributions, in the order, that k_rap's loops bring to this
// entry, so the sums are equal bit for bit.
__device__ __forceinline__ int lat_pos3(const int lo[3], const int cnt[3], int a, int b, int c) {
  return ((a - lo[0]) * cnt[1] + (b - lo[1])) * cnt[2] + (c - lo[2]);
}

// the first entry of row (z, y, x) of a tensor product of 1-D patterns: pre[a] entries precede
// the row's coordinate on axis a, which has cnt[a] entries; nnz1 per 1-D pattern
__device__ __forceinline__ long tp_row_start(const long pre[3], const int cnt[3], long nnz1) {
  return (pre[0] * nnz1 + cnt[0] * pre[1]) * nnz1 + (long)cnt[0] * cnt[1] * pre[2];
}

template <class T>
__global__ __launch_bounds__(256) void k_rap_q1(int Mc, int M, const T *__restrict__ av,
                                                const T *__restrict__ pv,
                                                const int *__restrict__ crp, double *acc,
                                                int upper) {
  // the 1-D patterns: A's row i holds i - 1, i, i + 1 within the lattice (3 M - 2 entries,
  // 3 i - 1 of them before row i > 0); P's row i holds i / 2 and, for odd i, i / 2 + 1
  // (3 Mc - 2 entries, i + i / 2 before row i)
  const long nA = 3L * M - 2, nP = 3L * Mc - 2;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long Nc = (long)Mc * Mc * Mc;
  if (t >= Nc * 27) return;
  const int k = (int)(t / 27), d = (int)(t - (long)k * 27);
  const int K[3] = {k / (Mc * Mc), (k / Mc) % Mc, k % Mc};
  const int D[3] = {d / 9 - 1, (d / 3) % 3 - 1, d % 3 - 1};
  int Nn[3], clo[3], ccnt[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    Nn[a] = K[a] + D[a];
    if (Nn[a] < 0 || Nn[a] >= Mc) return;
    clo[a] = K[a] > 0 ? -1 : 0;
    ccnt[a] = (K[a] < Mc - 1 ? 1 : 0) - clo[a] + 1;
  }
  const int n = (Nn[0] * Mc + Nn[1]) * Mc + Nn[2];
  if (upper && n < k) return;  // the lower triangle is mirrored afterwards
  double s = 0.0;
  const int l0[3] = {max(2 * K[0] - 1, 0), max(2 * K[1] - 1, 0), max(2 * K[2] - 1, 0)};
  const int l1[3] = {min(2 * K[0] + 1, M - 1), min(2 * K[1] + 1, M - 1), min(2 * K[2] + 1, M - 1)};
  for (int lz = l0[0]; lz <= l1[0]; lz++)
    for (int ly = l0[1]; ly <= l1[1]; ly++)
      for (int lx = l0[2]; lx <= l1[2]; lx++) {
        const int L[3] = {lz, ly, lx};
        // P_lk: row l holds floor(l / 2) (and ceil(l / 2) on odd coordinates) per axis
        int plo[3], pcnt[3], alo[3], acnt[3], m0[3], m1[3];
        long ppre[3], apre[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
          plo[a] = L[a] >> 1;
          pcnt[a] = (L[a] & 1) + 1;
          ppre[a] = L[a] + (L[a] >> 1);
          alo[a] = L[a] > 0 ? L[a] - 1 : 0;
          acnt[a] = min(L[a] + 1, M - 1) - alo[a] + 1;
          apre[a] = L[a] > 0 ? 3L * L[a] - 1 : 0;
          m0[a] = max(max(L[a] - 1, 2 * Nn[a] - 1), 0);
          m1[a] = min(min(L[a] + 1, 2 * Nn[a] + 1), M - 1);
        }
        const double plk =
            (double)pv[tp_row_start(ppre, pcnt, nP) + lat_pos3(plo, pcnt, K[0], K[1], K[2])];
        const long ab = tp_row_start(apre, acnt, nA);
        for (int mz = m0[0]; mz <= m1[0]; mz++)
          for (int my = m0[1]; my <= m1[1]; my++)
            for (int mx = m0[2]; mx <= m1[2]; mx++) {
              const double w = plk * (double)av[ab + lat_pos3(alo, acnt, mz, my, mx)];
              const int qlo[3] = {mz >> 1, my >> 1, mx >> 1};
              const int qcnt[3] = {(mz & 1) + 1, (my & 1) + 1, (mx & 1) + 1};
              const long qpre[3] = {mz + (mz >> 1), my + (my >> 1), mx + (mx >> 1)};
              s += w * (double)pv[tp_row_start(qpre, qcnt, nP) +
                                  lat_pos3(qlo, qcnt, Nn[0], Nn[1], Nn[2])];
            }
      }
  acc[crp[k] + lat_pos3(clo, ccnt, D[0], D[1], D[2])] = s;
}

// the longest row: a grid of at most 1024 blocks strides over the rows, each block reduces its
// maximum in LDS and issues one atomic (atomics to one address are serialised)
constexpr int kMaxRowBlocks = 1024;

__global__ __launch_bounds__(256) void k_max_row(int n, const int *__restrict__ rp, int *mx) {
  __shared__ int wm[4];
  int len = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    len = max(len, rp[i + 1] - rp[i]);
  for (int off = 32; off > 0; off >>= 1) len = max(len, __shfl_xor(len, off, 64));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = len;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int m = max(max(wm[0], wm[1]), max(wm[2], wm[3]));
    if (m > 0) atomicMax(mx, m);
  }
}

// the longest row of A's CSR (A's SELL may not be built yet)
static int max_row_length(const flgpu_mat *A) {
  if (A->nrows == 0) return 0;
  int *dm = (int *)dev_alloc(sizeof(int));
  FLGPU_HIP(hipMemsetAsync(dm, 0, sizeof(int), A->ctx->stream));
  const unsigned g = std::min<unsigned>(grid_for(A->nrows, 256), kMaxRowBlocks);
  k_max_row<<<g, 256, 0, A->ctx->stream>>>(A->nrows, A->rowptr, dm);
  FLGPU_HIP(hipGetLastError());
  int h = 0;
  FLGPU_HIP(hipMemcpyAsync(&h, dm, sizeof(int), hipMemcpyDeviceToHost, A->ctx->stream));
  FLGPU_HIP(hipStreamSynchronize(A->ctx->stream));
  dev_free(dm);
  return h;
}

template <class T>
__global__ void k_store(long n, const double *acc, T *v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (T)acc[i];
}

// asym[0] = 1 unless A's values are exactly symmetric (every A(i, j) has an equal A(j, i))
template <class T>
__global__ void k_csr_symcheck(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                               const T *__restrict__ v, int *asym) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool bad = false;
  for (int e = rp[i]; e < rp[i + 1]; e++) {
    int j = ci[e];
    if (j == i) continue;
    int lo = rp[j], hi = rp[j + 1] - 1;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (ci[mid] < i) lo = mid + 1;
      else hi = mid;
    }
    bad |= lo > hi || ci[lo] != i || !(v[lo] == v[e]);
  }
  if (bad) asym[0] = 1;
}

// The same with one lane per entry (32 lanes per row, rows of at most 32 entries): the binary
// searches of a row run side by side instead of one after another.
template <class T>
__global__ __launch_bounds__(256) void k_csr_symcheck32(int n, const int *__restrict__ rp,
                                                        const int *__restrict__ ci,
                                                        const T *__restrict__ v, int *asym) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(t >> 5), k = (int)(t & 31);
  if (i >= n) return;
  const int e = rp[i] + k;
  if (e >= rp[i + 1]) return;
  const int j = ci[e];
  if (j == i) return;
  int lo = rp[j], hi = rp[j + 1] - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ci[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  if (lo > hi || ci[lo] != i || !(v[lo] == v[e])) asym[0] = 1;
}

// the lower triangle of the coarse accumulators from the upper one (missing[0] = 1 when the
// pattern has no transposed position)
__global__ void k_mirror(int nc, const int *__restrict__ crp, const int *__restrict__ cci,
                         double *acc, int *missing) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nc) return;
  for (int e = crp[k]; e < crp[k + 1]; e++) {
    int m = cci[e];
    if (m >= k) break;
    int lo = crp[m], hi = crp[m + 1] - 1;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (cci[mid] < k) lo = mid + 1;
      else hi = mid;
    }
    if (lo > hi || cci[lo] != k) missing[0] = 1;
    else acc[e] = acc[lo];
  }
}

static bool csr_symmetric(const flgpu_mat *A) {
  if (A->nrows != A->ncols || A->nrows == 0) return false;
  flgpu_ctx *ctx = A->ctx;
  int *asym = (int *)dev_alloc(sizeof(int));
  FLGPU_HIP(hipMemsetAsync(asym, 0, sizeof(int), ctx->stream));
  const bool lanes = max_row_length(A) <= 32;
  const unsigned g32 = grid_for(32L * A->nrows, 256);
  if (A->prec == FLGPU_F64) {
    if (lanes)
      k_csr_symcheck32<double><<<g32, 256, 0, ctx->stream>>>(A->nrows, A->rowptr, A->colidx,
                                                              (const double *)A->vals, asym);
    else
      k_csr_symcheck<double><<<grid_for(A->nrows, 256), 256, 0, ctx->stream>>>(
          A->nrows, A->rowptr, A->colidx, (const double *)A->vals, asym);
  } else {
    if (lanes)
      k_csr_symcheck32<float><<<g32, 256, 0, ctx->stream>>>(A->nrows, A->rowptr, A->colidx,
                                                             (const float *)A->vals, asym);
    else
      k_csr_symcheck<float><<<grid_for(A->nrows, 256), 256, 0, ctx->stream>>>(
          A->nrows, A->rowptr, A->colidx, (const float *)A->vals, asym);
  }
  FLGPU_HIP(hipGetLastError());
  int h = 1;
  FLGPU_HIP(hipMemcpyAsync(&h, asym, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FLGPU_HIP(hipStreamSynchronize(ctx->stream));
  dev_free(asym);
  return h == 0;
}

// Ac's values <- P^T A P (pattern of Ac given); all three share one precision.  When A is
// exactly symmetric, only the upper triangle is accumulated and mirrored, so Ac is exactly
// symmetric too (the two triangles of P^T A P otherwise round differently), which lets the
// stencil kernels stream half the coefficients.
void galerkin_product(const flgpu_mat *P, const flgpu_mat *A, flgpu_mat *Ac, bool lattice) {
  FLGPU_REQUIRE(P->nrows == A->nrows && A->nrows == A->ncols, "galerkin: P and A do not match");
  FLGPU_REQUIRE(Ac->nrows == P->ncols && Ac->ncols == P->ncols, "galerkin: Ac does not match P");
  FLGPU_REQUIRE(P->prec == A->prec && A->prec == Ac->prec, "galerkin: precisions differ");
  flgpu_ctx *ctx = A->ctx;
  double *acc = (double *)dev_alloc(sizeof(double) * std::max<long>(Ac->nnz, 1));
  int *bad = (int *)dev_alloc(sizeof(int));
  FLGPU_HIP(hipMemsetAsync(bad, 0, sizeof(int), ctx->stream));
  int nc = Ac->nrows;
  int *missing = (int *)dev_alloc(sizeof(int));
  // the structured product: 3D Q1, the fine lattice of an even n, the patterns' sizes as the
  // lattices give them
  const long nf = A->lat_n, ncl = nf / 2;
  const bool q1 = lattice && is_q1_lattice(A) && nf % 2 == 0 &&
                  A->nnz == (3 * nf + 1) * (3 * nf + 1) * (3 * nf + 1) &&
                  Ac->nrows == (ncl + 1) * (ncl + 1) * (ncl + 1) &&
                  Ac->nnz == (3 * ncl + 1) * (3 * ncl + 1) * (3 * ncl + 1) &&
                  P->nnz == (3 * ncl + 1) * (3 * ncl + 1) * (3 * ncl + 1);
  if (!q1) mat_build_transpose(P);  // the general product walks the rows of P^T
  const flgpu_mat *PT = P->transpose;
  const bool lds = !q1 && max_row_length(Ac) <= kRapW;
  auto rap = [&](int upper) {
    const unsigned gl = (unsigned)((nc + kRapB - 1) / kRapB);
    if (q1 && A->prec == FLGPU_F64)
      k_rap_q1<double><<<grid_for(27L * nc, 256), 256, 0, ctx->stream>>>(
          (int)ncl + 1, (int)nf + 1, (const double *)A->vals, (const double *)P->vals,
          Ac->rowptr, acc, upper);
    else if (q1)
      k_rap_q1<float><<<grid_for(27L * nc, 256), 256, 0, ctx->stream>>>(
          (int)ncl + 1, (int)nf + 1, (const float *)A->vals, (const float *)P->vals,
          Ac->rowptr, acc, upper);
    else if (lds && A->prec == FLGPU_F64)
      k_rap_lds<double><<<gl, kRapB, 0, ctx->stream>>>(
          nc, PT->rowptr, PT->colidx, (const double *)PT->vals, A->rowptr, A->colidx,
          (const double *)A->vals, P->rowptr, P->colidx, (const double *)P->vals, Ac->rowptr,
          Ac->colidx, acc, bad, upper);
    else if (lds)
      k_rap_lds<float><<<gl, kRapB, 0, ctx->stream>>>(
          nc, PT->rowptr, PT->colidx, (const float *)PT->vals, A->rowptr, A->colidx,
          (const float *)A->vals, P->rowptr, P->colidx, (const float *)P->vals, Ac->rowptr,
          Ac->colidx, acc, bad, upper);
    else if (A->prec == FLGPU_F64)
      k_rap<double><<<grid_for(nc, 128), 128, 0, ctx->stream>>>(
          nc, PT->rowptr, PT->colidx, (const double *)PT->vals, A->rowptr, A->colidx,
          (const double *)A->vals, P->rowptr, P->colidx, (const double *)P->vals, Ac->rowptr,
          Ac->colidx, acc, bad, upper);
    else
      k_rap<float><<<grid_for(nc, 128), 128, 0, ctx->stream>>>(
          nc, PT->rowptr, PT->colidx, (const float *)PT->vals, A->rowptr, A->colidx,
          (const float *)A->vals, P->rowptr, P->colidx, (const float *)P->vals, Ac->rowptr,
          Ac->colidx, acc, bad, upper);
    FLGPU_HIP(hipGetLastError());
  };
  int hmiss = 1;
  // a lattice matrix's stencil form records whether its values are exactly symmetric (the
  // same comparison of A(i, j) with A(j, i) over the same entries, made when it was filled)
  const bool sym = (q1 && A->val27) ? A->st_sym : csr_symmetric(A);
  if (sym) {
    FLGPU_HIP(hipMemsetAsync(missing, 0, sizeof(int), ctx->stream));
    rap(1);
    k_mirror<<<grid_for(nc, 128), 128, 0, ctx->stream>>>(nc, Ac->rowptr, Ac->colidx, acc,
                                                          missing);
    FLGPU_HIP(hipGetLastError());
    FLGPU_HIP(hipMemcpyAsync(&hmiss, missing, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FLGPU_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (hmiss) rap(0);  // not symmetric, or the coarse pattern is not: the full product
  if (A->prec == FLGPU_F64)
    k_store<double><<<grid_for(Ac->nnz, 256), 256, 0, ctx->stream>>>(Ac->nnz, acc,
                                                                     (double *)Ac->vals);
  else
    k_store<float><<<grid_for(Ac->nnz, 256), 256, 0, ctx->stream>>>(Ac->nnz, acc,
                                                                    (float *)Ac->vals);
  FLGPU_HIP(hipGetLastError());
  int hbad = 0;
  FLGPU_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FLGPU_HIP(hipStreamSynchronize(ctx->stream));
  dev_free(acc);
  dev_free(bad);
  dev_free(missing);
  FLGPU_REQUIRE(!hbad, "galerkin: P^T A P has entries outside the coarse pattern");
}

}  // namespace flgpu

using namespace flgpu;

// ---- C-ABI ---------------------------------------------------------------------------

extern "C" int flgpu_galerkin_product(const flgpu_mat *P, const flgpu_mat *A, flgpu_mat *Ac) {
  return guard([&] {
    FLGPU_REQUIRE(P && A && Ac, "arguments must not be NULL");
    FLGPU_REQUIRE(P->ctx == A->ctx && A->ctx == Ac->ctx, "matrices of different contexts");
    CTX_LOCK(A->ctx);
    galerkin_product(P, A, Ac);
    mat_refresh_values(Ac);
    FLGPU_HIP(hipStreamSynchronize(A->ctx->stream));
    return 0;
  });
}

extern "C" int flgpu_csr_upload(flgpu_ctx *ctx, int nrows, int ncols, long nnz, const int *rowptr,
                                const int *colidx, const double *vals, int prec,
                                flgpu_mat **out) {
  return guard([&] {
    FLGPU_REQUIRE(ctx && out && rowptr, "ctx, rowptr and out must not be NULL");
    FLGPU_REQUIRE(nnz == 0 || (colidx && vals), "colidx and vals must not be NULL");
    FLGPU_REQUIRE(nrows >= 0 && ncols >= 0, "negative dimension");
    FLGPU_REQUIRE(rowptr[0] == 0 && rowptr[nrows] == nnz, "rowptr must run from 0 to nnz");
    for (int r = 0; r < nrows; r++) {
      FLGPU_REQUIRE(rowptr[r + 1] >= rowptr[r], "rowptr must be non-decreasing");
      for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
        FLGPU_REQUIRE(colidx[e] >= 0 && colidx[e] < ncols, "column index out of range");
        FLGPU_REQUIRE(e == rowptr[r] || colidx[e] > colidx[e - 1],
                      "column indices must ascend strictly within a row");
      }
    }
    CTX_LOCK(ctx);
    flgpu_mat *A = mat_new(ctx, nrows, ncols, nnz, prec);
    try {
      FLGPU_HIP(hipMemcpyAsync(A->rowptr, rowptr, sizeof(int) * (nrows + 1), hipMemcpyHostToDevice,
                               ctx->stream));
      if (nnz)
        FLGPU_HIP(hipMemcpyAsync(A->colidx, colidx, sizeof(int) * nnz, hipMemcpyHostToDevice,
                                 ctx->stream));
      upload_vals(A, vals);
      mat_finish(A);
      FLGPU_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      mat_free(A);
      throw;
    }
    *out = A;
    return 0;
  });
}

extern "C" int flgpu_csr_update_values(flgpu_mat *A, const double *vals) {
  return guard([&] {
    FLGPU_REQUIRE(A && (vals || A->nnz == 0), "arguments must not be NULL");
    CTX_LOCK(A->ctx);
    drop_alpha(A);
    upload_vals(A, vals);
    mat_refresh_values(A);
    FLGPU_HIP(hipStreamSynchronize(A->ctx->stream));
    return 0;
  });
}

extern "C" int flgpu_csr_download(const flgpu_mat *A, int *rowptr, int *colidx, double *vals) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    CTX_LOCK(A->ctx);
    std::vector<int> rp, ci;
    std::vector<double> v;
    download_csr(A, rp, ci, vals ? &v : nullptr);
    if (rowptr) std::memcpy(rowptr, rp.data(), sizeof(int) * rp.size());
    if (colidx) std::memcpy(colidx, ci.data(), sizeof(int) * ci.size());
    if (vals) std::memcpy(vals, v.data(), sizeof(double) * v.size());
    return 0;
  });
}

extern "C" int flgpu_mat_info(const flgpu_mat *A, int *nrows, int *ncols, long *nnz, int *prec) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    if (nrows) *nrows = A->nrows;
    if (ncols) *ncols = A->ncols;
    if (nnz) *nnz = A->nnz;
    if (prec) *prec = A->prec;
    return 0;
  });
}

extern "C" int flgpu_mat_convert(const flgpu_mat *A, int prec, flgpu_mat **out) {
  return guard([&] {
    FLGPU_REQUIRE(A && out, "arguments must not be NULL");
    CTX_LOCK(A->ctx);
    *out = mat_convert(A, prec);
    FLGPU_HIP(hipStreamSynchronize(A->ctx->stream));
    return 0;
  });
}

extern "C" int flgpu_mat_storage(flgpu_mat *A, int max_form, int *form) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    CTX_LOCK(A->ctx);
    FLGPU_HIP(hipStreamSynchronize(A->ctx->stream));
    if (max_form < FLGPU_STORAGE_CELL) drop_alpha(A);
    if (max_form < FLGPU_STORAGE_STENCIL) {
      drop_stencil(A);
      tp_drop(A);
    }
    long lat_rows = 1;
    for (int k = 0; k < A->lat_dim; k++) lat_rows *= (long)A->lat_order * A->lat_n + 1;
    if (!A->alpha && !A->val27 && A->lat_dim > 0 && A->ncolours == 0 && A->nrows == lat_rows)
      lattice_colouring(A, A->lat_dim, A->lat_order, A->lat_n);
    if (form)
      *form = A->alpha                   ? FLGPU_STORAGE_CELL
              : (A->val27 || A->tp_kind) ? FLGPU_STORAGE_STENCIL
                                         : FLGPU_STORAGE_SELL;
    return 0;
  });
}

extern "C" int flgpu_mat_stencil_planes(const flgpu_mat *A, int *planes) {
  return guard([&] {
    FLGPU_REQUIRE(A && planes, "arguments must not be NULL");
    *planes = (A->alpha || !A->val27) ? 0 : (A->st_sym ? 14 : 27);
    return 0;
  });
}

extern "C" int flgpu_mat_destroy(flgpu_mat *A) {
  return guard([&] {
    if (!A) return 0;
    CTX_LOCK(A->ctx);
    FLGPU_HIP(hipStreamSynchronize(A->ctx->stream));
    mat_free(A);
    return 0;
  });
}

static void check_vec(const flgpu_mat *A, const flgpu_vec *v, long n, const char *what) {
  FLGPU_REQUIRE(v != nullptr, std::string(what) + " must not be NULL");
  FLGPU_REQUIRE(v->ctx == A->ctx, std::string(what) + " belongs to another context");
  FLGPU_REQUIRE(v->n == n, std::string(what) + ": length mismatch");
  FLGPU_REQUIRE(v->prec == A->prec, std::string(what) + ": precision differs from the matrix");
}

extern "C" int flgpu_gemm_nn(double alpha, const flgpu_mat *A, const flgpu_vec *y, double beta,
                             flgpu_vec *x) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    check_vec(A, y, A->ncols, "y");
    check_vec(A, x, A->nrows, "x");
    FLGPU_REQUIRE(x->d != y->d, "gemm!: x and y must be different vectors");
    CTX_LOCK(A->ctx);
    spmv(A, alpha, y->d, beta, x->d, x->d);
    return 0;
  });
}

extern "C" int flgpu_gemm_tn(double alpha, const flgpu_mat *A, const flgpu_vec *y, double beta,
                             flgpu_vec *x) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    check_vec(A, y, A->nrows, "y");
    check_vec(A, x, A->ncols, "x");
    FLGPU_REQUIRE(x->d != y->d, "gemm!: x and y must be different vectors");
    CTX_LOCK(A->ctx);
    mat_build_transpose(A);
    spmv(A->transpose, alpha, y->d, beta, x->d, x->d);
    return 0;
  });
}

extern "C" int flgpu_compute_residual(const flgpu_mat *A, const flgpu_vec *x, const flgpu_vec *b,
                                      flgpu_vec *r) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    FLGPU_REQUIRE(A->nrows == A->ncols, "compute-residual needs a square matrix");
    check_vec(A, x, A->ncols, "x");
    check_vec(A, b, A->nrows, "b");
    check_vec(A, r, A->nrows, "r");
    FLGPU_REQUIRE(x->d != r->d, "compute-residual: x and r must be different vectors");
    CTX_LOCK(A->ctx);
    spmv(A, -1.0, x->d, 1.0, b->d, r->d);
    return 0;
  });
}

extern "C" int flgpu_mat_set_colouring(flgpu_mat *A, int ncolours, const int *colour) {
  return guard([&] {
    FLGPU_REQUIRE(A && colour, "arguments must not be NULL");
    CTX_LOCK(A->ctx);
    FLGPU_HIP(hipStreamSynchronize(A->ctx->stream));
    drop_alpha(A);  // sweeps follow the caller's colours from now on
    drop_stencil(A);
    set_colouring(A, ncolours, colour);
    return 0;
  });
}

extern "C" int flgpu_mat_colour_greedy(flgpu_mat *A, int *ncolours) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    FLGPU_REQUIRE(A->nrows == A->ncols, "colouring needs a square matrix");
    CTX_LOCK(A->ctx);
    std::vector<int> rp, ci;
    download_csr(A, rp, ci, nullptr);
    std::vector<int> colour(A->nrows, -1);
    int nc = 0;
    std::vector<int> used;
    for (int r = 0; r < A->nrows; r++) {
      used.assign(nc + 1, 0);
      for (int e = rp[r]; e < rp[r + 1]; e++) {
        int c = ci[e];
        if (c != r && colour[c] >= 0) used[colour[c]] = 1;
      }
      // symmetric pattern assumed for the reverse couplings; check columns
      int k = 0;
      while (k < nc && used[k]) k++;
      colour[r] = k;
      if (k == nc) nc++;
    }
    FLGPU_HIP(hipStreamSynchronize(A->ctx->stream));
    drop_alpha(A);
    drop_stencil(A);
    set_colouring(A, nc, colour.data());
    if (ncolours) *ncolours = nc;
    return 0;
  });
}

extern "C" int flgpu_jacobi_steps(const flgpu_mat *A, double damp, int steps, flgpu_vec *x,
                                  const flgpu_vec *b, flgpu_vec *tmp) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    FLGPU_REQUIRE(A->nrows == A->ncols, "<jacobi> needs a square matrix");
    FLGPU_REQUIRE(steps >= 0, "negative number of steps");
    check_vec(A, x, A->nrows, "x");
    check_vec(A, b, A->nrows, "b");
    check_vec(A, tmp, A->nrows, "tmp");
    FLGPU_REQUIRE(x != tmp && b != tmp && x != b, "x, b and tmp must be distinct vectors");
    CTX_LOCK(A->ctx);
    jacobi_steps(A, damp, steps, x->d, b->d, tmp->d);
    return 0;
  });
}

extern "C" int flgpu_jacobi_iterate(const flgpu_mat *A, double damp, flgpu_vec *x, flgpu_vec *r) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    FLGPU_REQUIRE(A->nrows == A->ncols, "<jacobi> needs a square matrix");
    check_vec(A, x, A->nrows, "x");
    check_vec(A, r, A->nrows, "r");
    CTX_LOCK(A->ctx);
    jacobi_iterate(A, damp, x->d, r->d);
    return 0;
  });
}

extern "C" int flgpu_sor_iterate(const flgpu_mat *A, double omega, int backwards, flgpu_vec *x,
                                 const flgpu_vec *b, flgpu_vec *r) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    check_vec(A, x, A->nrows, "x");
    check_vec(A, b, A->nrows, "b");
    check_vec(A, r, A->nrows, "r");
    CTX_LOCK(A->ctx);
    sor_sweeps(A, omega, 1, backwards, x->d, b->d, r->d);
    return 0;
  });
}
