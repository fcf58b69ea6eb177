// Internal declarations of libflgpu.so (gfx950).  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/flgpu.h"

namespace flgpu {

// ---- errors ---------------------------------------------------------------

void set_error(const std::string &msg);
void clear_error();

struct Error {
  int code;
  std::string msg;
};

#define FLGPU_HIP(call)                                                        \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw ::flgpu::Error{-2, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define FLGPU_REQUIRE(cond, msg)                                               \
  do {                                                                         \
    if (!(cond)) throw ::flgpu::Error{-1, (msg)};                              \
  } while (0)

// Runs f, converting exceptions to status codes and the thread's last error.
template <class F>
int guard(F &&f) {
  try {
    clear_error();
    return f();
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return -3;
  } catch (const std::exception &e) {
    set_error(e.what());
    return -4;
  }
}

// Locks a context for the rest of the scope and makes its device current.
#define CTX_LOCK(ctx)                                        \
  std::lock_guard<std::recursive_mutex> lk_((ctx)->lock);    \
  FLGPU_HIP(hipSetDevice((ctx)->device))

// ---- device buffers ---------------------------------------------------------

void *dev_alloc(size_t bytes);
void dev_free(void *p);
inline size_t prec_bytes(int prec) { return prec == FLGPU_F64 ? 8 : 4; }

// ---- launch geometry --------------------------------------------------------

constexpr int kWave = 64;
constexpr int kSlice = 64;  // SELL slice height = one wavefront

inline unsigned grid_for(long n, int block) {
  long g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

// Remaps the hardware block index so that consecutive logical blocks run on
// the same XCD: blocks are dealt round-robin over the 8 XCDs, so block b runs
// on XCD b % 8; logical = (b % 8) * (G8/8) + b / 8 keeps contiguous row
// ranges (and their shared vector neighbours) in one XCD's L2.  Logical
// indices beyond the work are skipped by the caller.
// The grid must be launched with a multiple of 8 blocks (xcd_grid) so that
// the map is a bijection of [0, gridDim.x).
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned G8) {
  unsigned per = G8 / 8u;
  return (b % 8u) * per + b / 8u;
}
inline unsigned xcd_grid(unsigned g) { return (g + 7u) & ~7u; }

// n / d for 0 <= n < 2^31 and d >= 2 by one multiply (Granlund & Montgomery 1994, theorem
// 4.2 with N = 31): l = ceil(log2 d), m = ceil(2^(31+l) / d) < 2^32, n / d = umulhi(n, m) >>
// (l - 1).  m d - 2^(31+l) lies in [0, d) and d <= 2^l, which is the theorem's condition.
struct FastDiv {
  unsigned d, m;
  int s;
};
inline FastDiv fast_div(unsigned d) {
  int l = 0;
  while ((1ull << l) < d) l++;
  unsigned long long m = ((1ull << (31 + l)) + d - 1) / d;
  return FastDiv{d, (unsigned)m, l - 1};
}
__device__ __forceinline__ unsigned fdiv(unsigned n, FastDiv f) { return __umulhi(n, f.m) >> f.s; }

}  // namespace flgpu

// ---- handles ------------------------------------------------------------------

struct flgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::recursive_mutex lock;
  int cus = 0;
  double *red_partial = nullptr;  // device, reduction partials
  double *red_out = nullptr;      // device, reduction results (64 slots)
  double *host_out = nullptr;     // pinned host copy of red_out
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  hipEvent_t ev_fetch = nullptr;  // marks the end of an asynchronous slot fetch
  // flgpu_solve's work vectors, kept between solves (the context lock serialises solves)
  void *solve_buf = nullptr;
  size_t solve_bytes = 0;
  std::string name;
};

struct flgpu_vec {
  flgpu_ctx *ctx = nullptr;
  long n = 0;
  int prec = FLGPU_F64;
  void *d = nullptr;
};

// SELL-64: rows grouped in slices of 64 (one wavefront), entries of a slice
// stored column-major (entry k of lane l at slice_ptr[s] + 64 k + l) so that a
// wavefront reads 64 consecutive values per step.  `rows` maps slice lanes to
// matrix rows (nullptr = identity); padding has value 0 and a valid column.
struct Sell {
  int nrows = 0;
  int nslices = 0;
  int *rows = nullptr;
  long *slice_ptr = nullptr;
  int *col = nullptr;
  void *val = nullptr;
  long nstored = 0;
  int maxw = 0;
};

struct flgpu_mat {
  flgpu_ctx *ctx = nullptr;
  int nrows = 0, ncols = 0;
  long nnz = 0;
  int prec = FLGPU_F64;
  int *rowptr = nullptr;  // device CSR
  int *colidx = nullptr;
  void *vals = nullptr;
  int *diagpos = nullptr;  // CSR position of the diagonal (-1 if absent)
  void *diag = nullptr;    // diagonal values (prec)
  Sell sell;
  int ncolours = 0;
  std::vector<Sell> colour_sell;
  flgpu_mat *transpose = nullptr;
  // lattice description (0 if the matrix was not assembled on a lattice)
  int lat_dim = 0, lat_order = 0, lat_n = 0;
  double lat_f = 1.0;
  // 3D Q1 with a coefficient constant per cell: element matrix of cell c is
  // alpha[c] * Kref, so A_ij = kt[|d|] * (sum of alpha over the cells shared
  // by DOFs i and j), d = lattice offset; kernels apply A from alpha
  void *alpha = nullptr;  // n^3 cell factors (prec), nullptr if not available
  double kt[8] = {};      // Kref[0][b], b = bz*4 + by*2 + bx
  // with alpha: the <sor> sweep's strips per plane (0: one launch per colour) and its copy of
  // x and b around the strip boundaries, 2 (M (S - 1) 8 M) values (prec), nullptr if S < 2
  int sor_strips = 0, sor_h = 0;
  void *sor_ghost = nullptr;
  // 3D Q1 lattice matrices otherwise: the 27 values of every row, SoA
  // val27[k][row], k = (dz+1)*9 + (dy+1)*3 + (dx+1); columns are implied
  void *val27 = nullptr;
  bool st_sym = false;  // val27 exactly symmetric: kernels stream 14 of the 27 planes
  // order-1 3D lattice interpolation from a (tp_nc+1)^3 to a (2 tp_nc+1)^3 lattice with the
  // full trilinear pattern (1: the interpolation, 2: its transpose): products read the CSR
  // values with the columns and row starts implied by the pattern (0: not such a matrix)
  int tp_kind = 0, tp_nc = 0;
  // the values again as 27 planes over the coarse nodes c: for the transpose (tp_kind 2)
  // tp27[k][c], k = (dz+1)*9 + (dy+1)*3 + (dx+1), the entry of fine neighbour 2c + d; for the
  // interpolation (tp_kind 1) the entries of the fine nodes of c's block by parity class
  // (k_tp8_fill); 0 where there is no such entry
  void *tp27 = nullptr;
  // block smoothers on a lattice matrix: inverses of the vertex-centred blocks, per vertex in
  // lattice vertex order, blk_inv[v][j][i] = (A_BB^-1)_ij (prec); nullptr until needed and for
  // order 1, whose blocks are single DOFs
  void *blk_inv = nullptr;
  // per-block partial sums of the per-cell product's fused dot (spmv_dot), dot_np entries
  double *dot_part = nullptr;
  long dot_np = 0;
  // counts the changes of the buffers and forms the kernels read (a form dropped or built, a
  // buffer freed or allocated, the colour subsets rebuilt, the stencil's symmetry changed);
  // captured V-cycles address those buffers and are dropped when any count has moved
  unsigned long layout = 0;
};

namespace flgpu {

inline void layout_changed(flgpu_mat *A) { A->layout++; }

// matrices (flgpu_sparse.hip)
flgpu_mat *mat_new(flgpu_ctx *ctx, int nrows, int ncols, long nnz, int prec);
void mat_free(flgpu_mat *A);
void mat_finish(flgpu_mat *A);  // SELL + diagonal from the device CSR
void mat_refresh_values(flgpu_mat *A);  // SELL/diag/colour values after a CSR value change
void mat_build_transpose(const flgpu_mat *A);
void drop_alpha(flgpu_mat *A);
// the strips and boundary copy of the per-cell <sor> sweep, after alpha is allocated
void sor_strips_alloc(flgpu_mat *A);
void tp_prepare(flgpu_mat *T);  // (re)fill tp27 of an interpolation or its transpose
void tp_drop(flgpu_mat *A);     // products of A (and its transpose) go back to SELL
bool is_q1_lattice(const flgpu_mat *A);
void stencil_refresh(flgpu_mat *A);
void drop_stencil(flgpu_mat *A);
flgpu_mat *mat_convert(const flgpu_mat *A, int prec);
void sell_free(Sell &s);
Sell sell_build_rows(flgpu_mat *A, const std::vector<int> &rows);
void set_colouring(flgpu_mat *A, int ncolours, const int *colour_host);
// lattice: P is a 3D Q1 lattice interpolation and A, Ac hold the lattice patterns of the fine
// and the coarse lattice (the product then computes positions instead of searching them)
void galerkin_product(const flgpu_mat *P, const flgpu_mat *A, flgpu_mat *Ac, bool lattice = false);

// out[row] = beta*z[row] + alpha*(A x)[row]  (z may be out; beta == 0 ignores z)
void spmv(const flgpu_mat *A, double alpha, const void *x, double beta,
          const void *z, void *out);
void spmv_t(const flgpu_mat *A, double alpha, const void *x, double beta,
            void *out);
// out = A x and the dot (out, x) into reduction slot `slot` in one pass over A, when A is
// applied in the per-cell form; false (nothing launched) otherwise
bool spmv_dot(const flgpu_mat *A, const void *x, void *out, int slot);
void jacobi_iterate(const flgpu_mat *A, double damp, void *x, void *r);
void sor_sweep(const flgpu_mat *A, double omega, int backwards, void *x,
               const void *b, void *r);
// `steps` <jacobi> steps with fresh residuals, x ping-ponged through tmp; x_zero: x is zero
// and need not be read (the first step is x = damp b / diag)
// x64 (FP64, n values) may receive a copy of the final x: true when the last step wrote it
bool jacobi_steps(const flgpu_mat *A, double damp, int steps, void *x, const void *b, void *tmp,
                  bool x_zero = false, int done = 0, double *x64 = nullptr);
// rhs = P^T r (beta 0) and, in the same pass, xo = damp (rhs / diag(Ac)): the first <jacobi>
// step from zero on the coarse level; false (nothing launched) unless PT is a structured
// transpose (tp_kind 2)
bool restrict_first_step(const flgpu_mat *PT, const void *r, void *rhs, const flgpu_mat *Ac,
                         double damp, void *xo);
// b = r converted to A's precision and xo = damp (b / diag(A)), the first <jacobi> step from
// zero, in one pass
void stage_first_step(const flgpu_mat *A, int rprec, const void *r, void *b, double damp,
                      void *xo);
// vertex-centred block smoothers (flgpu_block.hip): (re)compute the block inverses of a
// lattice matrix, release them, and one <geometric-ssc> (ssc) / <geometric-psc> iterate
// (backwards: colours last to first; b may be r)
void blocks_prepare(flgpu_mat *A);
void blocks_drop(flgpu_mat *A);
void block_sweep(const flgpu_mat *A, bool ssc, double omega, int backwards, void *x,
                 const void *b, void *r);
// `steps` sweeps
void sor_sweeps(const flgpu_mat *A, double omega, int steps, int backwards, void *x,
                const void *b, void *r);

// lattices (flgpu_lattice.hip)
flgpu_mat *lattice_assemble(flgpu_ctx *ctx, int dim, int order, int n, const flgpu_vec *kappa,
                            int nk, double f, int prec, int mode, flgpu_vec **rhs);
flgpu_mat *lattice_interpolation(flgpu_ctx *ctx, int dim, int order, int nc, int prec);
flgpu_mat *lattice_projection(flgpu_ctx *ctx, int dim, int order, int nc, int prec);
void lattice_colouring(flgpu_mat *A, int dim, int order, int n);
flgpu_mat *lattice_pattern(flgpu_ctx *ctx, int dim, int order, int n, int prec);
flgpu_mat *lattice_galerkin(const flgpu_mat *P, const flgpu_mat *Afine, int dim, int order,
                            int nc);
void lattice_dirichlet_identity(flgpu_mat *A, int dim, int order, int n);
// operator-dependent interpolation values of P (3D Q1) from the fine matrix A
void lattice_blackbox(const flgpu_mat *A, flgpu_mat *P);
void lattice_reassemble(flgpu_mat *A, const flgpu_vec *kappa, int nk, int mode);

// vectors (flgpu_core.hip)
flgpu_vec *vec_new(flgpu_ctx *ctx, long n, int prec);
void vec_free(flgpu_vec *v);
void vfill(flgpu_ctx *ctx, long n, int prec, void *x, double value);
void vcopy(flgpu_ctx *ctx, long n, int sprec, const void *src, int dprec,
           void *dst);
void vscal(flgpu_ctx *ctx, long n, int prec, double a, void *x);
void vaxpy(flgpu_ctx *ctx, long n, int prec, double a, const void *x, void *y);
// result stays on device in ctx->red_out[slot]; dot_host syncs and returns it
void vdot_dev(flgpu_ctx *ctx, long n, int prec, const void *x, const void *y,
              int slot);
// the sum of np per-block partials into reduction slot `slot`
void dot_final(flgpu_ctx *ctx, const double *partial, int np, int slot);
// 1 to 3 dots in one pass, each identical to vdot_dev's
void vdot_multi_dev(flgpu_ctx *ctx, long n, int prec, int count, const void *const *x,
                    const void *const *y, const int *slots);
void vdot_qr_f32(flgpu_ctx *ctx, long n, const float *z, const double *r, const double *a,
                 const int *slots);
double vdot_host(flgpu_ctx *ctx, long n, int prec, const void *x,
                 const void *y);
void fetch_slots(flgpu_ctx *ctx, int nslots);  // red_out[0..n) -> host_out
// the same copy queued on the stream; host_out is valid after fetch_wait, while work queued
// in between runs on
void fetch_slots_async(flgpu_ctx *ctx, int nslots);
void fetch_wait(flgpu_ctx *ctx);

}  // namespace flgpu
