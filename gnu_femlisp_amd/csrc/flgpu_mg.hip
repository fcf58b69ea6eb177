// Geometric multigrid and the Krylov/solver loop of libflgpu.so.
//
// Restated femlisp sources:
//   <mg-iteration> iterator, lmgc, CS restrict/prolongate
//                          src/iteration/multigrid.lisp:173-288,303-393
//   multilevel decomposition src/iteration/geomg.lisp:109-142
//   product-iterator       src/iteration/linit.lisp:94-120
//   <lu> coarse iterator   src/iteration/linit.lisp:132-162
//   <cg>                   src/iteration/krylow.lisp:67-132
//   <linear-solver> loop   src/iteration/linsolve.lisp:58-73,
//                          src/iteration/solve.lisp:126-165,
//                          src/iteration/iterate.lisp:239-250
//
// The V-cycle keeps femlisp's bookkeeping (residual-p flags per level, which
// decide when a residual is recomputed) so that the operator applied is the
// reference's, including its (iterate q w w) aliasing when asked for.
// Scalars of the Krylov recurrences stay on the device; the host reads them
// once per iteration (the defect norm femlisp's loop observes).
#include <chrono>
#include <cmath>
#include <cstring>

#include "flgpu_internal.h"
#include "flgpu_learned.h"

namespace flgpu {

// A^-1 of a small square matrix, compressed: a DOF whose row is the unit row e_k and whose
// column is zero apart from its diagonal (a Dirichlet DOF of a lattice matrix) is trivial, and
// A^-1 keeps e_k there (x_k = r_k).  The other m DOFs (cidx, ascending) form the m x m matrix
// C whose inverse is kept (FP64, row-major).
// For an FP32 cycle the inverse is kept as FP32 64 x 64 tiles of its upper triangle (tiles
// (I, J), I <= J, in the order of tI / tJ; inv freed), applied by two passes through part.
struct DenseInverse {
  double *inv = nullptr;
  float *tiles = nullptr;
  int *tI = nullptr, *tJ = nullptr;  // the tiles' block row and column
  double *part = nullptr;            // ntile x ntile x 64 partial products
  int ntile = 0;
  int *cidx = nullptr;  // the m non-trivial DOFs
  int *tidx = nullptr;  // the nt trivial DOFs
  int m = 0, nt = 0;
  void release() {
    dev_free(inv);
    dev_free(tiles);
    dev_free(tI);
    dev_free(tJ);
    dev_free(part);
    dev_free(cidx);
    dev_free(tidx);
    inv = nullptr;
    tiles = nullptr;
    tI = tJ = nullptr;
    part = nullptr;
    cidx = tidx = nullptr;
    m = nt = ntile = 0;
  }
};

struct MGLevel {
  flgpu_mat *A = nullptr;  // owned
  flgpu_mat *I = nullptr;  // owned: interpolation level -> level+1 (transpose attached)
  flgpu_mat *R = nullptr;  // owned, FAS only: projection level+1 -> level (:fas-r-vec)
  void *sol = nullptr, *rhs = nullptr, *res = nullptr;  // owned work vectors
  void *tmp = nullptr;  // <jacobi> steps: the other half of x's ping-pong (smoother Jacobi)
  bool residual_p = false;
  bool x_zero = false;  // sol is zero but not stored: the next smoother step or coarse solve
                        // writes it instead of updating it
  bool step1 = false;   // with x_zero: tmp already holds the first <jacobi> step from zero
                        // (written by the pass that produced rhs)
  DenseInverse cinv;  // base level: dense inverse (FP64, compressed)
  long N = 0;
};

}  // namespace flgpu

struct flgpu_mg {
  flgpu_ctx *ctx = nullptr;
  flgpu_mg_params p{};
  std::vector<flgpu::MGLevel> lv;
  int top = 0;
  // FP64 staging when an FP32 cycle is applied to FP64 vectors
  void *xs = nullptr, *bs = nullptr, *rs = nullptr;
  // active top-level pointers during one cycle
  void *sol_top = nullptr, *rhs_top = nullptr, *res_top = nullptr;
  // lattice description (re-assembly for a new coefficient)
  int lat_dim = 0, lat_order = 0, lat_n0 = 0, asm_mode = FLGPU_ASM_F64;
  // attached learned correction
  flgpu_unet *unet = nullptr;
  const flgpu_vec *unet_kappa = nullptr;
  // the learned correction's hook (its buffers) for the matrix hook_A of hook_N rows, kept
  // between solves; dropped when another network is attached
  std::unique_ptr<flgpu::LearnedHook> hook;
  const flgpu_mat *hook_A = nullptr;
  long hook_N = 0;
  int unet_n = 0;
  // an FP32 cycle for FP64 vectors: where the last top-level step may also store its result
  // in FP64 (set for one cycle; x64_done tells whether it did)
  double *x64 = nullptr;
  bool x64_done = false;
  // captured cycles: one HIP graph per (x, b, r, precision) of the caller
  struct Graph {
    void *x, *b, *r;
    int vprec, mode;  // mg_run's mode: 0 x += V(r); 1-4 x = V(b) from zero (staged, FP32 kept)
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  int use_graphs = 1;
  // the level matrices' layout counts when the cached graphs were captured
  unsigned long graph_layout = 0;
};

namespace flgpu {

// ---- coarse solve: x += A^-1 r (FP64 dense inverse, compressed) ------------------------------
// A block stages r at the m non-trivial DOFs in LDS (FP64), then each of its waves takes one
// row: loads issued in batches of 16 per lane before the products are summed (lane l sums
// j = l, l + 64, ... in order), a wave reduction, and lane 0 updates x at the row's DOF.  The
// blocks after the dense rows take the trivial DOFs (x_k = r_k).
// femlisp's <lu> iterate also leaves A^-1 r in r; nothing reads r afterwards (residual-p is
// cleared), and r is read by every row of this pass, so it keeps the restricted residual.
template <class T, class TA, int U>
__global__ __launch_bounds__(256) void k_coarse_solve(int m, const TA *__restrict__ Ainv,
                                                      const int *__restrict__ cidx, int nt,
                                                      const int *__restrict__ tidx,
                                                      const T *__restrict__ r, T *__restrict__ x,
                                                      int x_zero) {
  extern __shared__ double rc[];  // m values
  const int nbd = (m + 3) / 4;    // blocks of the dense rows
  if ((int)blockIdx.x >= nbd) {
    const int t = ((int)blockIdx.x - nbd) * 256 + (int)threadIdx.x;
    if (t < nt) {
      const int k = tidx[t];
      const T v = r[k];
      x[k] = x_zero ? v : x[k] + v;
    }
    return;
  }
  // 16 values per thread and round, their index loads and then their value loads all in
  // flight together (the index clamped into the list, the values beyond it not stored)
  for (int j0 = 0; j0 < m; j0 += 256 * 16) {
    int ci[16];
#pragma unroll
    for (int u = 0; u < 16; u++) ci[u] = cidx[min(j0 + (int)threadIdx.x + 256 * u, m - 1)];
    T rv[16];
#pragma unroll
    for (int u = 0; u < 16; u++) rv[u] = r[ci[u]];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int j = j0 + (int)threadIdx.x + 256 * u;
      if (j < m) rc[j] = (double)rv[u];
    }
  }
  __syncthreads();
  int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= m) return;
  int lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int j0 = 0; j0 < m; j0 += 64 * U) {  // batches of 64 U columns
    TA a[U];
    // every load unconditional (the index clamped into the row) so that all are in flight
    // together; the terms beyond the row are dropped from the sum, which keeps its order
#pragma unroll
    for (int u = 0; u < U; u++) a[u] = Ainv[(long)row * m + min(j0 + lane + 64 * u, m - 1)];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int j = j0 + lane + 64 * u;
      acc = j < m ? acc + (double)a[u] * rc[j] : acc;
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) {
    const int k = cidx[row];
    T v = (T)acc;
    x[k] = x_zero ? v : x[k] + v;
  }
}

// ---- dense FP64 inverse on the GPU -------------------------------------------------------
// Gauss-Jordan elimination without pivoting (the matrices inverted here -- coarse levels and
// the <lu> fallback's system -- are SPD after the Dirichlet elimination) in the textbook form:
// for each pivot k, row k is divided by a_kk (a_kk becomes 1 / a_kk) and a_ik-multiples of it
// are subtracted from the other rows (a_ik becoming -a_ik / a_kk); after the last pivot the
// array holds the inverse.  Blocked by kGjB pivots: with K the pivot block and D = C_KK^-1
// (the unblocked elimination of the block alone), the pivots of K together take
//   C_iK <- -C_iK D,  C_KJ <- D C_KJ,  C_iJ <- C_iJ - C_iK D C_KJ,  C_KK <- D
// for rows i and columns J outside K, the rank-kGjB update on FP64 MFMA.

__global__ void k_dense_zero(long nn, double *M) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nn) M[i] = 0.0;
}

template <class T>
__global__ void k_dense_scatter(int n, const int *__restrict__ rowptr,
                                const int *__restrict__ colidx, const T *__restrict__ vals,
                                double *M) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int e = rowptr[r]; e < rowptr[r + 1]; e++) M[(long)r * n + colidx[e]] = (double)vals[e];
}

// entry b of the new row k, and entry (a, b) after the pivot k, from the values before it (one
// expression each, shared by both kernels below so that they round alike)
__device__ __forceinline__ double gj_row(double skb, double skk, bool b_is_k) {
  return (b_is_k ? 1.0 : skb) / skk;
}
__device__ __forceinline__ double gj_entry(double sab, double sak, double rkb, bool a_is_k,
                                           bool b_is_k) {
  return a_is_k ? rkb : (b_is_k ? 0.0 : sab) - sak * rkb;
}

// trivial[k] = 1 when row k is the unit row e_k and column k is +0 apart from its diagonal
// (a Dirichlet row and column of a lattice matrix).  Such a pivot is skipped: row k stays e_k
// under every pivot, every other row subtracts zero multiples of it, and its column keeps
// zeros, so the inverse is the same.  With positive pivots (SPD matrices) it is the same bit
// for bit, since an entry that starts at +0 only ever has +0 or -0 subtracted from it; a
// negative pivot can turn a zero of column k into -0, which the skipped pivot would have made
// +0 again.
__global__ __launch_bounds__(256) void k_dense_trivial(int n, const double *__restrict__ M,
                                                       int *trivial) {
  const int k = blockIdx.x;  // one block per pivot, its threads stride over row and column k
  bool t = true;
  for (int j = threadIdx.x; j < n; j += blockDim.x)
    t = t && (j == k ? M[(long)k * n + k] == 1.0
                     : __double_as_longlong(M[(long)k * n + j]) == 0 &&
                           __double_as_longlong(M[(long)j * n + k]) == 0);
  t = __syncthreads_and(t);
  if (threadIdx.x == 0) trivial[k] = t ? 1 : 0;
}

__global__ void k_dense_gather(int m, int n, const int *__restrict__ cidx,
                               const double *__restrict__ M, double *__restrict__ C) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * m) return;
  const int a = (int)(e / m), b = (int)(e - (long)a * m);
  C[e] = M[(long)cidx[a] * n + cidx[b]];
}

constexpr int kGjB = 64;  // pivots per block
typedef double v4d __attribute__((ext_vector_type(4)));

// D = C_KK^-1 (bs x bs, bs <= kGjB, zero-padded to kGjB x kGjB) by the unblocked elimination,
// pivot by pivot; one workgroup.  Thread t holds the entries (t / 64 + 4 u, t % 64), u < 16, in
// registers; per pivot k the holders of column k and row k publish their values before it
// (LDS, two buffers by the parity of k, so one barrier per pivot), and every thread forms its
// new entries with gj_row / gj_entry.
__global__ __launch_bounds__(256) void k_gj_diag(int m, int k0, int bs,
                                                 const double *__restrict__ C,
                                                 double *__restrict__ D, int *bad) {
  __shared__ double colk[2][kGjB], rowk[2][kGjB];
  constexpr int kPer = kGjB * kGjB / 256;
  const int b = threadIdx.x % kGjB, a0 = threadIdx.x / kGjB;
  double v[kPer];
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int a = a0 + 4 * u;
    v[u] = (a < bs && b < bs) ? C[(long)(k0 + a) * m + k0 + b] : 0.0;
  }
  for (int k = 0; k < bs; k++) {
    const int par = k & 1;
    if (b == k) {
#pragma unroll
      for (int u = 0; u < kPer; u++) colk[par][a0 + 4 * u] = v[u];
    }
    if (a0 == (k & 3)) {
#pragma unroll
      for (int u = 0; u < kPer; u++)
        if (u == (k >> 2)) rowk[par][b] = v[u];
    }
    __syncthreads();
    const double piv = colk[par][k];
    if (threadIdx.x == 0 && piv == 0.0) *bad = 1;
    const double rkb = gj_row(rowk[par][b], piv, b == k);
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      const int a = a0 + 4 * u;
      v[u] = gj_entry(v[u], colk[par][a], rkb, a == k, b == k);
    }
  }
#pragma unroll
  for (int u = 0; u < kPer; u++) D[(a0 + 4 * u) * kGjB + b] = v[u];
}

// U (m x kGjB) = C_iK D for the rows i outside K and D's rows inside K; R (kGjB x m) = C_KJ
// before the update (rows and columns beyond bs are zero)
__global__ __launch_bounds__(256) void k_gj_panels(int m, int k0, int bs,
                                                   const double *__restrict__ C,
                                                   const double *__restrict__ D,
                                                   double *__restrict__ U, double *__restrict__ R) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nu = (long)m * kGjB;
  if (e < nu) {
    const int i = (int)(e / kGjB), c = (int)(e % kGjB);
    double v = 0.0;
    if (i >= k0 && i < k0 + bs) {
      v = D[(i - k0) * kGjB + c];
    } else if (c < bs) {
      for (int t = 0; t < bs; t++) v += C[(long)i * m + k0 + t] * D[t * kGjB + c];
    }
    U[e] = v;
  } else if (e < 2 * nu) {
    const long f = e - nu;
    const int t = (int)(f / m), j = (int)(f - (long)t * m);
    R[f] = t < bs ? C[(long)(k0 + t) * m + j] : 0.0;
  }
}

// one kGjB x kGjB tile (I, J) of C per workgroup, K = the pivot block:
//   J == K: C_iK = D_i (i in K) or -U_i (i outside K)
//   J != K: C_iJ = D_i R_J (i in K) or C_iJ - U_i R_J (i outside K)
// U_I R_J on FP64 MFMA 16x16x4 from LDS (lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and receives C[l / 16 + 4 v][l % 16]); wave w forms the 32 x 32 quarter
// (w / 2, w % 2) of the tile
__global__ __launch_bounds__(256) void k_gj_update(int m, int k0, const double *__restrict__ U,
                                                   const double *__restrict__ R,
                                                   double *__restrict__ C) {
  __shared__ double us[kGjB][kGjB + 1];
  __shared__ double rs[kGjB][kGjB + 2];
  const int ntile = (m + kGjB - 1) / kGjB;
  const int I = (int)blockIdx.x / ntile, J = (int)blockIdx.x % ntile, kb = k0 / kGjB;
  const int i0 = I * kGjB, j0 = J * kGjB;
  constexpr int kPer = kGjB * kGjB / 256;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, kg = lane >> 4;
  const int r0 = (w >> 1) * 32, c0 = (w & 1) * 32;
  // this thread's 16 entries of the tile, loaded first (index clamped into C) so that their
  // loads overlap the staging below
  double cv[2][2][4];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = min(i0 + r0 + 16 * p + kg + 4 * v, m - 1);
        const int j = min(j0 + c0 + 16 * q + li, m - 1);
        cv[p][q][v] = C[(long)i * m + j];
      }
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int e = threadIdx.x + 256 * u, a = e / kGjB, b = e % kGjB;
    us[a][b] = i0 + a < m ? U[(long)(i0 + a) * kGjB + b] : 0.0;
    rs[a][b] = j0 + b < m ? R[(long)a * m + j0 + b] : 0.0;
  }
  __syncthreads();
  v4d acc[2][2];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int q = 0; q < 2; q++) acc[p][q] = v4d{0.0, 0.0, 0.0, 0.0};
  if (J != kb) {
    for (int k = 0; k < kGjB; k += 4)
#pragma unroll
      for (int p = 0; p < 2; p++)
#pragma unroll
        for (int q = 0; q < 2; q++)
          acc[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              us[r0 + 16 * p + li][k + kg], rs[k + kg][c0 + 16 * q + li], acc[p][q], 0, 0, 0);
  }
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int a = r0 + 16 * p + kg + 4 * v, b = c0 + 16 * q + li;
        const int i = i0 + a, j = j0 + b;
        if (i < m && j < m) {
          double out;
          if (J == kb) out = I == kb ? us[a][b] : -us[a][b];
          else out = I == kb ? acc[p][q][v] : cv[p][q][v] - acc[p][q][v];
          C[(long)i * m + j] = out;
        }
      }
}

// A^-1 of a square matrix of at most 8192 rows, compressed
static DenseInverse dense_inverse_dev(const flgpu_mat *A) {
  FLGPU_REQUIRE(A->nrows == A->ncols, "dense inverse of a non-square matrix");
  FLGPU_REQUIRE(A->nrows <= 8192, "matrix too large for the dense inverse (> 8192 rows)");
  const int n = A->nrows;
  const long nn = (long)n * n;
  hipStream_t s = A->ctx->stream;
  DenseInverse out;
  double *M0 = nullptr, *D = nullptr, *U = nullptr, *R = nullptr;
  int *bad = nullptr, *dt = nullptr;
  try {
    if (n == 0) return out;
    M0 = (double *)dev_alloc(sizeof(double) * nn);
    dt = (int *)dev_alloc(sizeof(int) * n);
    k_dense_zero<<<grid_for(nn, 256), 256, 0, s>>>(nn, M0);
    if (A->prec == FLGPU_F64)
      k_dense_scatter<double><<<grid_for(n, 256), 256, 0, s>>>(n, A->rowptr, A->colidx,
                                                               (const double *)A->vals, M0);
    else
      k_dense_scatter<float><<<grid_for(n, 256), 256, 0, s>>>(n, A->rowptr, A->colidx,
                                                              (const float *)A->vals, M0);
    k_dense_trivial<<<n, 256, 0, s>>>(n, M0, dt);
    FLGPU_HIP(hipGetLastError());
    std::vector<int> trivial(n, 0), piv, triv;
    FLGPU_HIP(hipMemcpyAsync(trivial.data(), dt, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    FLGPU_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < n; k++) (trivial[k] ? triv : piv).push_back(k);
    const int m = (int)piv.size();
    out.m = m;
    out.nt = (int)triv.size();
    if (out.nt > 0) {
      out.tidx = (int *)dev_alloc(sizeof(int) * out.nt);
      FLGPU_HIP(hipMemcpyAsync(out.tidx, triv.data(), sizeof(int) * out.nt,
                               hipMemcpyHostToDevice, s));
    }
    if (m > 0) {
      const long mm = (long)m * m;
      out.cidx = (int *)dev_alloc(sizeof(int) * m);
      FLGPU_HIP(hipMemcpyAsync(out.cidx, piv.data(), sizeof(int) * m, hipMemcpyHostToDevice, s));
      out.inv = (double *)dev_alloc(sizeof(double) * mm);
      k_dense_gather<<<grid_for(mm, 256), 256, 0, s>>>(m, n, out.cidx, M0, out.inv);
      FLGPU_HIP(hipGetLastError());
      D = (double *)dev_alloc(sizeof(double) * kGjB * kGjB);
      U = (double *)dev_alloc(sizeof(double) * (size_t)m * kGjB);
      R = (double *)dev_alloc(sizeof(double) * (size_t)m * kGjB);
      bad = (int *)dev_alloc(sizeof(int));
      FLGPU_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
      const int ntile = (m + kGjB - 1) / kGjB;
      for (int k0 = 0; k0 < m; k0 += kGjB) {
        const int bs = std::min(kGjB, m - k0);
        k_gj_diag<<<1, 256, 0, s>>>(m, k0, bs, out.inv, D, bad);
        k_gj_panels<<<grid_for(2L * m * kGjB, 256), 256, 0, s>>>(m, k0, bs, out.inv, D, U, R);
        k_gj_update<<<(unsigned)(ntile * ntile), 256, 0, s>>>(m, k0, U, R, out.inv);
      }
      FLGPU_HIP(hipGetLastError());
      int hbad = 0;
      FLGPU_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
      FLGPU_HIP(hipStreamSynchronize(s));
      FLGPU_REQUIRE(hbad == 0, "matrix is singular (zero pivot in the dense inverse)");
    }
    FLGPU_HIP(hipStreamSynchronize(s));
  } catch (...) {
    dev_free(M0);
    dev_free(dt);
    dev_free(D);
    dev_free(U);
    dev_free(R);
    dev_free(bad);
    out.release();
    throw;
  }
  dev_free(M0);
  dev_free(dt);
  dev_free(D);
  dev_free(U);
  dev_free(R);
  dev_free(bad);
  return out;
}

// The symmetric inverse from the tiles of its upper triangle, pass 1: one workgroup per tile
// (I, J) stages it in LDS with r at its rows and columns, and writes T r_J (row sums) to
// part[I][J] and, for I < J, T^T r_I (column sums) to part[J][I]; four threads per row or
// column, 16 products each in order, then combined in a fixed order (deterministic)
constexpr int kSymB = 64;
template <class T>
__global__ __launch_bounds__(256) void k_coarse_sym1(int m, int ntile,
                                                     const float *__restrict__ tiles,
                                                     const int *__restrict__ tI,
                                                     const int *__restrict__ tJ,
                                                     const int *__restrict__ cidx,
                                                     const T *__restrict__ r,
                                                     double *__restrict__ part) {
  __shared__ float tl[kSymB][kSymB + 1];
  __shared__ double rI[kSymB], rJ[kSymB];
  const int q = blockIdx.x, I = tI[q], J = tJ[q];
  const float4 *tq = (const float4 *)(tiles + (long)q * kSymB * kSymB);
  float4 tv[kSymB * kSymB / 1024];  // 16-byte loads, all in flight
#pragma unroll
  for (int u = 0; u < kSymB * kSymB / 1024; u++) tv[u] = tq[threadIdx.x + 256 * u];
#pragma unroll
  for (int u = 0; u < kSymB * kSymB / 1024; u++) {
    const int e = 4 * (threadIdx.x + 256 * u), a = e / kSymB, b = e % kSymB;
    tl[a][b] = tv[u].x;
    tl[a][b + 1] = tv[u].y;
    tl[a][b + 2] = tv[u].z;
    tl[a][b + 3] = tv[u].w;
  }
  if (threadIdx.x < kSymB) {
    const int i = I * kSymB + threadIdx.x;
    rI[threadIdx.x] = i < m ? (double)r[cidx[i]] : 0.0;
  } else if (threadIdx.x < 2 * kSymB) {
    const int j = J * kSymB + threadIdx.x - kSymB;
    rJ[threadIdx.x - kSymB] = j < m ? (double)r[cidx[j]] : 0.0;
  }
  __syncthreads();
  const int a = threadIdx.x >> 2, c = threadIdx.x & 3;
  double sr = 0.0;
#pragma unroll
  for (int b = 16 * c; b < 16 * c + 16; b++) sr += (double)tl[a][b] * rJ[b];
  sr += __shfl_xor(sr, 1, 64);
  sr += __shfl_xor(sr, 2, 64);
  if (c == 0) part[((long)I * ntile + J) * kSymB + a] = sr;
  if (I < J) {
    double sc = 0.0;
#pragma unroll
    for (int b = 16 * c; b < 16 * c + 16; b++) sc += (double)tl[b][a] * rI[b];
    sc += __shfl_xor(sc, 1, 64);
    sc += __shfl_xor(sc, 2, 64);
    if (c == 0) part[((long)J * ntile + I) * kSymB + a] = sc;
  }
}

// pass 2: row i of block row I sums part[I][0..ntile)[i mod 64] in order and updates x at its
// DOF; the blocks after the dense rows take the trivial DOFs (x_k = r_k)
template <class T>
__global__ __launch_bounds__(256) void k_coarse_sym2(int m, int ntile,
                                                     const double *__restrict__ part,
                                                     const int *__restrict__ cidx, int nt,
                                                     const int *__restrict__ tidx,
                                                     const T *__restrict__ r, T *__restrict__ x,
                                                     int x_zero) {
  const int nbd = (m + 255) / 256;
  if ((int)blockIdx.x >= nbd) {
    const int t = ((int)blockIdx.x - nbd) * 256 + (int)threadIdx.x;
    if (t < nt) {
      const int k = tidx[t];
      const T v = r[k];
      x[k] = x_zero ? v : x[k] + v;
    }
    return;
  }
  const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (i >= m) return;
  const double *pi = part + (long)(i / kSymB) * ntile * kSymB + i % kSymB;
  // 16 partials per batch, their loads all in flight (the index clamped), summed in order
  double acc = 0.0;
  for (int J0 = 0; J0 < ntile; J0 += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = pi[(long)min(J0 + u, ntile - 1) * kSymB];
#pragma unroll
    for (int u = 0; u < 16; u++) acc = J0 + u < ntile ? acc + v[u] : acc;
  }
  const int k = cidx[i];
  const T v = (T)acc;
  x[k] = x_zero ? v : x[k] + v;
}

// x = A^-1 r (x_zero) or x += A^-1 r
static void dense_inverse_apply(hipStream_t s, const DenseInverse &D, int prec, const void *r,
                                void *x, int x_zero) {
  if (D.tiles) {
    FLGPU_REQUIRE(prec == FLGPU_F32, "the FP32 tiles of the inverse are for FP32 vectors");
    const int nq = D.ntile * (D.ntile + 1) / 2;
    k_coarse_sym1<float><<<nq, 256, 0, s>>>(D.m, D.ntile, D.tiles, D.tI, D.tJ, D.cidx,
                                            (const float *)r, D.part);
    const unsigned g2 = (unsigned)((D.m + 255) / 256 + (D.nt + 255) / 256);
    k_coarse_sym2<float><<<g2, 256, 0, s>>>(D.m, D.ntile, D.part, D.cidx, D.nt, D.tidx,
                                            (const float *)r, (float *)x, x_zero);
    FLGPU_HIP(hipGetLastError());
    return;
  }
  const unsigned g = (unsigned)((D.m + 3) / 4 + (D.nt + 255) / 256);
  if (g == 0) return;
  const size_t lds = sizeof(double) * (size_t)D.m;  // m <= 8192: at most 64 KiB
  // one or two batches of loads per row: 64 registers of the inverse per lane
  if (prec == FLGPU_F64) {
    k_coarse_solve<double, double, 32><<<g, 256, lds, s>>>(D.m, D.inv, D.cidx, D.nt, D.tidx,
                                                           (const double *)r, (double *)x,
                                                           x_zero);
  } else {
    k_coarse_solve<float, double, 32><<<g, 256, lds, s>>>(D.m, D.inv, D.cidx, D.nt, D.tidx,
                                                          (const float *)r, (float *)x,
                                                          x_zero);
  }
  FLGPU_HIP(hipGetLastError());
}

// tile q = (I, J) of the upper triangle, rounded to FP32 (zero beyond m)
__global__ void k_pack_sym(int m, const double *__restrict__ inv, const int *__restrict__ tI,
                           const int *__restrict__ tJ, float *__restrict__ tiles) {
  const int q = blockIdx.x;
  const int I = tI[q], J = tJ[q];
  for (int e = threadIdx.x; e < kSymB * kSymB; e += blockDim.x) {
    const int i = I * kSymB + e / kSymB, j = J * kSymB + e % kSymB;
    tiles[(long)q * kSymB * kSymB + e] = (i < m && j < m) ? (float)inv[(long)i * m + j] : 0.0f;
  }
}

// the base level's inverse, computed in FP64; an FP32 cycle applies it rounded to FP32 from
// the tiles of its upper triangle (its products summed in FP64): a quarter of the bytes
static void coarse_setup(flgpu_mg *mg, MGLevel &L) {
  FLGPU_REQUIRE(L.A->nrows <= 8192,
                "base level too large for the dense coarse solve (> 8192 DOFs)");
  L.cinv = dense_inverse_dev(L.A);
  DenseInverse &D = L.cinv;
  if (mg->p.prec == FLGPU_F32 && D.m > 0) {
    hipStream_t s = mg->ctx->stream;
    D.ntile = (D.m + kSymB - 1) / kSymB;
    std::vector<int> hI, hJ;
    for (int I = 0; I < D.ntile; I++)
      for (int J = I; J < D.ntile; J++) {
        hI.push_back(I);
        hJ.push_back(J);
      }
    const int nq = (int)hI.size();
    D.tI = (int *)dev_alloc(sizeof(int) * nq);
    D.tJ = (int *)dev_alloc(sizeof(int) * nq);
    D.tiles = (float *)dev_alloc(sizeof(float) * (size_t)nq * kSymB * kSymB);
    D.part = (double *)dev_alloc(sizeof(double) * (size_t)D.ntile * D.ntile * kSymB);
    FLGPU_HIP(hipMemcpyAsync(D.tI, hI.data(), sizeof(int) * nq, hipMemcpyHostToDevice, s));
    FLGPU_HIP(hipMemcpyAsync(D.tJ, hJ.data(), sizeof(int) * nq, hipMemcpyHostToDevice, s));
    k_pack_sym<<<nq, 256, 0, s>>>(D.m, D.inv, D.tI, D.tJ, D.tiles);
    FLGPU_HIP(hipGetLastError());
    FLGPU_HIP(hipStreamSynchronize(s));
    dev_free(D.inv);
    D.inv = nullptr;
  }
}

static void coarse_apply(flgpu_mg *mg, MGLevel &L, void *x, void *r) {
  dense_inverse_apply(mg->ctx->stream, L.cinv, mg->p.prec, r, x, L.x_zero);
}

// ---- V-cycle ---------------------------------------------------------------------------------

static inline void *lsol(flgpu_mg *mg, int l) { return l == mg->top ? mg->sol_top : mg->lv[l].sol; }
static inline void *lrhs(flgpu_mg *mg, int l) { return l == mg->top ? mg->rhs_top : mg->lv[l].rhs; }
static inline void *lres(flgpu_mg *mg, int l) { return l == mg->top ? mg->res_top : mg->lv[l].res; }

// ensure-mg-residual (multigrid.lisp:173-177)
static void ensure_residual(flgpu_mg *mg, int l) {
  MGLevel &L = mg->lv[l];
  if (L.residual_p) return;
  spmv(L.A, -1.0, lsol(mg, l), 1.0, lrhs(mg, l), lres(mg, l));
  L.residual_p = true;
}

// smooth (multigrid.lisp:179-189) with the product-iterator of the smoother.  last: the
// cycle's final operation (top level, post-smoothing, multiplicative)
static void smooth(flgpu_mg *mg, int l, bool post, bool last = false) {
  const flgpu_mg_params &p = mg->p;
  int steps = post ? p.post_steps : p.pre_steps;
  if (steps <= 0) return;
  MGLevel &L = mg->lv[l];
  void *x = lsol(mg, l), *b = lrhs(mg, l), *r = lres(mg, l);
  if (p.smoother == FLGPU_SMOOTH_JACOBI && b != r && L.tmp) {
    // residual recomputed inside each step; only femlisp's aliased rhs = residual (below)
    // needs the in-place form that divides the residual vector itself
    bool w = jacobi_steps(L.A, p.omega, steps, x, b, L.tmp, L.x_zero, L.step1 ? 1 : 0,
                          last ? mg->x64 : nullptr);
    if (last) mg->x64_done = w;
    L.x_zero = false;
    L.step1 = false;
  } else if (p.smoother == FLGPU_SMOOTH_JACOBI) {
    ensure_residual(mg, l);  // <jacobi> has residual-before t
    for (int k = 0; k < steps; k++) {
      jacobi_iterate(L.A, p.omega, x, r);
      if (steps > 1) spmv(L.A, -1.0, x, 1.0, b, r);  // residual-before and not residual-after
    }
  } else if (p.smoother == FLGPU_SMOOTH_SSC || p.smoother == FLGPU_SMOOTH_PSC) {
    // <block-iteration> has residual-before t and the product-iterator recomputes the residual
    // after each of several steps (linit.lisp:116-119).  <ssc> overwrites the residual of every
    // block DOF from b, so those computations are seen only when b is r (femlisp's aliased
    // call); <psc> reads them.
    bool ssc = p.smoother == FLGPU_SMOOTH_SSC;
    bool seen = !ssc || b == r;
    if (seen) ensure_residual(mg, l);
    for (int k = 0; k < steps; k++) {
      block_sweep(L.A, ssc, p.omega, post && p.symmetric, x, b, r);
      if (steps > 1 && seen) spmv(L.A, -1.0, x, 1.0, b, r);
    }
  } else {
    sor_sweeps(L.A, p.omega, steps, post && p.symmetric, x, b, r);
  }
  L.residual_p = false;  // residual-after nil
}

static void lmgc(flgpu_mg *mg, int l) {
  MGLevel &L = mg->lv[l];
  if (l == mg->p.base_level) {
    coarse_apply(mg, L, lsol(mg, l), lres(mg, l));
    L.x_zero = false;
    L.residual_p = false;
    return;
  }
  // :multiplicative smooths before restricting; :additive restricts the residual first and
  // smooths the same level from it (multigrid.lisp:262-269)
  const bool add = mg->p.additive != 0;
  if (!add) smooth(mg, l, false);
  // restrict (multigrid.lisp:202-224): res_{l-1} = I^T res_l
  ensure_residual(mg, l);
  MGLevel &C = mg->lv[l - 1];
  // :after methods: x_{l-1} = 0, rhs_{l-1} = res_{l-1}, residual-p t.  Two shortcuts keep
  // the values: the coarse solve and fused <jacobi> steps write x instead of updating a
  // stored zero, and fused steps recompute the residual from rhs, so the restriction goes
  // straight into rhs there (the coarse solve reads only res)
  bool base = (l - 1 == mg->p.base_level);
  bool fused = !add && mg->p.smoother == FLGPU_SMOOTH_JACOBI && C.tmp && mg->p.pre_steps > 0;
  if (mg->p.fas) {
    // <fas> :after (multigrid.lisp:226-232): rhs_{l-1} = res_{l-1}, sol_{l-1} = R sol_l,
    // rhs_{l-1} += A_{l-1} sol_{l-1}, residual-p t
    spmv(C.I->transpose, 1.0, lres(mg, l), 0.0, C.res, C.res);
    vcopy(mg->ctx, C.N, mg->p.prec, C.res, mg->p.prec, C.rhs);
    spmv(C.R, 1.0, lsol(mg, l), 0.0, C.sol, C.sol);
    spmv(C.A, 1.0, C.sol, 1.0, C.rhs, C.rhs);
    C.x_zero = false;
    C.step1 = false;
    C.residual_p = true;
  } else if (fused && !base) {
    // the first coarse <jacobi> step from zero is pointwise in rhs: taken in the same pass
    // where the transfer allows it
    C.step1 = restrict_first_step(C.I->transpose, lres(mg, l), C.rhs, C.A, mg->p.omega, C.tmp);
    if (!C.step1) spmv(C.I->transpose, 1.0, lres(mg, l), 0.0, C.rhs, C.rhs);
    C.x_zero = true;
    C.residual_p = false;
  } else {
    spmv(C.I->transpose, 1.0, lres(mg, l), 0.0, C.res, C.res);
    if (base) {
      C.x_zero = true;
    } else {
      vfill(mg->ctx, C.N, mg->p.prec, C.sol, 0.0);
      vcopy(mg->ctx, C.N, mg->p.prec, C.res, mg->p.prec, C.rhs);
    }
    C.residual_p = true;
  }
  if (add) smooth(mg, l, false);
  // gamma coarse cycles, one when the coarse level is the base level, none for gamma 0
  // (multigrid.lisp:268-273).  femlisp brings the coarse residual up to date between them;
  // here every path that reads it recomputes it first (the smoothers' residual-before, the
  // restriction's ensure-mg-residual), so that step is left out.
  int reps = mg->p.gamma == 0 ? 0 : (base ? 1 : mg->p.gamma);
  for (int i = 0; i < reps; i++) lmgc(mg, l - 1);
  if (reps == 0) {  // the coarse correction is zero: sol_{l-1} must hold it
    if (C.x_zero) vfill(mg->ctx, C.N, mg->p.prec, C.sol, 0.0);
    C.x_zero = false;
    C.step1 = false;
  }
  if (add) smooth(mg, l, true);  // :additive post-smooths before the prolongation
  if (mg->p.fas) {
    // <fas> prolongate (multigrid.lisp:241-247): res_{l-1} = sol_{l-1} - R sol_l,
    // sol_l += I res_{l-1}
    spmv(C.R, -1.0, lsol(mg, l), 1.0, C.sol, C.res);
    spmv(C.I, 1.0, C.res, 1.0, lsol(mg, l), lsol(mg, l));
  } else {
    // prolongate (multigrid.lisp:237-252): sol_l += I sol_{l-1}
    spmv(C.I, 1.0, C.sol, 1.0, lsol(mg, l), lsol(mg, l));
  }
  L.residual_p = false;
  if (!add) smooth(mg, l, true, l == mg->top);
}

// the <mg-iteration> iterator on device pointers of the hierarchy's precision
static void mg_cycle(flgpu_mg *mg, void *x, void *b, void *r) {
  mg->sol_top = x;
  mg->rhs_top = b;
  mg->res_top = r;
  mg->lv[mg->top].residual_p = true;
  lmgc(mg, mg->top);
}

template <class S, class D>
__global__ void k_add_conv(long n, const S *x, D *y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += (D)x[i];
}

// y += x with precision conversion
static void vadd_converted(flgpu_ctx *ctx, long n, int sp, const void *src, int dp, void *dst) {
  unsigned g = grid_for(n, 256);
  if (sp == FLGPU_F32 && dp == FLGPU_F64)
    k_add_conv<float, double><<<g, 256, 0, ctx->stream>>>(n, (const float *)src, (double *)dst);
  else if (sp == FLGPU_F64 && dp == FLGPU_F32)
    k_add_conv<double, float><<<g, 256, 0, ctx->stream>>>(n, (const double *)src, (float *)dst);
  else if (sp == FLGPU_F64)
    k_add_conv<double, double><<<g, 256, 0, ctx->stream>>>(n, (const double *)src, (double *)dst);
  else
    k_add_conv<float, float><<<g, 256, 0, ctx->stream>>>(n, (const float *)src, (float *)dst);
  FLGPU_HIP(hipGetLastError());
}

// x <- x + V(r) (femlisp's iterate on (x, b, r)); vectors of precision vprec
void mg_apply_ptr(flgpu_mg *mg, int vprec, void *x, void *b, void *r) {
  if (vprec == mg->p.prec) {
    mg_cycle(mg, x, b, r);
    return;
  }
  // cycle in the hierarchy's precision on staged copies: the correction
  // scheme solves A e = r from e = 0 and adds e to x
  FLGPU_REQUIRE(b != r, "a mixed-precision cycle needs separate rhs and residual vectors");
  long N = mg->lv[mg->top].N;
  vfill(mg->ctx, N, mg->p.prec, mg->xs, 0.0);
  vcopy(mg->ctx, N, vprec, r, mg->p.prec, mg->rs);
  vcopy(mg->ctx, N, mg->p.prec, mg->rs, mg->p.prec, mg->bs);
  mg_cycle(mg, mg->xs, mg->bs, mg->rs);
  vadd_converted(mg->ctx, N, mg->p.prec, mg->xs, vprec, x);
}

// The cycle's launch sequence depends only on the parameters (the residual-p
// bookkeeping restarts at every cycle), so it is captured once per vector
// triple into a HIP graph and replayed: one launch instead of ~40 per level.
// The cycle clears with fill kernels, never with memset nodes.
static void graphs_clear(flgpu_mg *mg) {
  for (auto &g : mg->graphs) (void)hipGraphExecDestroy(g.exec);
  mg->graphs.clear();
}

// <cg>'s preconditioner call (copy! r w) (x<-0 q) (iterate q w w2) without its copies: the
// cycle never writes its top-level rhs, so it reads r in place (converted once for a cycle of
// another precision), and q is written instead of cleared and updated.  work (n values of
// vprec) is the top-level residual; it is seeded from r only when the first smoothing step
// would read it (an in-place <jacobi> iterate, or no pre-smoothing).
// true when an FP32 (or FP64) cycle for vectors of the other precision stages r by the pass
// that also takes the first top-level <jacobi> step (stage_first_step): then a caller that
// produces r anyway may write that pass's two outputs itself (mg_run mode 2)
static bool mg_stages_first(const flgpu_mg *mg, int vprec) {
  const flgpu_mg_params &p = mg->p;
  const MGLevel &T = mg->lv[mg->top];
  bool base_top = mg->top == p.base_level;
  bool jac = p.smoother == FLGPU_SMOOTH_JACOBI && T.tmp && p.pre_steps > 0 && !base_top;
  return vprec != p.prec && jac && p.additive == 0;
}

// staged: the caller has already written bs (r converted) and the top level's tmp (the first
// <jacobi> step from zero) as stage_first_step would (only when mg_stages_first)
// keep32 (an FP32 cycle under FP64 vectors): V(b) is left in FP32 in mg->xs, q not written
static void mg_apply_zero(flgpu_mg *mg, int vprec, void *q, void *r, void *work,
                          bool staged = false, bool keep32 = false) {
  const flgpu_mg_params &p = mg->p;
  MGLevel &T = mg->lv[mg->top];
  long N = T.N;
  bool base_top = mg->top == p.base_level;  // one level: the coarse solve reads r
  bool jac = p.smoother == FLGPU_SMOOTH_JACOBI && T.tmp && p.pre_steps > 0 && !base_top;
  // :additive restricts before smoothing: the top residual is read and x must be stored
  bool add = p.additive != 0;
  bool sweeps = p.smoother == FLGPU_SMOOTH_SOR || p.smoother == FLGPU_SMOOTH_SSC;
  bool reads_r = base_top || add || !(jac || (sweeps && p.pre_steps > 0));
  bool zero = (jac && !add) || base_top;  // the first step writes x
  FLGPU_REQUIRE(!staged || mg_stages_first(mg, vprec), "this cycle does not take a staged rhs");
  if (vprec == p.prec) {
    if (zero) T.x_zero = true;
    else vfill(mg->ctx, N, vprec, q, 0.0);
    if (reads_r) vcopy(mg->ctx, N, vprec, r, vprec, work);
    mg_cycle(mg, q, r, work);
    return;
  }
  if (zero && jac && !reads_r) {
    // the staging pass also takes the first top-level <jacobi> step from zero
    if (!staged) stage_first_step(T.A, vprec, r, mg->bs, p.omega, T.tmp);
    T.step1 = true;
  } else {
    vcopy(mg->ctx, N, vprec, r, p.prec, mg->bs);
  }
  if (zero) T.x_zero = true;
  else vfill(mg->ctx, N, p.prec, mg->xs, 0.0);
  if (reads_r) vcopy(mg->ctx, N, p.prec, mg->bs, p.prec, mg->rs);
  mg->x64 = (vprec == FLGPU_F64 && !keep32) ? (double *)q : nullptr;
  mg->x64_done = false;
  mg_cycle(mg, mg->xs, mg->bs, mg->rs);
  if (!keep32 && !mg->x64_done) vcopy(mg->ctx, N, p.prec, mg->xs, vprec, q);
  mg->x64 = nullptr;
  mg->x64_done = false;
}

static void mg_run(flgpu_mg *mg, int mode, int vprec, void *x, void *b, void *r) {
  if (mode >= 1 && mode <= 4) mg_apply_zero(mg, vprec, x, b, r, mode == 2 || mode == 4, mode >= 3);
  else mg_apply_ptr(mg, vprec, x, b, r);
}

// mode 0: x <- x + V(r) on (x, b, r); mode 1: x <- V(b) from zero, b read only, r work; mode
// 2: mode 1 with b already staged (mg_apply_zero's `staged`); modes 3 and 4: modes 1 and 2
// with V(b) left in FP32 in mg->xs (mg_apply_zero's `keep32`).
// With graphs on, the first call on a (mode, x, b, r, precision) is captured and replayed.
// Sum of the layout counts of every matrix a cycle reads.  The counts only grow, so the sum
// moves whenever one of them does: a level matrix handed out by flgpu_mg_level_matrix whose
// form, colouring or buffers the caller changed.
static unsigned long mg_layout(const flgpu_mg *mg) {
  unsigned long sum = 0;
  auto add = [&](const flgpu_mat *M) {
    if (!M) return;
    sum += M->layout;
    if (M->transpose) sum += M->transpose->layout;
  };
  for (const MGLevel &L : mg->lv) {
    add(L.A);
    add(L.I);
    add(L.R);
  }
  return sum;
}

static void mg_run_cached(flgpu_mg *mg, int mode, int vprec, void *x, void *b, void *r) {
  if (!mg->use_graphs) {
    mg_run(mg, mode, vprec, x, b, r);
    return;
  }
  hipStream_t s = mg->ctx->stream;
  unsigned long lay = mg_layout(mg);
  if (lay != mg->graph_layout) {  // captured cycles may address freed or replaced buffers
    FLGPU_HIP(hipStreamSynchronize(s));
    graphs_clear(mg);
    mg->graph_layout = lay;
  }
  for (auto &g : mg->graphs)
    if (g.x == x && g.b == b && g.r == r && g.vprec == vprec && g.mode == mode) {
      FLGPU_HIP(hipGraphLaunch(g.exec, s));
      return;
    }
  if (mg->graphs.size() >= 8) {  // bounded cache: drop the oldest
    FLGPU_HIP(hipStreamSynchronize(s));
    (void)hipGraphExecDestroy(mg->graphs.front().exec);
    mg->graphs.erase(mg->graphs.begin());
  }
  hipGraph_t graph = nullptr;
  FLGPU_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    mg_run(mg, mode, vprec, x, b, r);
  } catch (...) {
    (void)hipStreamEndCapture(s, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  FLGPU_HIP(hipStreamEndCapture(s, &graph));
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  FLGPU_HIP(e);
  mg->graphs.push_back({x, b, r, vprec, mode, exec});
  FLGPU_HIP(hipGraphLaunch(exec, s));
}

void mg_apply_cached(flgpu_mg *mg, int vprec, void *x, void *b, void *r) {
  mg_run_cached(mg, 0, vprec, x, b, r);
}

static bool block_smoother(const flgpu_mg_params &p) {
  return p.smoother == FLGPU_SMOOTH_SSC || p.smoother == FLGPU_SMOOTH_PSC;
}

static void check_params(const flgpu_mg_params *p, int nlevels) {
  FLGPU_REQUIRE(p != nullptr, "params must not be NULL");
  FLGPU_REQUIRE(p->smoother == FLGPU_SMOOTH_JACOBI || p->smoother == FLGPU_SMOOTH_SOR ||
                    p->smoother == FLGPU_SMOOTH_SSC || p->smoother == FLGPU_SMOOTH_PSC,
                "unknown smoother");
  FLGPU_REQUIRE(p->pre_steps >= 0 && p->post_steps >= 0, "negative smoothing steps");
  FLGPU_REQUIRE(p->base_level >= 0 && p->base_level < nlevels, "base_level out of range");
  FLGPU_REQUIRE(p->prec == FLGPU_F64 || p->prec == FLGPU_F32, "unknown precision");
  FLGPU_REQUIRE(p->transfer == FLGPU_TRANSFER_GEOMETRIC || p->transfer == FLGPU_TRANSFER_BLACKBOX,
                "unknown transfer");
  FLGPU_REQUIRE(p->gamma >= 0 && p->gamma <= 16, "gamma must be in 0..16");
  FLGPU_REQUIRE(p->additive == 0 || p->additive == 1, "additive must be 0 or 1");
  FLGPU_REQUIRE(p->fas == 0 || p->fas == 1, "fas must be 0 or 1");
}

static void mg_alloc_work(flgpu_mg *mg) {
  for (auto &L : mg->lv) {
    L.N = L.A->nrows;
    size_t b = prec_bytes(mg->p.prec) * L.N;
    L.sol = dev_alloc(b);
    L.rhs = dev_alloc(b);
    L.res = dev_alloc(b);
    vfill(mg->ctx, L.N, mg->p.prec, L.sol, 0.0);
    vfill(mg->ctx, L.N, mg->p.prec, L.rhs, 0.0);
    vfill(mg->ctx, L.N, mg->p.prec, L.res, 0.0);
    if (mg->p.smoother == FLGPU_SMOOTH_JACOBI) L.tmp = dev_alloc(b);
  }
  long N = mg->lv[mg->top].N;
  size_t b = prec_bytes(mg->p.prec) * N;
  mg->xs = dev_alloc(b);
  mg->bs = dev_alloc(b);
  mg->rs = dev_alloc(b);
  coarse_setup(mg, mg->lv[mg->p.base_level]);
}

void mg_unet(flgpu_mg *mg, flgpu_unet **u, const flgpu_vec **kappa, int *n) {
  *u = mg->unet;
  *kappa = mg->unet_kappa;
  *n = mg->unet_n;
}

int mg_level_cells(const flgpu_mg *mg, int n) {
  if (mg->lat_dim == 0) return -1;
  for (int l = 0; l <= mg->top; l++)
    if ((mg->lat_n0 << l) == n) return l;
  return -1;
}
int mg_top_level(const flgpu_mg *mg) { return mg->top; }
int mg_precision(const flgpu_mg *mg) { return mg->p.prec; }
long mg_level_rows(const flgpu_mg *mg, int l) { return mg->lv[l].N; }

// coarse = I^T fine from level l to l - 1, as the cycle restricts (beta 0)
void mg_restrict(flgpu_mg *mg, int l, const void *fine, void *coarse) {
  const MGLevel &C = mg->lv[l - 1];
  spmv(C.I->transpose, 1.0, fine, 0.0, coarse, coarse);
}

// fine = I coarse from level l - 1 to l (beta 0)
void mg_prolong(flgpu_mg *mg, int l, const void *coarse, void *fine) {
  const MGLevel &C = mg->lv[l - 1];
  spmv(C.I, 1.0, coarse, 0.0, fine, fine);
}

void mg_free(flgpu_mg *mg) {
  if (!mg) return;
  graphs_clear(mg);
  for (auto &L : mg->lv) {
    mat_free(L.A);
    mat_free(L.I);
    mat_free(L.R);
    dev_free(L.sol);
    dev_free(L.rhs);
    dev_free(L.res);
    dev_free(L.tmp);
    L.cinv.release();
  }
  dev_free(mg->xs);
  dev_free(mg->bs);
  dev_free(mg->rs);
  delete mg;
}

// ---- Krylov kernels with device scalars ----------------------------------------------------

// p <- (a/b) p + q   (<cg>: scal! (/ new-alpha alpha) p, m+! q p)
template <class T>
__global__ void k_cg_p(long n, const double *a, const double *b, T *p, const T *__restrict__ q) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T sc = (T)(*a / *b);
  p[i] = sc * p[i] + q[i];
}

// p <- q + beta p, beta = -(q.a)/(a.p) (flexible CG): Polak-Ribiere's q.(r - r_old)/(q_old.r_old)
// with r - r_old = -lam a and lam = (q_old.r_old)/(a.p), so that r_old need not be kept
template <class T, class TQ = T>
__global__ void k_fcg_p(long n, const double *qa, const double *ap, T *p,
                        const TQ *__restrict__ q) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T beta = (T)(-*qa / *ap);
  p[i] = (T)q[i] + beta * p[i];
}

// <cg>'s update in one pass: x <- x + lam p, r <- r - lam a with lam = num / den rounded to T
// once per sign (femlisp's axpy! with (/ alpha beta))
// STAGE: also the next preconditioner call's staging pass on the new r (stage_first_step, for
// an FP32 hierarchy under FP64 vectors): bs = (float)r, xo = damp (bs / diag), the values that
// pass would write
template <class T, bool STAGE = false>
__global__ void k_cg_update(long n, const double *num, const double *den, const T *__restrict__ p,
                            const T *__restrict__ a, T *__restrict__ x, T *__restrict__ r,
                            float *__restrict__ bs = nullptr,
                            const float *__restrict__ diag = nullptr, float damp = 0.0f,
                            float *__restrict__ xo = nullptr) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double q = *num / *den;
  T lp = (T)(1.0 * q), lm = (T)(-1.0 * q);
  T ri = r[i];
  x[i] = lp * p[i] + x[i];
  T rn = lm * a[i] + ri;
  r[i] = rn;
  if (STAGE) {
    float bi = (float)rn;
    bs[i] = bi;
    xo[i] = damp * (bi / diag[i]);
  }
}

}  // namespace flgpu

using namespace flgpu;

// ---- C-ABI: hierarchy ---------------------------------------------------------------------

extern "C" int flgpu_mg_create(flgpu_ctx *ctx, int nlevels, flgpu_mat *const *A,
                               flgpu_mat *const *I, const flgpu_mg_params *params,
                               flgpu_mg **out) {
  return guard([&] {
    FLGPU_REQUIRE(ctx && A && out, "ctx, A and out must not be NULL");
    FLGPU_REQUIRE(nlevels >= 1, "nlevels must be >= 1");
    FLGPU_REQUIRE(nlevels == 1 || I, "I must not be NULL");
    check_params(params, nlevels);
    FLGPU_REQUIRE(!params->galerkin && params->transfer == FLGPU_TRANSFER_GEOMETRIC,
                  "galerkin = 1 and black-box transfers are for lattice hierarchies; build "
                  "coarse matrices with flgpu_galerkin_product");
    for (int l = 0; l < nlevels; l++) {
      FLGPU_REQUIRE(A[l] && A[l]->ctx == ctx, "level matrix missing or of another context");
      FLGPU_REQUIRE(A[l]->nrows == A[l]->ncols, "level matrices must be square");
      if (l + 1 < nlevels) {
        FLGPU_REQUIRE(I[l] && I[l]->ctx == ctx, "interpolation missing or of another context");
        FLGPU_REQUIRE(A[l + 1] && I[l]->nrows == A[l + 1]->nrows && I[l]->ncols == A[l]->nrows,
                      "interpolation dimensions do not match the level matrices");
      }
    }
    FLGPU_REQUIRE(!params->fas,
                  "FAS needs the projection matrices of a lattice hierarchy "
                  "(flgpu_mg_create_lattice)");
    CTX_LOCK(ctx);
    flgpu_mg *mg = new flgpu_mg;
    mg->ctx = ctx;
    mg->p = *params;
    mg->top = nlevels - 1;
    mg->lv.resize(nlevels);
    try {
      for (int l = 0; l < nlevels; l++) {
        mg->lv[l].A = mat_convert(A[l], params->prec);
        if (params->smoother == FLGPU_SMOOTH_SOR && mg->lv[l].A->ncolours == 0 &&
            l > params->base_level) {
          int nc = 0;
          FLGPU_REQUIRE(flgpu_mat_colour_greedy(mg->lv[l].A, &nc) == 0, "greedy colouring failed");
        }
        // block smoothers find their blocks from the lattice (blocks_prepare checks it)
        if (block_smoother(*params) && l > params->base_level) blocks_prepare(mg->lv[l].A);
        if (l + 1 < nlevels) {
          mg->lv[l].I = mat_convert(I[l], params->prec);
          mat_build_transpose(mg->lv[l].I);
        }
      }
      mg_alloc_work(mg);
    } catch (...) {
      mg_free(mg);
      throw;
    }
    *out = mg;
    return 0;
  });
}

extern "C" int flgpu_mg_create_lattice(flgpu_ctx *ctx, int dim, int order, int n0, int nlevels,
                                       const flgpu_vec *kappa, int nk, double f, int asm_mode,
                                       const flgpu_mg_params *params, flgpu_mg **out,
                                       flgpu_vec **rhs) {
  return guard([&] {
    FLGPU_REQUIRE(ctx && out, "ctx and out must not be NULL");
    FLGPU_REQUIRE(nlevels >= 1 && n0 >= 1, "nlevels and n0 must be >= 1");
    check_params(params, nlevels);
    FLGPU_REQUIRE(!kappa || kappa->ctx == ctx, "kappa belongs to another context");
    FLGPU_REQUIRE(params->transfer == FLGPU_TRANSFER_GEOMETRIC || (dim == 3 && order == 1),
                  "black-box transfers are defined for 3D Q1 lattices");
    CTX_LOCK(ctx);
    flgpu_mg *mg = new flgpu_mg;
    mg->ctx = ctx;
    mg->p = *params;
    mg->top = nlevels - 1;
    mg->lv.resize(nlevels);
    mg->lat_dim = dim;
    mg->lat_order = order;
    mg->lat_n0 = n0;
    mg->asm_mode = asm_mode;
    flgpu_vec *b = nullptr;
    try {
      for (int l = 0; l < nlevels; l++) {
        int n = n0 << l;
        bool is_top = (l == nlevels - 1);
        if (!is_top) mg->lv[l].I = lattice_interpolation(ctx, dim, order, n, params->prec);
        if (!is_top && params->fas)
          mg->lv[l].R = lattice_projection(ctx, dim, order, n, params->prec);
      }
      for (int l = nlevels - 1; l >= 0; l--) {
        int n = n0 << l;
        bool is_top = (l == nlevels - 1);
        bool bbox = params->transfer == FLGPU_TRANSFER_BLACKBOX;
        if (is_top || (!params->galerkin && !bbox)) {
          mg->lv[l].A = lattice_assemble(ctx, dim, order, n, kappa, nk, f, params->prec, asm_mode,
                                         (is_top && rhs) ? &b : nullptr);
        } else {
          if (bbox) lattice_blackbox(mg->lv[l + 1].A, mg->lv[l].I);
          mg->lv[l].A = lattice_galerkin(mg->lv[l].I, mg->lv[l + 1].A, dim, order, n);
        }
        // matrices in the per-cell form sweep the lattice colours directly
        if (params->smoother == FLGPU_SMOOTH_SOR && l > params->base_level &&
            !mg->lv[l].A->alpha && !mg->lv[l].A->val27)
          lattice_colouring(mg->lv[l].A, dim, order, n);
        if (block_smoother(*params) && l > params->base_level) blocks_prepare(mg->lv[l].A);
      }
      mg_alloc_work(mg);
    } catch (...) {
      vec_free(b);
      mg_free(mg);
      throw;
    }
    if (rhs) *rhs = b;
    *out = mg;
    return 0;
  });
}

extern "C" int flgpu_mg_reassemble(flgpu_mg *mg, const flgpu_vec *kappa, int nk) {
  return guard([&] {
    FLGPU_REQUIRE(mg, "hierarchy must not be NULL");
    FLGPU_REQUIRE(mg->lat_dim > 0, "the hierarchy was not assembled on a lattice");
    FLGPU_REQUIRE(!kappa || kappa->ctx == mg->ctx, "kappa belongs to another context");
    CTX_LOCK(mg->ctx);
    for (int l = mg->top; l >= 0; l--) {
      MGLevel &L = mg->lv[l];
      bool bbox = mg->p.transfer == FLGPU_TRANSFER_BLACKBOX;
      if (l == mg->top || (!mg->p.galerkin && !bbox)) {
        lattice_reassemble(L.A, kappa, nk, mg->asm_mode);
      } else {
        if (bbox) lattice_blackbox(mg->lv[l + 1].A, L.I);
        galerkin_product(L.I, mg->lv[l + 1].A, L.A, true);
        lattice_dirichlet_identity(L.A, mg->lat_dim, mg->lat_order, mg->lat_n0 << l);
        mat_refresh_values(L.A);
      }
    }
    graphs_clear(mg);  // the coarse inverse is reallocated below
    MGLevel &B = mg->lv[mg->p.base_level];
    B.cinv.release();
    coarse_setup(mg, B);
    FLGPU_HIP(hipStreamSynchronize(mg->ctx->stream));
    return 0;
  });
}

extern "C" int flgpu_mg_attach_unet(flgpu_mg *mg, flgpu_unet *u, const flgpu_vec *kappa, int n) {
  return guard([&] {
    FLGPU_REQUIRE(mg, "hierarchy must not be NULL");
    FLGPU_REQUIRE(mg->lat_dim == 3 && mg->lat_order == 1,
                  "the learned correction is defined for 3D Q1 lattices");
    FLGPU_REQUIRE(!u || mg_level_cells(mg, n) >= 0, "n does not match a level of the hierarchy");
    FLGPU_REQUIRE(!u || n % 4 == 0, "the network grid needs n divisible by 4");
    FLGPU_REQUIRE(!kappa || (kappa->prec == FLGPU_F64 && kappa->n == (long)n * n * n),
                  "kappa must be an FP64 vector of n^3 cell values");
    FLGPU_REQUIRE(!u || unet_ctx(u) == mg->ctx, "the network belongs to another context");
    FLGPU_REQUIRE(!kappa || kappa->ctx == mg->ctx, "kappa belongs to another context");
    CTX_LOCK(mg->ctx);
    mg->hook.reset();  // the next learned solve builds the hook for this network
    mg->hook_A = nullptr;
    mg->hook_N = 0;
    mg->unet = u;
    mg->unet_kappa = kappa;
    mg->unet_n = n;
    return 0;
  });
}

extern "C" int flgpu_mg_destroy(flgpu_mg *mg) {
  return guard([&] {
    if (!mg) return 0;
    CTX_LOCK(mg->ctx);
    FLGPU_HIP(hipStreamSynchronize(mg->ctx->stream));
    mg_free(mg);
    return 0;
  });
}

extern "C" int flgpu_mg_top_matrix(flgpu_mg *mg, flgpu_mat **A) {
  return guard([&] {
    FLGPU_REQUIRE(mg && A, "arguments must not be NULL");
    *A = mg->lv[mg->top].A;
    return 0;
  });
}

extern "C" int flgpu_mg_level_matrix(flgpu_mg *mg, int level, int interp, flgpu_mat **A) {
  return guard([&] {
    FLGPU_REQUIRE(mg && A, "arguments must not be NULL");
    FLGPU_REQUIRE(level >= 0 && level <= mg->top, "level out of range");
    FLGPU_REQUIRE(interp >= 0 && interp <= 2, "interp must be 0, 1 or 2");
    FLGPU_REQUIRE(!interp || level < mg->top, "the top level has no interpolation");
    FLGPU_REQUIRE(interp != 2 || mg->lv[level].R, "no projection: the hierarchy is not FAS");
    *A = interp == 2 ? mg->lv[level].R : interp ? mg->lv[level].I : mg->lv[level].A;
    return 0;
  });
}

extern "C" int flgpu_mg_apply(flgpu_mg *mg, flgpu_vec *x, flgpu_vec *b, flgpu_vec *r) {
  return guard([&] {
    FLGPU_REQUIRE(mg && x && b && r, "arguments must not be NULL");
    long N = mg->lv[mg->top].N;
    FLGPU_REQUIRE(x->n == N && b->n == N && r->n == N, "vector length does not match the top level");
    FLGPU_REQUIRE(x->prec == b->prec && b->prec == r->prec, "vectors must share one precision");
    FLGPU_REQUIRE(x->d != r->d && x->d != b->d, "x must differ from b and r");
    FLGPU_REQUIRE(x->ctx == mg->ctx && b->ctx == mg->ctx && r->ctx == mg->ctx,
                  "x, b and r must belong to the hierarchy's context");
    CTX_LOCK(mg->ctx);
    mg_apply_cached(mg, x->prec, x->d, b->d, r->d);
    return 0;
  });
}

extern "C" int flgpu_mg_use_graphs(flgpu_mg *mg, int enable) {
  return guard([&] {
    FLGPU_REQUIRE(mg, "hierarchy must not be NULL");
    CTX_LOCK(mg->ctx);
    FLGPU_HIP(hipStreamSynchronize(mg->ctx->stream));
    graphs_clear(mg);
    mg->use_graphs = enable ? 1 : 0;
    return 0;
  });
}

// ---- C-ABI: solver --------------------------------------------------------------------------

// The learned correction is added only when its energy-optimal step adds at least this fraction
// of the energy of the cycle's own step (decided on the device, within the step), and the
// first step where it does not drops the network for the rest of the solve: the network is
// trained on residuals after one V-cycle, which is what it sees in the first step; a network
// that adds little there is not expected to add more later, and each application costs about
// a cycle.  The decision comes back in reduction slot kLearnedFlag.
#ifndef FLGPU_LEARNED_MIN_GAIN
#define FLGPU_LEARNED_MIN_GAIN 1e-2
#endif
constexpr double kLearnedMinGain = FLGPU_LEARNED_MIN_GAIN;
constexpr int kLearnedFlag = 14;

namespace {

struct SolveWork {
  flgpu_ctx *ctx;
  long N;
  int prec;
  // t: the true residual b - A x when it is checked
  void *r = nullptr, *p = nullptr, *a = nullptr, *w = nullptr, *w2 = nullptr, *q = nullptr,
       *t = nullptr;
  // the vectors are slices of the context's solve buffer, grown when too small and kept for
  // the next solve: no allocation or release per solve once the size is reached
  SolveWork(flgpu_ctx *c, long n, int pr) : ctx(c), N(n), prec(pr) {
    const size_t b = (prec_bytes(pr) * (size_t)std::max(n, 1L) + 255) & ~(size_t)255;
    if (c->solve_bytes < 7 * b) {
      FLGPU_HIP(hipStreamSynchronize(c->stream));
      dev_free(c->solve_buf);
      c->solve_buf = nullptr;
      c->solve_bytes = 0;
      c->solve_buf = dev_alloc(7 * b);
      c->solve_bytes = 7 * b;
    }
    char *base = (char *)c->solve_buf;
    int k = 0;
    for (void **v : {&r, &p, &a, &w, &w2, &q, &t}) {
      *v = base + (k++) * b;
      vfill(c, n, pr, *v, 0.0);
    }
  }
};

}  // namespace

namespace flgpu {

// preconditioner call of <cg> (krylow.lisp:118-121): (copy! r w) (x<-0 q)
// (funcall iterate q w w); alias = 0 passes a second copy as the residual
static void precondition(flgpu_mg *mg, const flgpu_solve_params *p, SolveWork &W, void *r,
                         void *q, LearnedHook *learned, bool staged = false,
                         bool keep32 = false) {
  if (!p->alias) {
    FLGPU_REQUIRE(!keep32 || !learned, "the learned correction reads q in FP64");
    mg_run_cached(mg, keep32 ? (staged ? 4 : 3) : (staged ? 2 : 1), W.prec, q, r, W.w2);
    if (learned) learned->correct(r, q, kLearnedMinGain, kLearnedFlag);
    return;
  }
  // femlisp's literal (iterate q w w): rhs and residual are one vector
  vcopy(W.ctx, W.N, W.prec, r, W.prec, W.w);
  vfill(W.ctx, W.N, W.prec, q, 0.0);
  mg_apply_cached(mg, W.prec, q, W.w, W.w);
  if (learned) learned->correct(r, q, kLearnedMinGain, kLearnedFlag);
}

}  // namespace flgpu

extern "C" int flgpu_solve(flgpu_mg *mg, const flgpu_mat *A, const flgpu_solve_params *p,
                           const flgpu_vec *b, flgpu_vec *x, flgpu_solve_report *report,
                           double *defnorms, int max_defnorms) {
  return guard([&] {
    FLGPU_REQUIRE(A && p && b && x, "A, p, b and x must not be NULL");
    FLGPU_REQUIRE(A->nrows == A->ncols, "the system matrix must be square");
    FLGPU_REQUIRE(b->n == A->nrows && x->n == A->nrows, "vector length mismatch");
    FLGPU_REQUIRE(b->prec == A->prec && x->prec == A->prec, "vector precision differs from A");
    FLGPU_REQUIRE(p->outer >= FLGPU_OUTER_MG && p->outer <= FLGPU_OUTER_FCG, "unknown outer iteration");
    FLGPU_REQUIRE(p->outer == FLGPU_OUTER_CG_PLAIN || mg, "this iteration needs a hierarchy");
    if (mg) FLGPU_REQUIRE(mg->lv[mg->top].N == A->nrows, "hierarchy does not match A");
    FLGPU_REQUIRE(!(p->alias && mg && mg->p.prec != A->prec),
                  "alias mode needs the hierarchy in the precision of A");
    FLGPU_REQUIRE(!p->learned || p->outer == FLGPU_OUTER_FCG, "the learned correction needs FCG");
    FLGPU_REQUIRE(p->restart_cycle >= 0, "restart_cycle must not be negative");
    FLGPU_REQUIRE(p->fallback == FLGPU_FALLBACK_NONE || p->fallback == FLGPU_FALLBACK_LU,
                  "unknown fallback iteration");
    FLGPU_REQUIRE(p->fallback == FLGPU_FALLBACK_NONE || A->nrows <= 8192,
                  "the <lu> fallback needs a system of at most 8192 unknowns");
    FLGPU_REQUIRE(b->ctx == A->ctx && x->ctx == A->ctx, "b and x must belong to A's context");
    FLGPU_REQUIRE(!mg || mg->ctx == A->ctx, "the hierarchy must belong to A's context");
    flgpu_ctx *ctx = A->ctx;
    CTX_LOCK(ctx);
    long N = A->nrows;
    int prec = A->prec;
    hipStream_t s = ctx->stream;
    SolveWork W(ctx, N, prec);
    LearnedHook *learned = nullptr;
    if (p->learned) {
      if (!mg->hook || mg->hook_A != A || mg->hook_N != A->nrows) {
        mg->hook.reset();
        mg->hook = learned_hook(mg, A);
        mg->hook_A = A;
        mg->hook_N = A->nrows;
      }
      learned = mg->hook.get();
    }
    double *slot = ctx->red_out;
    unsigned g = grid_for(N, 256);
    auto t0 = std::chrono::steady_clock::now();
    vfill(ctx, N, prec, x->d, 0.0);  // initially: x := 0 (solve.lisp:126-132)
    bool residual_p = false, initialized = false;
    int step = 0, status = 1;
    double initial = -1.0, defnorm = 0.0, alpha = 0.0;
    // a = A p and its dot with p for the current p are already queued: they modify nothing
    // the iteration keeps, so they are queued ahead of the host's wait for the stopping
    // test and simply dropped when the iteration stops or restarts
    bool ap_ready = false;
    // the current alpha's slot: 0 after a restart, then alternating 2 / 0 (the new alpha goes
    // to the other one, so no copy is needed between steps)
    int ca = 0;
    int fell_back = 0;
    double best_true = -1.0;
    int best_step = 0;
    int count = 0;  // <cg>'s steps for restart-cycle (krylow.lisp:101-105,128-131)
    bool learned_added = false;  // the network's correction has been added in some step
    // <safe-linear-solver> (linsolve.lisp:45-56,75-89): the <lu> fallback's inverse of A
    bool fallback_p = false;
    struct InverseHolder {
      DenseInverse d;
      ~InverseHolder() { d.release(); }
    } lu_inv;
    for (;;) {
      // intermediate :before: defect and its norm (solve.lisp:139-155)
      if (!residual_p) {
        spmv(A, -1.0, x->d, 1.0, b->d, W.r);
        residual_p = true;
        vdot_dev(ctx, N, prec, W.r, W.r, 3);
        fetch_slots(ctx, 4);
      }
      defnorm = std::sqrt(ctx->host_out[3]);
      if (initial < 0) initial = defnorm;
      if (defnorms && step < max_defnorms) defnorms[step] = defnorm;
      double reduction = (defnorm == 0.0) ? 0.0 : defnorm / initial;
      bool ok = (p->stop == FLGPU_STOP_REDUCTION) ? (reduction < p->tol) : (defnorm < p->tol);
      if (ok && p->true_residual && p->outer != FLGPU_OUTER_MG) {
        // confirm with the FP64 true residual; replace the recurrence if it lags
        spmv(A, -1.0, x->d, 1.0, b->d, W.t);
        vdot_dev(ctx, N, prec, W.t, W.t, 5);
        fetch_slots(ctx, 6);
        double tn = std::sqrt(ctx->host_out[5]);
        double tr = (tn == 0.0) ? 0.0 : tn / initial;
        ok = (p->stop == FLGPU_STOP_REDUCTION) ? (tr < p->tol) : (tn < p->tol);
        if (!ok) {
          vcopy(ctx, N, prec, W.t, prec, W.r);
          initialized = false;  // restart the recurrence from the true residual
          ap_ready = false;
        }
      }
      if (ok) {
        status = 0;
        break;
      }
      if (step > p->maxsteps) {
        if (p->fallback == FLGPU_FALLBACK_NONE) {
          status = 1;
          break;
        }
        if (fallback_p) {  // "The fallback iteration has failed." (linsolve.lisp:84-85)
          status = 2;
          break;
        }
        // the iterator becomes <lu>'s; the step count and the residual at hand stay
        fallback_p = true;
        lu_inv.d = dense_inverse_dev(A);
      }
      if (fallback_p) {  // <lu> iterate (linit.lisp:150-160): x += A^-1 r, residual-after nil
        dense_inverse_apply(s, lu_inv.d, prec, W.r, x->d, 0);
        residual_p = false;
        ap_ready = false;
        step++;
        continue;
      }
      // <cg> restarts at the start of the step after a completed restart cycle
      if (p->restart_cycle > 0 && count > 0 && count % p->restart_cycle == 0) {
        initialized = false;
        ap_ready = false;
      }
      // learned stall fallback: plain GMG-FCG takes over when the defect has
      // not dropped 10x within stall_window steps
      if (learned && p->stall_window > 0) {
        if (best_true < 0 || defnorm < 0.1 * best_true) {
          best_true = defnorm;
          best_step = step;
        } else if (step - best_step >= p->stall_window) {
          learned = nullptr;
          fell_back = 1;
          initialized = false;
          ap_ready = false;
          best_true = defnorm;
          best_step = step;
        }
      }
      // next-step
      switch (p->outer) {
        case FLGPU_OUTER_MG:
          mg_apply_cached(mg, prec, x->d, b->d, W.r);
          residual_p = false;
          break;
        case FLGPU_OUTER_CG:
        case FLGPU_OUTER_CG_PLAIN:
        case FLGPU_OUTER_FCG: {
          bool pc = (p->outer != FLGPU_OUTER_CG_PLAIN);
          bool fcg = (p->outer == FLGPU_OUTER_FCG);
          if (!initialized) {  // restart (krylow.lisp:88-96)
            if (pc) precondition(mg, p, W, W.r, W.p, learned);
            else vcopy(ctx, N, prec, W.r, prec, W.p);
            vdot_dev(ctx, N, prec, W.p, W.r, 0);
            fetch_slots(ctx, learned ? kLearnedFlag + 1 : 1);
            if (learned) {
              // not added: p is the cycle's own output, so the recurrence goes on without it
              if (ctx->host_out[kLearnedFlag] != 0.0) learned_added = true;
              else {
                learned = nullptr;
                fell_back = 1;
              }
            }
            alpha = ctx->host_out[0];
            ca = 0;
            initialized = true;
          }
          if (alpha != 0.0) {
            if (!ap_ready && !spmv_dot(A, W.p, W.a, 1)) {
              spmv(A, 1.0, W.p, 0.0, W.a, W.a);         // gemm! 1 A p 0 a
              vdot_dev(ctx, N, prec, W.a, W.p, 1);       // beta = a.p
            }
            ap_ready = false;
            // x += (alpha/beta) p, r -= (alpha/beta) a (FCG keeps the old r first); for an
            // FP32 hierarchy under FP64 vectors the same pass stages the new r for the cycle
            const bool staged = pc && !p->alias && prec == FLGPU_F64 &&
                                mg->p.prec == FLGPU_F32 && mg_stages_first(mg, prec);
            if (staged) {
              MGLevel &T = mg->lv[mg->top];
              k_cg_update<double, true><<<g, 256, 0, s>>>(
                  N, slot + ca, slot + 1, (const double *)W.p, (const double *)W.a,
                  (double *)x->d, (double *)W.r, (float *)mg->bs, (const float *)T.A->diag,
                  (float)mg->p.omega, (float *)T.tmp);
            } else if (prec == FLGPU_F64)
              k_cg_update<double><<<g, 256, 0, s>>>(N, slot + ca, slot + 1, (const double *)W.p,
                                                    (const double *)W.a, (double *)x->d,
                                                    (double *)W.r);
            else
              k_cg_update<float><<<g, 256, 0, s>>>(N, slot + ca, slot + 1, (const float *)W.p,
                                                   (const float *)W.a, (float *)x->d,
                                                   (float *)W.r);
            FLGPU_HIP(hipGetLastError());
            // FCG with an FP32 cycle under FP64 vectors reads the cycle's output in FP32 (the
            // same values as their FP64 copy, which is then not written)
            const bool z32 = pc && fcg && !learned && !p->alias && prec == FLGPU_F64 &&
                             mg->p.prec == FLGPU_F32;
            if (pc) precondition(mg, p, W, W.r, W.q, learned, staged, z32);
            else vcopy(ctx, N, prec, W.r, prec, W.q);
            const int na = 2 - ca;  // the new alpha's slot
            if (z32) {  // new alpha = z.r, defect norm^2 = r.r, FCG's z.a: one pass
              const int slots[3] = {na, 3, 4};
              vdot_qr_f32(ctx, N, (const float *)mg->xs, (const double *)W.r,
                          (const double *)W.a, slots);
            } else {  // new alpha = q.r, defect norm^2 = r.r, FCG's q.a: one pass
              const void *dx[3] = {W.q, W.r, W.q}, *dy[3] = {W.r, W.r, W.a};
              const int slots[3] = {na, 3, 4};
              vdot_multi_dev(ctx, N, prec, fcg ? 3 : 2, dx, dy, slots);
            }
            if (z32) {
              k_fcg_p<double, float><<<g, 256, 0, s>>>(N, slot + 4, slot + 1, (double *)W.p,
                                                       (const float *)mg->xs);
            } else if (fcg) {
              if (prec == FLGPU_F64)
                k_fcg_p<double><<<g, 256, 0, s>>>(N, slot + 4, slot + 1, (double *)W.p,
                                                  (const double *)W.q);
              else
                k_fcg_p<float><<<g, 256, 0, s>>>(N, slot + 4, slot + 1, (float *)W.p,
                                                 (const float *)W.q);
            } else {
              if (prec == FLGPU_F64)
                k_cg_p<double><<<g, 256, 0, s>>>(N, slot + na, slot + ca, (double *)W.p,
                                                 (const double *)W.q);
              else
                k_cg_p<float><<<g, 256, 0, s>>>(N, slot + na, slot + ca, (float *)W.p,
                                                (const float *)W.q);
            }
            FLGPU_HIP(hipGetLastError());
            ca = na;
            // with the network on, its decision (kLearnedFlag) comes along
            fetch_slots_async(ctx, learned ? kLearnedFlag + 1 : 4);
            if (!spmv_dot(A, W.p, W.a, 1)) {  // the next step's product, ahead of the wait
              spmv(A, 1.0, W.p, 0.0, W.a, W.a);
              vdot_dev(ctx, N, prec, W.a, W.p, 1);
            }
            ap_ready = true;
            fetch_wait(ctx);
            alpha = ctx->host_out[ca];
            if (learned) {
              if (ctx->host_out[kLearnedFlag] != 0.0) {
                learned_added = true;
              } else {
                // the step went on with the cycle's own output; the network is dropped.  A
                // direction formed with the network's correction in an earlier step is not
                // kept: GMG-FCG restarts from the residual at hand (keeping it was measured
                // worse, 138 against 97 steps at 128^3: flexible CG tolerates the change of
                // preconditioner, but that direction still slows the steps after it).
                learned = nullptr;
                fell_back = 1;
                if (learned_added) {
                  initialized = false;
                  ap_ready = false;
                }
              }
            }
          } else {
            vdot_dev(ctx, N, prec, W.r, W.r, 3);
            fetch_slots(ctx, 4);
          }
          // a completed restart cycle leaves no :residual-after: the loop recomputes r
          residual_p = !(p->restart_cycle > 0 && ++count % p->restart_cycle == 0);
          if (!residual_p) ap_ready = false;
          break;
        }
      }
      step++;
    }
    auto t1 = std::chrono::steady_clock::now();
    // true residual in the precision of A
    spmv(A, -1.0, x->d, 1.0, b->d, W.t);
    double tn = std::sqrt(vdot_host(ctx, N, prec, W.t, W.t));
    if (report) {
      report->status = status;
      report->steps = step;
      report->initial_defnorm = initial;
      report->final_defnorm = defnorm;
      report->true_residual_norm = tn;
      report->fell_back = fell_back;
      report->seconds = std::chrono::duration<double>(t1 - t0).count();
      report->fallback_used = fallback_p ? 1 : 0;
    }
    return status;
  });
}
