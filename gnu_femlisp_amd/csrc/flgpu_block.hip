// Vertex-centred block smoothers: <geometric-ssc> and <geometric-psc>
// (src/iteration/blockit.lisp:41-203, src/iteration/geoblock.lisp:37-157).
//
// femlisp's vertex-centred block of a mesh vertex v holds v and every DOF of a cell containing
// v that is not Dirichlet (find-vertex-centered-block, geoblock.lisp:65-91).  On the lattice of
// a Q`order` discretisation (M = order n + 1 points per axis) v sits at order a, a in
// [0, n]^dim, and its block is the points v + d with |d_k| <= order - 1: W = (2 order - 1)^dim
// slots, slot s = the digits of s in base 2 order - 1, axis 0 most significant.  A slot is
// absent when its point lies outside the lattice or is Dirichlet (v itself always counts).
//
// Blocks are swept colour by colour, colour = the parity of a per axis (axis 0 most
// significant).  The vertices of two blocks of one colour are at least 2 order lattice steps
// apart along some axis.  The rows of a block couple only DOFs of the cells around its vertex,
// within order steps of it, while the other block's DOFs are at least order + 1 steps away:
// neither block reads what the other writes, and the sequential sweep over one colour equals
// its parallel update.
//
// The blocks are symmetric positive definite (principal submatrices of the SPD operator, the
// Dirichlet vertex an identity row), so their inverses are formed once by Gauss-Jordan without
// pivoting in FP64 and stored (femlisp's store-p t, blockit.lisp:118-121), per vertex in
// lattice vertex order as inv[v][j][i] = (A_BB^-1)_ij, identity rows and columns at absent
// slots.  One iterate then costs per block the block's matrix rows (<ssc> only) and the W x W
// inverse; order-1 blocks are single DOFs and take one lane each.
#include "flgpu_internal.h"

namespace flgpu {

namespace {

constexpr int kMaxW = 27;  // (2 order - 1)^dim for order <= 2, dim <= 3

struct BlkGeom {
  int dim, order, n, M, W, w1, h;
  int par[3];  // colour parity per axis
  int cnt[3];  // vertices of that parity per axis
};

BlkGeom blk_geom(const flgpu_mat *A, int colour) {
  BlkGeom g{};
  g.dim = A->lat_dim;
  g.order = A->lat_order;
  g.n = A->lat_n;
  g.M = g.order * g.n + 1;
  g.w1 = 2 * g.order - 1;
  g.h = g.order - 1;
  g.W = 1;
  for (int k = 0; k < g.dim; k++) g.W *= g.w1;
  for (int k = 0; k < 3; k++) {
    g.par[k] = 0;
    g.cnt[k] = 1;
  }
  for (int k = 0; k < g.dim; k++) {
    g.par[k] = (colour >> (g.dim - 1 - k)) & 1;
    g.cnt[k] = (g.n + 2 - g.par[k]) / 2;
  }
  return g;
}

long blk_count(const BlkGeom &g) {
  long c = 1;
  for (int k = 0; k < g.dim; k++) c *= g.cnt[k];
  return c;
}

// vertex coordinates a of the w-th vertex of the colour (last axis fastest)
__device__ __forceinline__ void blk_vertex(const BlkGeom &g, long w, int *a) {
  for (int k = g.dim - 1; k >= 0; k--) {
    int j = (int)(w % g.cnt[k]);
    w /= g.cnt[k];
    a[k] = 2 * j + g.par[k];
  }
}

__device__ __forceinline__ long blk_vertex_id(const BlkGeom &g, const int *a) {
  long v = 0;
  for (int k = 0; k < g.dim; k++) v = v * (g.n + 1) + a[k];
  return v;
}

// the DOF of slot s of vertex a, or -1
__device__ __forceinline__ int blk_key(const BlkGeom &g, const int *a, int s) {
  int c[3];
  bool centre = true, bdry = false;
  for (int k = g.dim - 1; k >= 0; k--) {
    int d = s % g.w1;
    s /= g.w1;
    c[k] = g.order * a[k] + d - g.h;
    if (c[k] < 0 || c[k] >= g.M) return -1;
    if (d != g.h) centre = false;
    if (c[k] == 0 || c[k] == g.M - 1) bdry = true;
  }
  if (bdry && !centre) return -1;
  long key = 0;
  for (int k = 0; k < g.dim; k++) key = key * g.M + c[k];
  return (int)key;
}

// Block inverses, one wave per vertex (grid: all (n+1)^dim vertices, colour 0's geometry with
// the parities ignored).  The block matrix [A_BB | I] is gathered into LDS from the CSR rows
// of the block's DOFs (a column belongs to the block when its offset from v is within the
// block), then reduced by Gauss-Jordan.
template <class T>
__global__ __launch_bounds__(64) void k_blk_invert(BlkGeom g, const int *__restrict__ rowptr,
                                                   const int *__restrict__ colidx,
                                                   const T *__restrict__ vals, T *__restrict__ inv) {
  __shared__ double a[kMaxW][2 * kMaxW + 1];
  __shared__ double fcol[kMaxW];
  const int W = g.W, lane = threadIdx.x;
  int av[3];
  long v = blockIdx.x;
  {
    long t = v;
    for (int k = g.dim - 1; k >= 0; k--) {
      av[k] = (int)(t % (g.n + 1));
      t /= (g.n + 1);
    }
  }
  for (int e = lane; e < W * 2 * W; e += 64) {
    int i = e / (2 * W), j = e % (2 * W);
    a[i][j] = (j - W == i) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int s = 0; s < W; s++) {
    int ks = blk_key(g, av, s);
    if (ks < 0) {
      if (lane == 0) a[s][s] = 1.0;  // absent slot: identity row
      continue;
    }
    for (int e = rowptr[ks] + lane; e < rowptr[ks + 1]; e += 64) {
      int col = colidx[e];
      int t = 0;
      bool in = true, centre = true, bdry = false;
      for (int k = g.dim - 1, mul = 1; k >= 0; k--, mul *= g.w1) {
        int ck = col % g.M;
        col /= g.M;
        int d = ck - (g.order * av[k] - g.h);
        if (d < 0 || d >= g.w1) in = false;
        if (d != g.h) centre = false;
        if (ck == 0 || ck == g.M - 1) bdry = true;
        t += d * mul;
      }
      if (in && (centre || !bdry)) a[s][t] = (double)vals[e];
    }
  }
  __syncthreads();
  for (int k = 0; k < W; k++) {
    double p = a[k][k];
    __syncthreads();
    for (int j = lane; j < 2 * W; j += 64) a[k][j] /= p;
    for (int i = lane; i < W; i += 64) fcol[i] = a[i][k];
    __syncthreads();
    for (int e = lane; e < W * 2 * W; e += 64) {
      int i = e / (2 * W), j = e % (2 * W);
      if (i != k) a[i][j] -= fcol[i] * a[k][j];
    }
    __syncthreads();
  }
  T *out = inv + v * W * W;
  for (int e = lane; e < W * W; e += 64) {
    int j = e / W, i = e % W;
    out[e] = (T)a[i][W + j];
  }
}

__device__ __forceinline__ double lane_read(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ float lane_read(float v, int src) { return __shfl(v, src, 64); }

template <class T>
__device__ __forceinline__ T wave_allsum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One colour of an iterate with blocks of W > 1 slots, one wave per block.
// <ssc>: r_k = b_k - (A x)_k for the block's DOFs k (stored in r, as femlisp stores the block
// residual), then x_B += omega A_BB^-1 r_B.  <psc>: x_B += omega A_BB^-1 r_B with r as given.
// b may be r.
template <class T, bool SSC>
__global__ __launch_bounds__(256) void k_blk_sweep(BlkGeom g, long cnt,
                                                   const int *__restrict__ rowptr,
                                                   const int *__restrict__ colidx,
                                                   const T *__restrict__ vals,
                                                   const T *__restrict__ inv, T omega, T *x,
                                                   const T *b, T *r) {
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  long w = (long)lb * 4 + (threadIdx.x >> 6);
  if (w >= cnt) return;  // uniform over the wave
  const int lane = threadIdx.x & 63, W = g.W;
  int av[3];
  blk_vertex(g, w, av);
  int key = lane < W ? blk_key(g, av, lane) : -1;
  T myr = 0;
  if (SSC) {
    for (int s = 0; s < W; s++) {
      int ks = __builtin_amdgcn_readlane(key, s);
      if (ks < 0) continue;
      T acc = 0;
      for (int e = rowptr[ks] + lane; e < rowptr[ks + 1]; e += 64) acc += vals[e] * x[colidx[e]];
      acc = wave_allsum(acc);
      if (lane == s) myr = b[ks] - acc;
    }
  } else if (key >= 0) {
    myr = r[key];
  }
  const T *iv = inv + blk_vertex_id(g, av) * W * W;
  T y = 0;
  for (int j = 0; j < W; j++) {
    T rj = lane_read(myr, j);
    if (lane < W) y += iv[j * W + lane] * rj;
  }
  if (key >= 0) {
    if (SSC) r[key] = myr;
    x[key] += omega * y;
  }
}

// One colour of an iterate with one-DOF blocks (order 1), one lane per block: gesv! of the
// 1 x 1 block is the division by the diagonal.
template <class T, bool SSC>
__global__ __launch_bounds__(256) void k_blk_point(BlkGeom g, long cnt,
                                                   const int *__restrict__ rowptr,
                                                   const int *__restrict__ colidx,
                                                   const T *__restrict__ vals,
                                                   const T *__restrict__ diag, T omega, T *x,
                                                   const T *b, T *r) {
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  long w = (long)lb * 256 + threadIdx.x;
  if (w >= cnt) return;
  int av[3];
  blk_vertex(g, w, av);
  int key = blk_key(g, av, 0);
  T res;
  if (SSC) {
    T acc = 0;
    for (int e = rowptr[key]; e < rowptr[key + 1]; e++) acc += vals[e] * x[colidx[e]];
    res = b[key] - acc;
    r[key] = res;
  } else {
    res = r[key];
  }
  x[key] += omega * (res / diag[key]);
}

}  // namespace

void blocks_prepare(flgpu_mat *A) {
  FLGPU_REQUIRE(A->lat_dim >= 1 && A->lat_dim <= 3 && A->lat_order >= 1 && A->lat_order <= 2,
                "block smoothers need a matrix assembled on a lattice (dim 1-3, order 1-2)");
  BlkGeom g = blk_geom(A, 0);
  long nv = 1;
  for (int k = 0; k < g.dim; k++) nv *= g.n + 1;
  long N = 1;
  for (int k = 0; k < g.dim; k++) N *= g.M;
  FLGPU_REQUIRE(A->nrows == N && A->ncols == N, "matrix size differs from its lattice");
  if (g.W == 1) return;  // one-DOF blocks use the diagonal
  size_t bytes = (size_t)nv * g.W * g.W * prec_bytes(A->prec);
  if (!A->blk_inv) A->blk_inv = dev_alloc(bytes);
  if (A->prec == FLGPU_F64)
    k_blk_invert<double><<<(unsigned)nv, 64, 0, A->ctx->stream>>>(
        g, A->rowptr, A->colidx, (const double *)A->vals, (double *)A->blk_inv);
  else
    k_blk_invert<float><<<(unsigned)nv, 64, 0, A->ctx->stream>>>(
        g, A->rowptr, A->colidx, (const float *)A->vals, (float *)A->blk_inv);
  FLGPU_HIP(hipGetLastError());
}

void blocks_drop(flgpu_mat *A) {
  dev_free(A->blk_inv);
  A->blk_inv = nullptr;
}

template <class T>
static void block_sweep_t(const flgpu_mat *A, bool ssc, T omega, int backwards, T *x, const T *b,
                          T *r) {
  int ncol = 1 << A->lat_dim;
  for (int k = 0; k < ncol; k++) {
    BlkGeom g = blk_geom(A, backwards ? ncol - 1 - k : k);
    long cnt = blk_count(g);
    hipStream_t st = A->ctx->stream;
    if (g.W == 1) {
      unsigned grid = xcd_grid(grid_for(cnt, 256));
      if (ssc)
        k_blk_point<T, true><<<grid, 256, 0, st>>>(g, cnt, A->rowptr, A->colidx,
                                                   (const T *)A->vals, (const T *)A->diag, omega,
                                                   x, b, r);
      else
        k_blk_point<T, false><<<grid, 256, 0, st>>>(g, cnt, A->rowptr, A->colidx,
                                                    (const T *)A->vals, (const T *)A->diag,
                                                    omega, x, b, r);
    } else {
      unsigned grid = xcd_grid(grid_for(cnt, 4));
      if (ssc)
        k_blk_sweep<T, true><<<grid, 256, 0, st>>>(g, cnt, A->rowptr, A->colidx,
                                                   (const T *)A->vals, (const T *)A->blk_inv,
                                                   omega, x, b, r);
      else
        k_blk_sweep<T, false><<<grid, 256, 0, st>>>(g, cnt, A->rowptr, A->colidx,
                                                    (const T *)A->vals, (const T *)A->blk_inv,
                                                    omega, x, b, r);
    }
    FLGPU_HIP(hipGetLastError());
  }
}

void block_sweep(const flgpu_mat *A, bool ssc, double omega, int backwards, void *x,
                 const void *b, void *r) {
  BlkGeom g = blk_geom(A, 0);
  if (g.W > 1 && !A->blk_inv) blocks_prepare(const_cast<flgpu_mat *>(A));
  if (A->prec == FLGPU_F64)
    block_sweep_t<double>(A, ssc, omega, backwards, (double *)x, (const double *)b, (double *)r);
  else
    block_sweep_t<float>(A, ssc, (float)omega, backwards, (float *)x, (const float *)b,
                         (float *)r);
}

}  // namespace flgpu

using namespace flgpu;

static void check_block_vec(const flgpu_mat *A, const flgpu_vec *v, const char *what) {
  FLGPU_REQUIRE(v != nullptr, std::string(what) + " must not be NULL");
  FLGPU_REQUIRE(v->ctx == A->ctx, std::string(what) + " belongs to another context");
  FLGPU_REQUIRE(v->n == A->nrows, std::string(what) + ": length mismatch");
  FLGPU_REQUIRE(v->prec == A->prec, std::string(what) + ": precision differs from the matrix");
}

extern "C" int flgpu_block_iterate(const flgpu_mat *A, int kind, double omega, int backwards,
                                   flgpu_vec *x, const flgpu_vec *b, flgpu_vec *r) {
  return guard([&] {
    FLGPU_REQUIRE(A, "matrix must not be NULL");
    FLGPU_REQUIRE(kind == FLGPU_SMOOTH_SSC || kind == FLGPU_SMOOTH_PSC,
                  "kind must be FLGPU_SMOOTH_SSC or FLGPU_SMOOTH_PSC");
    check_block_vec(A, x, "x");
    check_block_vec(A, b, "b");
    check_block_vec(A, r, "r");
    FLGPU_REQUIRE(x != b && x != r, "x must differ from b and r");
    CTX_LOCK(A->ctx);
    flgpu_mat *M = const_cast<flgpu_mat *>(A);
    if (!M->blk_inv) blocks_prepare(M);
    block_sweep(A, kind == FLGPU_SMOOTH_SSC, omega, backwards, x->d, b->d, r->d);
    return 0;
  });
}
