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

// the DOF of the slot with digits dg (axis k's offset dg[k] - h) of vertex a, or -1
__device__ __forceinline__ int blk_key_d(const BlkGeom &g, const int *a, const int *dg) {
  int key = 0;
  bool centre = true, bdry = false;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (k >= g.dim) break;
    int c = g.order * a[k] + dg[k] - g.h;
    if (c < 0 || c >= g.M) return -1;
    if (dg[k] != g.h) centre = false;
    if (c == 0 || c == g.M - 1) bdry = true;
    key = key * g.M + c;
  }
  return (bdry && !centre) ? -1 : key;
}

// the DOF of slot s of vertex a, or -1
__device__ __forceinline__ int blk_key(const BlkGeom &g, const int *a, int s) {
  int dg[3] = {0, 0, 0};
  for (int k = g.dim - 1; k >= 0; k--) {
    dg[k] = s % g.w1;
    s /= g.w1;
  }
  return blk_key_d(g, a, dg);
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

// v broadcast from lane j (j uniform over the wave)
__device__ __forceinline__ float lane_bcast(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ double lane_bcast(double v, int j) {
  long long u = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), j);
  int hi = __builtin_amdgcn_readlane((int)(u >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

constexpr int kBlkWaves = 2;  // waves per workgroup of k_blk_sweep

// Order-2 blocks in DIM dimensions: W = 3^DIM slots, slot digits base 3, last axis fastest.
// Slots 3q, 3q + 1, 3q + 2 are three lattice neighbours along the last axis: their CSR rows
// are contiguous, a run of at most 11 * 3^(DIM-1-z) * 5^z entries, z = the number of the first
// DIM - 1 digits of q that sit on the vertex's planes.
constexpr int pow_i(int b, int e) { return e == 0 ? 1 : b * pow_i(b, e - 1); }
template <int DIM>
constexpr int run_chunks(int q) {  // 64-entry chunks of run q
  int z = 0;
  for (int k = 0; k < DIM - 1; k++, q /= 3) z += (q % 3) == 1;
  return (11 * pow_i(3, DIM - 1 - z) * pow_i(5, z) + 63) / 64;
}
template <int DIM>
constexpr int range_chunks(int q0, int q1) {
  int c = 0;
  for (int q = q0; q < q1; q++) c += run_chunks<DIM>(q);
  return c;
}

// runs [Q0, Q1) of a block: every chunk's entries loaded, then the x values they name, then
// the products added into the slots' partial sums
template <class T, int DIM, int Q0, int Q1>
__device__ __forceinline__ void blk_runs(const int *__restrict__ colidx,
                                         const T *__restrict__ vals, const T *x, int lane,
                                         const int *ta, const int *e0, const int *e1,
                                         const int *e2, const int *e3, T *acc) {
  constexpr int NCH = range_chunks<DIM>(Q0, Q1);
  int cc[NCH];
  T vv[NCH];
#pragma unroll
  for (int q = Q0, c = 0; q < Q1; q++)
#pragma unroll
    for (int i = 0; i < run_chunks<DIM>(q); i++, c++) {
      int e = e0[q] + 64 * i + lane;
      int ec = max(min(e, e3[q] - 1), 0);
      cc[c] = colidx[ec];
      vv[c] = vals[ec];
    }
  T xv[NCH];
#pragma unroll
  for (int c = 0; c < NCH; c++) xv[c] = x[cc[c]];
#pragma unroll
  for (int q = Q0, c = 0; q < Q1; q++)
#pragma unroll
    for (int i = 0; i < run_chunks<DIM>(q); i++, c++) {
      int e = e0[q] + 64 * i + lane;
      T p = e < e3[q] ? vv[c] * xv[c] : (T)0;
      int t = ta[q] + (e >= e1[q]) + (e >= e2[q]);
      acc[3 * q] += t == 0 ? p : (T)0;
      acc[3 * q + 1] += t == 1 ? p : (T)0;
      acc[3 * q + 2] += t == 2 ? p : (T)0;
    }
}

// One colour of an iterate with order-2 blocks, one wave per block.
// <ssc>: r_k = b_k - (A x)_k for the block's DOFs k (stored in r, as femlisp stores the block
// residual), then x_B += omega A_BB^-1 r_B.  <psc>: x_B += omega A_BB^-1 r_B with r as given.
// b may be r.  The memory reads of a block are issued in batches of independent loads: the
// row bounds; the CSR entries of the runs, 64 at a time and every lane its own entry; the x
// values they name; then the block inverse, while the partial sums are reduced.  Every lane adds its products into per-slot
// partial sums, and the 64 partials of each slot are summed in lane order through LDS by the
// slot's own lane.
template <class T, bool SSC, int DIM>
__global__ __launch_bounds__(64 * kBlkWaves) void k_blk_sweep(BlkGeom g, long cnt,
                                                              const int *__restrict__ rowptr,
                                                              const int *__restrict__ colidx,
                                                              const T *__restrict__ vals,
                                                              const T *__restrict__ inv, T omega,
                                                              T *x, const T *b, T *r) {
  constexpr int W = pow_i(3, DIM), NRUN = W / 3;
  __shared__ T part[kBlkWaves][64][W + 1];
  unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int wv = threadIdx.x >> 6;
  long w = (long)lb * kBlkWaves + wv;
  if (w >= cnt) return;  // uniform over the wave
  const int lane = threadIdx.x & 63;
  int av[3];
  blk_vertex(g, w, av);
  int key = lane < W ? blk_key(g, av, lane) : -1;
  const T *iv = inv + blk_vertex_id(g, av) * W * W;
  T ivr[W];
  auto load_inverse = [&]() {
#pragma unroll
    for (int j = 0; j < W; j++) ivr[j] = lane < W ? iv[j * W + lane] : (T)0;
  };
  T myr = 0;
  if (SSC) {
    int rs = 0, re = 0;
    if (key >= 0) {
      rs = rowptr[key];
      re = rowptr[key + 1];
    }
    // per run: first present slot, and the entry bounds [e0, e3) with the ends e1, e2 of its
    // first two rows (present rows are consecutive lattice points; none present: empty run)
    int ta[NRUN], e0[NRUN], e1[NRUN], e2[NRUN], e3[NRUN];
#pragma unroll
    for (int q = 0; q < NRUN; q++) {
      int k0 = __builtin_amdgcn_readlane(key, 3 * q);
      int k1 = __builtin_amdgcn_readlane(key, 3 * q + 1);
      int k2 = __builtin_amdgcn_readlane(key, 3 * q + 2);
      int a = k0 >= 0 ? 0 : (k1 >= 0 ? 1 : (k2 >= 0 ? 2 : -1));
      int z = k2 >= 0 ? 2 : (k1 >= 0 ? 1 : 0);
      ta[q] = a < 0 ? 0 : a;
      e0[q] = a < 0 ? 0 : __builtin_amdgcn_readlane(rs, 3 * q + a);
      e3[q] = a < 0 ? 0 : __builtin_amdgcn_readlane(re, 3 * q + z);
      e1[q] = (a >= 0 && a < z) ? __builtin_amdgcn_readlane(re, 3 * q + a) : e3[q];
      e2[q] = (a >= 0 && a + 1 < z) ? __builtin_amdgcn_readlane(re, 3 * q + a + 1) : e3[q];
    }
    T acc[W];
#pragma unroll
    for (int s = 0; s < W; s++) acc[s] = 0;
    blk_runs<T, DIM, 0, NRUN>(colidx, vals, x, lane, ta, e0, e1, e2, e3, acc);
#pragma unroll
    for (int s = 0; s < W; s++) part[wv][lane][s] = acc[s];
    load_inverse();  // in flight during the reduction, in the registers acc held
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (key >= 0) {
      T sum = 0;
      for (int l = 0; l < 64; l++) sum += part[wv][l][lane];
      myr = b[key] - sum;
    }
  } else {
    load_inverse();
    if (key >= 0) myr = r[key];
  }
  T y = 0;
#pragma unroll
  for (int j = 0; j < W; j++) y += ivr[j] * lane_bcast(myr, j);
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
  if (!A->blk_inv) {
    A->blk_inv = dev_alloc(bytes);
    layout_changed(A);
  }
  if (A->prec == FLGPU_F64)
    k_blk_invert<double><<<(unsigned)nv, 64, 0, A->ctx->stream>>>(
        g, A->rowptr, A->colidx, (const double *)A->vals, (double *)A->blk_inv);
  else
    k_blk_invert<float><<<(unsigned)nv, 64, 0, A->ctx->stream>>>(
        g, A->rowptr, A->colidx, (const float *)A->vals, (float *)A->blk_inv);
  FLGPU_HIP(hipGetLastError());
}

void blocks_drop(flgpu_mat *A) {
  if (A->blk_inv) layout_changed(A);
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
      unsigned grid = xcd_grid(grid_for(cnt, kBlkWaves));
      auto launch = [&](auto kern) {
        kern<<<grid, 64 * kBlkWaves, 0, st>>>(g, cnt, A->rowptr, A->colidx, (const T *)A->vals,
                                              (const T *)A->blk_inv, omega, x, b, r);
      };
      if (g.dim == 3)
        launch(ssc ? k_blk_sweep<T, true, 3> : k_blk_sweep<T, false, 3>);
      else if (g.dim == 2)
        launch(ssc ? k_blk_sweep<T, true, 2> : k_blk_sweep<T, false, 2>);
      else
        launch(ssc ? k_blk_sweep<T, true, 1> : k_blk_sweep<T, false, 1>);
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
