// Learned coarse correction: a 3-level 3D U-Net (16/32/64 channels, 3x3x3
// kernels) with BF16 weights and FP32 accumulation on MFMA, trained on the
// GPU by its own forward and backward kernels with Adam.
//
// No femlisp counterpart: femlisp has no learned components.  The network is
// used by flgpu_solve as an extra correction inside the flexible CG: after
// the V-cycle z1 = V(r), the residual r1 = r - A z1 is fed to the network and
// z = z1 + N(r1); plain GMG-FCG takes over when the defect stalls.
//
// Tensors are channels-last [batch][z][y][x][c] in FP32.  Convolutions are
// implicit GEMMs on v_mfma_f32_16x16x16bf16_1k: M = voxels (tiles of 16),
// N = output channels (tiles of 16), K = 27 taps x input channels (chunks of
// 16); activations and weights are rounded to BF16 when they enter the
// matrix cores, products accumulate in FP32.  Weight gradients are implicit
// GEMMs on the same instruction (voxels as K) with per-chunk partials reduced
// in a fixed order.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "flgpu_learned.h"

namespace flgpu {

constexpr int NL = 9;
static const int L_CIN[NL] = {-1, 16, 16, 32, 32, 64, 96, 48, 16};
static const int L_COUT[NL] = {16, 16, 32, 32, 64, 64, 32, 16, 1};
static const int L_LEVEL[NL] = {0, 0, 1, 1, 2, 2, 1, 0, 0};

typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_bits(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// ---- convolution forward / input gradient (implicit GEMM on MFMA) --------------------

// out[v][o] = bias[o] + sum_t sum_c in[v + off(t) * sgn][c] * W[t][c][o]
// with W given in BF16 as wk[t][o][c] (the K index c contiguous).  sgn = +1
// for the forward convolution; sgn = -1 with wk = W^T computes the input
// gradient (taps mirrored).  One wavefront: 16 voxels x all output channels.
template <int NT>
__global__ __launch_bounds__(256) void k_conv_mfma(int batch, int n, int ci, int co,
                                                   const float *__restrict__ in,
                                                   const short *__restrict__ wk,
                                                   const float *__restrict__ bias, float *out,
                                                   int relu, int sgn) {
  long nv = (long)n * n * n;
  long total = batch * nv;
  long tile = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  long v0 = tile * 16;
  if (v0 >= total) return;  // wave-uniform
  int lane = threadIdx.x & 63;
  int i = lane & 15, g = lane >> 4;
  long v = v0 + i;
  bool vin = v < total;
  int b = 0, z = 0, y = 0, x = 0;
  if (vin) {
    b = (int)(v / nv);
    long r = v % nv;
    z = (int)(r / ((long)n * n));
    y = (int)((r / n) % n);
    x = (int)(r % n);
  }
  v4f acc[NT];
  for (int t = 0; t < NT; t++) acc[t] = (v4f){0.f, 0.f, 0.f, 0.f};
  bool vec_ok = (ci % 4) == 0;
  for (int tap = 0; tap < 27; tap++) {
    int dz = (tap / 9 - 1) * sgn, dy = ((tap / 3) % 3 - 1) * sgn, dx = (tap % 3 - 1) * sgn;
    int zz = z + dz, yy = y + dy, xx = x + dx;
    bool ok = vin && zz >= 0 && zz < n && yy >= 0 && yy < n && xx >= 0 && xx < n;
    const float *ip = in + ((long)b * nv + ((long)zz * n + yy) * n + xx) * ci;
    for (int c0 = 0; c0 < ci; c0 += 16) {
      int cb = c0 + 4 * g;
      v4s a;
      if (ok && vec_ok && cb < ci) {
        float4 f = *(const float4 *)(ip + cb);
        a[0] = bf16_bits(f.x);
        a[1] = bf16_bits(f.y);
        a[2] = bf16_bits(f.z);
        a[3] = bf16_bits(f.w);
      } else {
        for (int j = 0; j < 4; j++) a[j] = (ok && cb + j < ci) ? bf16_bits(ip[cb + j]) : (short)0;
      }
      for (int t = 0; t < NT; t++) {
        int o = t * 16 + i;
        v4s bw;
        const short *wp = wk + ((long)tap * co + o) * ci + cb;
        for (int j = 0; j < 4; j++) bw[j] = (o < co && cb + j < ci) ? wp[j] : (short)0;
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bw, acc[t], 0, 0, 0);
      }
    }
  }
  // D[4g+v][i]: voxel v0 + 4g + q, channel t*16 + i
  for (int t = 0; t < NT; t++) {
    int o = t * 16 + i;
    if (o >= co) continue;
    float bo = bias ? bias[o] : 0.f;
    for (int q = 0; q < 4; q++) {
      long vv = v0 + 4 * g + q;
      if (vv >= total) continue;
      float val = acc[t][q] + bo;
      if (relu && val < 0.f) val = 0.f;
      out[vv * co + o] = val;
    }
  }
}

// ---- weight gradients: gW[t][c][o] = sum_v in[v + off(t)][c] * gout[v][o] -------------
// An implicit GEMM on v_mfma_f32_16x16x16bf16_1k per tap: M = input channels (tiles of 16),
// N = output channels (tiles of 16), K = voxels (16 per product).  A block takes one tap and a
// chunk of kWgVox voxels; it stages kWgChunk voxels at a time in LDS (the tap-shifted inputs
// and the output gradients, rounded to BF16), and each wave accumulates up to kWgTiles
// (c, o) tiles in FP32.  Blocks of the 27 taps of a chunk are adjacent in the grid, so the
// chunk's values are read from HBM once and from L2 by the other taps.  The centre tap's
// blocks also sum gout per channel in FP32 (the bias gradient).  Every block writes its sums
// as a partial [chunk][27 ci co + co], reduced over the chunks in a fixed order.

constexpr int kWgChunk = 64;  // voxels staged per LDS round
constexpr int kWgVox = 4096;  // voxels per block (partial sum)
constexpr int kWgTiles = 4;   // (c, o) tiles per wave: 16 tiles of 16 x 16 for 4 waves

__global__ __launch_bounds__(256) void k_conv_wgrad(int batch, int n, int ci, int co,
                                                    const float *__restrict__ in,
                                                    const float *__restrict__ gout,
                                                    float *partial) {
  __shared__ short sin[kWgChunk * 96];  // [voxel][c], c padded to a multiple of 16
  __shared__ short sg[kWgChunk * 64];   // [voxel][o]
  __shared__ long soff[kWgChunk];       // element offset of the shifted input voxel, -1: zero
  __shared__ float sb[256];
  const int cip = (ci + 15) & ~15, cop = (co + 15) & ~15;
  const int tap = (int)(blockIdx.x % 27);
  const long chunk = blockIdx.x / 27;
  const long nv = (long)n * n * n, total = batch * nv;
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
  const int nto = cop / 16, ntiles = (cip / 16) * nto;
  v4f acc[kWgTiles];
  for (int t = 0; t < kWgTiles; t++) acc[t] = (v4f){0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;  // centre tap: this thread's channel o = threadIdx.x % cop
  const long vbeg = chunk * kWgVox, vend = min(vbeg + kWgVox, total);
  for (long vs = vbeg; vs < vend; vs += kWgChunk) {
    __syncthreads();
    if (threadIdx.x < kWgChunk) {
      const long v = vs + threadIdx.x;
      long off = -1;
      if (v < vend) {
        const int b = (int)(v / nv);
        const long r = v % nv;
        const int z = (int)(r / ((long)n * n)) + dz, y = (int)((r / n) % n) + dy,
                  x = (int)(r % n) + dx;
        if (z >= 0 && z < n && y >= 0 && y < n && x >= 0 && x < n)
          off = ((long)b * nv + ((long)z * n + y) * n + x) * ci;
      }
      soff[threadIdx.x] = off;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kWgChunk * cip; e += blockDim.x) {
      const int k = e / cip, c = e - k * cip;
      const long off = soff[k];
      sin[e] = bf16_bits(off >= 0 && c < ci ? in[off + c] : 0.f);
    }
    for (int e = threadIdx.x; e < kWgChunk * cop; e += blockDim.x) {
      const int k = e / cop, o = e - k * cop;
      const long v = vs + k;
      const float val = v < vend && o < co ? gout[v * co + o] : 0.f;
      bsum += val;
      sg[e] = bf16_bits(val);
    }
    __syncthreads();
#pragma unroll 1
    for (int k0 = 0; k0 < kWgChunk; k0 += 16) {
#pragma unroll
      for (int t = 0; t < kWgTiles; t++) {
        const int tile = w + 4 * t;
        if (tile >= ntiles) break;  // wave-uniform
        const int c0 = (tile / nto) * 16, o0 = (tile % nto) * 16;
        v4s a, bb;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          a[j] = sin[(k0 + 4 * g + j) * cip + c0 + i];   // A[c0 + i][voxel 4g + j]
          bb[j] = sg[(k0 + 4 * g + j) * cop + o0 + i];   // B[voxel 4g + j][o0 + i]
        }
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[t], 0, 0, 0);
      }
    }
  }
  const long npair = (long)ci * co, stride = 27 * npair + co;
  float *part = partial + chunk * stride;
  // D[4g + q][i]: input channel c0 + 4g + q, output channel o0 + i
#pragma unroll
  for (int t = 0; t < kWgTiles; t++) {
    const int tile = w + 4 * t;
    if (tile >= ntiles) break;
    const int c0 = (tile / nto) * 16, o = (tile % nto) * 16 + i;
    for (int q = 0; q < 4; q++) {
      const int c = c0 + 4 * g + q;
      if (c < ci && o < co) part[tap * npair + (long)c * co + o] = acc[t][q];
    }
  }
  if (tap == 13) {  // bias: threads with the same channel summed in thread order
    sb[threadIdx.x] = bsum;
    __syncthreads();
    if (threadIdx.x < co) {
      float s = 0.f;
      for (int k = threadIdx.x; k < 256; k += cop) s += sb[k];
      part[27 * npair + threadIdx.x] = s;
    }
  }
}

// gW[t][c][o] and gb[o] = sums over the chunks (fixed order) of the partials
__global__ void k_wgrad_reduce(long nchunks, int np27, int co, const float *partial, float *gw,
                               float *gb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np27 + co) return;
  const long stride = (long)np27 + co;
  float s = 0.f;
  for (long c = 0; c < nchunks; c++) s += partial[c * stride + p];
  if (p < np27) gw[p] = s;
  else gb[p - np27] = s;
}

// ---- elementwise / resampling ------------------------------------------------------------

__global__ void k_relu_bwd(long m, const float *act, float *g) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && act[i] <= 0.f) g[i] = 0.f;
}

// average pool 2x2x2: out (n/2) from in (n)
__global__ void k_pool(int batch, int n, int c, const float *in, float *out) {
  int h = n / 2;
  long total = (long)batch * h * h * h * c;
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  int k = (int)(e % c);
  long pv = e / c;
  int x = (int)(pv % h), y = (int)((pv / h) % h), z = (int)((pv / ((long)h * h)) % h);
  int b = (int)(pv / ((long)h * h * h));
  float s = 0.f;
  for (int d = 0; d < 8; d++) {
    int zz = 2 * z + (d >> 2), yy = 2 * y + ((d >> 1) & 1), xx = 2 * x + (d & 1);
    s += in[(((long)b * n + zz) * n * n + (long)yy * n + xx) * c + k];
  }
  out[e] = 0.125f * s;
}

// gin (n) += gout (n/2) / 8
__global__ void k_pool_bwd(int batch, int n, int c, const float *gout, float *gin) {
  long total = (long)batch * n * n * n * c;
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  int k = (int)(e % c);
  long v = e / c;
  int x = (int)(v % n), y = (int)((v / n) % n), z = (int)((v / ((long)n * n)) % n);
  int b = (int)(v / ((long)n * n * n));
  int h = n / 2;
  gin[e] += 0.125f * gout[(((long)b * h + z / 2) * h * h + (long)(y / 2) * h + x / 2) * c + k];
}

// cat (n, cu+cs) = [nearest-upsampled up (n/2, cu), skip (n, cs)]
__global__ void k_upcat(int batch, int n, int cu, int cs, const float *up, const float *skip,
                        float *cat) {
  int ct = cu + cs;
  long total = (long)batch * n * n * n * ct;
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  int k = (int)(e % ct);
  long v = e / ct;
  if (k >= cu) {
    cat[e] = skip[v * cs + (k - cu)];
    return;
  }
  int x = (int)(v % n), y = (int)((v / n) % n), z = (int)((v / ((long)n * n)) % n);
  int b = (int)(v / ((long)n * n * n));
  int h = n / 2;
  cat[e] = up[(((long)b * h + z / 2) * h * h + (long)(y / 2) * h + x / 2) * cu + k];
}

// backward of upcat: gup (n/2, cu) = sum over the 8 children; gskip (n, cs) += skip part
__global__ void k_upcat_bwd_up(int batch, int n, int cu, int cs, const float *gcat, float *gup) {
  int h = n / 2;
  long total = (long)batch * h * h * h * cu;
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  int k = (int)(e % cu);
  long pv = e / cu;
  int x = (int)(pv % h), y = (int)((pv / h) % h), z = (int)((pv / ((long)h * h)) % h);
  int b = (int)(pv / ((long)h * h * h));
  float s = 0.f;
  for (int d = 0; d < 8; d++) {
    int zz = 2 * z + (d >> 2), yy = 2 * y + ((d >> 1) & 1), xx = 2 * x + (d & 1);
    s += gcat[(((long)b * n + zz) * n * n + (long)yy * n + xx) * (cu + cs) + k];
  }
  gup[e] = s;
}

__global__ void k_upcat_bwd_skip(long nvox, int cu, int cs, const float *gcat, float *gskip) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nvox * cs) return;
  long v = e / cs;
  int k = (int)(e % cs);
  gskip[e] = gcat[v * (cu + cs) + cu + k];
}

// MSE: g = 2 (y - t) / m; per-block partial sums of (y - t)^2
__global__ __launch_bounds__(256) void k_mse(long m, const float *y, const float *t, float *g,
                                             double *partial) {
  __shared__ double s[256];
  double acc = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < m; i += (long)gridDim.x * 256) {
    float d = y[i] - t[i];
    g[i] = 2.f * d / (float)m;
    acc += (double)d * d;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void k_sum_partials(int np, const double *p, double *out, double scale) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < np; i++) s += p[i];
    *out = s * scale;
  }
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters
__global__ void k_adam(long np, float *theta, const float *g, float *m, float *v, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  theta[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 weight copies: wf[t][o][c] (forward) and wg[t][c][o] (input gradient)
__global__ void k_wcopies(int ci, int co, const float *W, short *wf, short *wg) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = 27L * ci * co;
  if (e >= tot) return;
  int t = (int)(e / ((long)ci * co));
  int c = (int)((e / co) % ci);
  int o = (int)(e % co);
  short b = bf16_bits(W[e]);  // W[t][c][o]
  wf[((long)t * co + o) * ci + c] = b;
  wg[((long)t * ci + c) * co + o] = b;
}

}  // namespace flgpu

// ---- network object ---------------------------------------------------------------------

struct flgpu_unet {
  flgpu_ctx *ctx = nullptr;
  int cin = 2;
  long np = 0;
  long woff[flgpu::NL], boff[flgpu::NL];
  float *theta = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  short *wf = nullptr, *wg = nullptr;  // BF16 copies, same offsets as W blocks
  int step = 0;
  // activation storage for the current (batch, n)
  int cap_batch = 0, cap_n = 0;
  float *a[flgpu::NL] = {}, *p0 = nullptr, *p1 = nullptr, *c1 = nullptr, *c0 = nullptr;
  float *g[flgpu::NL] = {}, *gp0 = nullptr, *gp1 = nullptr, *gc1 = nullptr, *gc0 = nullptr;
  float *gskip0 = nullptr, *gskip1 = nullptr, *partial = nullptr;
  long partial_cap = 0;
  double *lpart = nullptr;
};

namespace flgpu {

flgpu_ctx *unet_ctx(const flgpu_unet *u) { return u->ctx; }

static long vox(int batch, int n) { return (long)batch * n * n * n; }

static void unet_free_acts(flgpu_unet *u) {
  for (int l = 0; l < NL; l++) {
    dev_free(u->a[l]);
    dev_free(u->g[l]);
    u->a[l] = u->g[l] = nullptr;
  }
  for (float **p : {&u->p0, &u->p1, &u->c1, &u->c0, &u->gp0, &u->gp1, &u->gc1, &u->gc0,
                    &u->gskip0, &u->gskip1}) {
    dev_free(*p);
    *p = nullptr;
  }
  u->cap_batch = u->cap_n = 0;
}

static void unet_alloc_acts(flgpu_unet *u, int batch, int n, bool train) {
  if (u->cap_batch == batch && u->cap_n == n && (!train || u->g[0])) return;
  unet_free_acts(u);
  int nl[3] = {n, n / 2, n / 4};
  auto fa = [](long cnt) { return (float *)dev_alloc(sizeof(float) * cnt); };
  for (int l = 0; l < NL; l++) u->a[l] = fa(vox(batch, nl[L_LEVEL[l]]) * L_COUT[l]);
  u->p0 = fa(vox(batch, nl[1]) * 16);
  u->p1 = fa(vox(batch, nl[2]) * 32);
  u->c1 = fa(vox(batch, nl[1]) * 96);
  u->c0 = fa(vox(batch, nl[0]) * 48);
  if (train) {
    for (int l = 0; l < NL; l++) u->g[l] = fa(vox(batch, nl[L_LEVEL[l]]) * L_COUT[l]);
    u->gp0 = fa(vox(batch, nl[1]) * 16);
    u->gp1 = fa(vox(batch, nl[2]) * 32);
    u->gc1 = fa(vox(batch, nl[1]) * 96);
    u->gc0 = fa(vox(batch, nl[0]) * 48);
    u->gskip0 = fa(vox(batch, nl[0]) * 16);
    u->gskip1 = fa(vox(batch, nl[1]) * 32);
  }
  u->cap_batch = batch;
  u->cap_n = n;
}

static void conv(flgpu_unet *u, int l, int batch, int n, int ci, const float *in, float *out,
                 int relu) {
  int co = L_COUT[l];
  long total = vox(batch, n);
  unsigned gsz = grid_for((total + 15) / 16, 4);
  const short *wk = u->wf + u->woff[l];
  const float *bias = u->theta + u->boff[l];
  hipStream_t s = u->ctx->stream;
  int nt = (co + 15) / 16;
  switch (nt) {
    case 1: k_conv_mfma<1><<<gsz, 256, 0, s>>>(batch, n, ci, co, in, wk, bias, out, relu, 1); break;
    case 2: k_conv_mfma<2><<<gsz, 256, 0, s>>>(batch, n, ci, co, in, wk, bias, out, relu, 1); break;
    default: k_conv_mfma<4><<<gsz, 256, 0, s>>>(batch, n, ci, co, in, wk, bias, out, relu, 1); break;
  }
  FLGPU_HIP(hipGetLastError());
}

// gin = conv^T(gout) through the transposed weights (taps mirrored)
static void conv_gin(flgpu_unet *u, int l, int batch, int n, int ci, const float *gout,
                     float *gin) {
  int co = L_COUT[l];
  long total = vox(batch, n);
  unsigned gsz = grid_for((total + 15) / 16, 4);
  const short *wk = u->wg + u->woff[l];  // [t][c][o]: K index o contiguous
  hipStream_t s = u->ctx->stream;
  int nt = (ci + 15) / 16;
  // roles swap: "input" channels = co (K), "output" channels = ci (N)
  switch (nt) {
    case 1: k_conv_mfma<1><<<gsz, 256, 0, s>>>(batch, n, co, ci, gout, wk, nullptr, gin, 0, -1); break;
    case 2: k_conv_mfma<2><<<gsz, 256, 0, s>>>(batch, n, co, ci, gout, wk, nullptr, gin, 0, -1); break;
    case 3:
    case 4: k_conv_mfma<4><<<gsz, 256, 0, s>>>(batch, n, co, ci, gout, wk, nullptr, gin, 0, -1); break;
    default: k_conv_mfma<6><<<gsz, 256, 0, s>>>(batch, n, co, ci, gout, wk, nullptr, gin, 0, -1); break;
  }
  FLGPU_HIP(hipGetLastError());
}

static void conv_wgrad(flgpu_unet *u, int l, int batch, int n, int ci, const float *in,
                       const float *gout) {
  int co = L_COUT[l];
  long total = vox(batch, n);
  long nchunks = (total + kWgVox - 1) / kWgVox;
  const int np27 = 27 * ci * co;
  long need = nchunks * (np27 + (long)co);
  if (need > u->partial_cap) {
    dev_free(u->partial);
    u->partial = (float *)dev_alloc(sizeof(float) * need);
    u->partial_cap = need;
  }
  const int cip = (ci + 15) & ~15, cop = (co + 15) & ~15;
  FLGPU_REQUIRE(cip <= 96 && cop <= 64 && (cip / 16) * (cop / 16) <= 4 * kWgTiles &&
                    256 % cop == 0,
                "layer too wide for the weight-gradient kernel");
  hipStream_t s = u->ctx->stream;
  k_conv_wgrad<<<(unsigned)(nchunks * 27), 256, 0, s>>>(batch, n, ci, co, in, gout, u->partial);
  k_wgrad_reduce<<<grid_for(np27 + co, 256), 256, 0, s>>>(nchunks, np27, co, u->partial,
                                                          u->grad + u->woff[l],
                                                          u->grad + u->boff[l]);
  FLGPU_HIP(hipGetLastError());
}

static void refresh_bf16(flgpu_unet *u) {
  for (int l = 0; l < NL; l++) {
    int ci = l == 0 ? u->cin : L_CIN[l], co = L_COUT[l];
    long tot = 27L * ci * co;
    k_wcopies<<<grid_for(tot, 256), 256, 0, u->ctx->stream>>>(ci, co, u->theta + u->woff[l],
                                                             u->wf + u->woff[l], u->wg + u->woff[l]);
  }
  FLGPU_HIP(hipGetLastError());
}

static unsigned ew(long m) { return grid_for(m, 256); }

void unet_forward(flgpu_unet *u, int batch, int n, const float *x) {
  hipStream_t s = u->ctx->stream;
  int nl[3] = {n, n / 2, n / 4};
  conv(u, 0, batch, n, u->cin, x, u->a[0], 1);
  conv(u, 1, batch, n, 16, u->a[0], u->a[1], 1);
  k_pool<<<ew(vox(batch, nl[1]) * 16), 256, 0, s>>>(batch, n, 16, u->a[1], u->p0);
  conv(u, 2, batch, nl[1], 16, u->p0, u->a[2], 1);
  conv(u, 3, batch, nl[1], 32, u->a[2], u->a[3], 1);
  k_pool<<<ew(vox(batch, nl[2]) * 32), 256, 0, s>>>(batch, nl[1], 32, u->a[3], u->p1);
  conv(u, 4, batch, nl[2], 32, u->p1, u->a[4], 1);
  conv(u, 5, batch, nl[2], 64, u->a[4], u->a[5], 1);
  k_upcat<<<ew(vox(batch, nl[1]) * 96), 256, 0, s>>>(batch, nl[1], 64, 32, u->a[5], u->a[3], u->c1);
  conv(u, 6, batch, nl[1], 96, u->c1, u->a[6], 1);
  k_upcat<<<ew(vox(batch, n) * 48), 256, 0, s>>>(batch, n, 32, 16, u->a[6], u->a[1], u->c0);
  conv(u, 7, batch, n, 48, u->c0, u->a[7], 1);
  conv(u, 8, batch, n, 16, u->a[7], u->a[8], 0);
  FLGPU_HIP(hipGetLastError());
}

// loss (device slot) and gradients of all parameters into u->grad
void unet_backward(flgpu_unet *u, int batch, int n, const float *x, const float *t) {
  hipStream_t s = u->ctx->stream;
  int nl[3] = {n, n / 2, n / 4};
  long v0 = vox(batch, n), v1 = vox(batch, nl[1]), v2 = vox(batch, nl[2]);
  k_mse<<<256, 256, 0, s>>>(v0, u->a[8], t, u->g[8], u->lpart);
  k_sum_partials<<<1, 1, 0, s>>>(256, u->lpart, u->ctx->red_out + 8, 1.0 / (double)v0);
  conv_wgrad(u, 8, batch, n, 16, u->a[7], u->g[8]);
  conv_gin(u, 8, batch, n, 16, u->g[8], u->g[7]);
  k_relu_bwd<<<ew(v0 * 16), 256, 0, s>>>(v0 * 16, u->a[7], u->g[7]);
  conv_wgrad(u, 7, batch, n, 48, u->c0, u->g[7]);
  conv_gin(u, 7, batch, n, 48, u->g[7], u->gc0);
  k_upcat_bwd_up<<<ew(v1 * 32), 256, 0, s>>>(batch, n, 32, 16, u->gc0, u->g[6]);
  k_upcat_bwd_skip<<<ew(v0 * 16), 256, 0, s>>>(v0, 32, 16, u->gc0, u->gskip0);
  k_relu_bwd<<<ew(v1 * 32), 256, 0, s>>>(v1 * 32, u->a[6], u->g[6]);
  conv_wgrad(u, 6, batch, nl[1], 96, u->c1, u->g[6]);
  conv_gin(u, 6, batch, nl[1], 96, u->g[6], u->gc1);
  k_upcat_bwd_up<<<ew(v2 * 64), 256, 0, s>>>(batch, nl[1], 64, 32, u->gc1, u->g[5]);
  k_upcat_bwd_skip<<<ew(v1 * 32), 256, 0, s>>>(v1, 64, 32, u->gc1, u->gskip1);
  k_relu_bwd<<<ew(v2 * 64), 256, 0, s>>>(v2 * 64, u->a[5], u->g[5]);
  conv_wgrad(u, 5, batch, nl[2], 64, u->a[4], u->g[5]);
  conv_gin(u, 5, batch, nl[2], 64, u->g[5], u->g[4]);
  k_relu_bwd<<<ew(v2 * 64), 256, 0, s>>>(v2 * 64, u->a[4], u->g[4]);
  conv_wgrad(u, 4, batch, nl[2], 32, u->p1, u->g[4]);
  conv_gin(u, 4, batch, nl[2], 32, u->g[4], u->gp1);
  // a3 feeds the pool and the skip: gradients add
  k_pool_bwd<<<ew(v1 * 32), 256, 0, s>>>(batch, nl[1], 32, u->gp1, u->gskip1);
  k_relu_bwd<<<ew(v1 * 32), 256, 0, s>>>(v1 * 32, u->a[3], u->gskip1);
  conv_wgrad(u, 3, batch, nl[1], 32, u->a[2], u->gskip1);
  conv_gin(u, 3, batch, nl[1], 32, u->gskip1, u->g[2]);
  k_relu_bwd<<<ew(v1 * 32), 256, 0, s>>>(v1 * 32, u->a[2], u->g[2]);
  conv_wgrad(u, 2, batch, nl[1], 16, u->p0, u->g[2]);
  conv_gin(u, 2, batch, nl[1], 16, u->g[2], u->gp0);
  k_pool_bwd<<<ew(v0 * 16), 256, 0, s>>>(batch, n, 16, u->gp0, u->gskip0);
  k_relu_bwd<<<ew(v0 * 16), 256, 0, s>>>(v0 * 16, u->a[1], u->gskip0);
  conv_wgrad(u, 1, batch, n, 16, u->a[0], u->gskip0);
  conv_gin(u, 1, batch, n, 16, u->gskip0, u->g[0]);
  k_relu_bwd<<<ew(v0 * 16), 256, 0, s>>>(v0 * 16, u->a[0], u->g[0]);
  conv_wgrad(u, 0, batch, n, u->cin, x, u->g[0]);
  FLGPU_HIP(hipGetLastError());
}

static void check_shape(const flgpu_unet *u, const flgpu_vec *vec, int batch, int n, int ch,
                        const char *what) {
  FLGPU_REQUIRE(vec && vec->ctx == u->ctx, std::string(what) + " missing or of another context");
  FLGPU_REQUIRE(vec->prec == FLGPU_F32, std::string(what) + " must be an FP32 vector");
  FLGPU_REQUIRE(vec->n == vox(batch, n) * ch, std::string(what) + " has the wrong length");
}

}  // namespace flgpu

using namespace flgpu;

extern "C" int flgpu_unet_create(flgpu_ctx *ctx, int cin, unsigned long seed, flgpu_unet **out) {
  return guard([&] {
    FLGPU_REQUIRE(ctx && out, "arguments must not be NULL");
    FLGPU_REQUIRE(cin >= 1 && cin <= 16, "cin must be in 1..16");
    CTX_LOCK(ctx);
    flgpu_unet *u = new flgpu_unet;
    u->ctx = ctx;
    u->cin = cin;
    long off = 0;
    std::vector<float> h;
    std::mt19937_64 rng(seed);
    for (int l = 0; l < NL; l++) {
      int ci = l == 0 ? cin : L_CIN[l], co = L_COUT[l];
      u->woff[l] = off;
      long nw = 27L * ci * co;
      // He-normal initialisation for ReLU layers
      std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (27.0 * ci)));
      for (long i = 0; i < nw; i++) h.push_back((float)(l == NL - 1 ? 0.1 * nd(rng) : nd(rng)));
      off += nw;
      u->boff[l] = off;
      for (int i = 0; i < co; i++) h.push_back(0.f);
      off += co;
    }
    u->np = off;
    u->theta = (float *)dev_alloc(sizeof(float) * off);
    u->m = (float *)dev_alloc(sizeof(float) * off);
    u->v = (float *)dev_alloc(sizeof(float) * off);
    u->grad = (float *)dev_alloc(sizeof(float) * off);
    u->wf = (short *)dev_alloc(sizeof(short) * off);
    u->wg = (short *)dev_alloc(sizeof(short) * off);
    u->lpart = (double *)dev_alloc(sizeof(double) * 256);
    FLGPU_HIP(hipMemcpyAsync(u->theta, h.data(), sizeof(float) * off, hipMemcpyHostToDevice,
                             ctx->stream));
    vfill(ctx, off, FLGPU_F32, u->m, 0.0);
    vfill(ctx, off, FLGPU_F32, u->v, 0.0);
    vfill(ctx, off, FLGPU_F32, u->grad, 0.0);
    refresh_bf16(u);
    FLGPU_HIP(hipStreamSynchronize(ctx->stream));
    *out = u;
    return 0;
  });
}

extern "C" int flgpu_unet_destroy(flgpu_unet *u) {
  return guard([&] {
    if (!u) return 0;
    CTX_LOCK(u->ctx);
    FLGPU_HIP(hipStreamSynchronize(u->ctx->stream));
    unet_free_acts(u);
    for (void *p : {(void *)u->theta, (void *)u->m, (void *)u->v, (void *)u->grad, (void *)u->wf,
                    (void *)u->wg, (void *)u->partial, (void *)u->lpart})
      dev_free(p);
    delete u;
    return 0;
  });
}

extern "C" int flgpu_unet_nparams(const flgpu_unet *u, long *np) {
  return guard([&] {
    FLGPU_REQUIRE(u && np, "arguments must not be NULL");
    *np = u->np;
    return 0;
  });
}

// parameters (and Adam state) as doubles: checkpoint / restore
extern "C" int flgpu_unet_get_params(const flgpu_unet *u, double *theta, double *m, double *v,
                                     int *step) {
  return guard([&] {
    FLGPU_REQUIRE(u, "network must not be NULL");
    CTX_LOCK(u->ctx);
    std::vector<float> h(u->np);
    const float *src[3] = {u->theta, u->m, u->v};
    double *dst[3] = {theta, m, v};
    for (int k = 0; k < 3; k++) {
      if (!dst[k]) continue;
      FLGPU_HIP(hipMemcpy(h.data(), src[k], sizeof(float) * u->np, hipMemcpyDeviceToHost));
      for (long i = 0; i < u->np; i++) dst[k][i] = h[i];
    }
    if (step) *step = u->step;
    return 0;
  });
}

extern "C" int flgpu_unet_set_params(flgpu_unet *u, const double *theta, const double *m,
                                     const double *v, int step) {
  return guard([&] {
    FLGPU_REQUIRE(u && theta, "network and theta must not be NULL");
    CTX_LOCK(u->ctx);
    FLGPU_HIP(hipStreamSynchronize(u->ctx->stream));
    std::vector<float> h(u->np);
    const double *src[3] = {theta, m, v};
    float *dst[3] = {u->theta, u->m, u->v};
    for (int k = 0; k < 3; k++) {
      if (src[k]) {
        for (long i = 0; i < u->np; i++) h[i] = (float)src[k][i];
        FLGPU_HIP(hipMemcpy(dst[k], h.data(), sizeof(float) * u->np, hipMemcpyHostToDevice));
      } else if (k > 0) {
        vfill(u->ctx, u->np, FLGPU_F32, dst[k], 0.0);
      }
    }
    u->step = step;
    refresh_bf16(u);
    FLGPU_HIP(hipStreamSynchronize(u->ctx->stream));
    return 0;
  });
}

extern "C" int flgpu_unet_forward(flgpu_unet *u, int batch, int n, const flgpu_vec *x,
                                  flgpu_vec *y) {
  return guard([&] {
    FLGPU_REQUIRE(u, "network must not be NULL");
    FLGPU_REQUIRE(batch >= 1 && n >= 4 && n % 4 == 0, "n must be a multiple of 4");
    check_shape(u, x, batch, n, u->cin, "x");
    check_shape(u, y, batch, n, 1, "y");
    CTX_LOCK(u->ctx);
    unet_alloc_acts(u, batch, n, false);
    unet_forward(u, batch, n, (const float *)x->d);
    vcopy(u->ctx, vox(batch, n), FLGPU_F32, u->a[8], FLGPU_F32, y->d);
    return 0;
  });
}

// loss and gradient without an update (gradient check)
extern "C" int flgpu_unet_loss_grad(flgpu_unet *u, int batch, int n, const flgpu_vec *x,
                                    const flgpu_vec *t, double *loss, double *grad) {
  return guard([&] {
    FLGPU_REQUIRE(u && loss, "arguments must not be NULL");
    FLGPU_REQUIRE(batch >= 1 && n >= 4 && n % 4 == 0, "n must be a multiple of 4");
    check_shape(u, x, batch, n, u->cin, "x");
    check_shape(u, t, batch, n, 1, "t");
    CTX_LOCK(u->ctx);
    unet_alloc_acts(u, batch, n, true);
    unet_forward(u, batch, n, (const float *)x->d);
    unet_backward(u, batch, n, (const float *)x->d, (const float *)t->d);
    fetch_slots(u->ctx, 9);
    *loss = u->ctx->host_out[8];
    if (grad) {
      std::vector<float> h(u->np);
      FLGPU_HIP(hipMemcpy(h.data(), u->grad, sizeof(float) * u->np, hipMemcpyDeviceToHost));
      for (long i = 0; i < u->np; i++) grad[i] = h[i];
    }
    return 0;
  });
}

// one Adam training step on a batch: returns the batch loss before the update
extern "C" int flgpu_unet_train_step(flgpu_unet *u, int batch, int n, const flgpu_vec *x,
                                     const flgpu_vec *t, double lr, double *loss) {
  return guard([&] {
    FLGPU_REQUIRE(u, "network must not be NULL");
    FLGPU_REQUIRE(batch >= 1 && n >= 4 && n % 4 == 0, "n must be a multiple of 4");
    check_shape(u, x, batch, n, u->cin, "x");
    check_shape(u, t, batch, n, 1, "t");
    CTX_LOCK(u->ctx);
    unet_alloc_acts(u, batch, n, true);
    unet_forward(u, batch, n, (const float *)x->d);
    unet_backward(u, batch, n, (const float *)x->d, (const float *)t->d);
    u->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float c1 = 1.f - std::pow(b1, (float)u->step), c2 = 1.f - std::pow(b2, (float)u->step);
    k_adam<<<grid_for(u->np, 256), 256, 0, u->ctx->stream>>>(u->np, u->theta, u->grad, u->m, u->v,
                                                             (float)lr, b1, b2, eps, c1, c2);
    FLGPU_HIP(hipGetLastError());
    refresh_bf16(u);
    if (loss) {
      fetch_slots(u->ctx, 9);
      *loss = u->ctx->host_out[8];
    }
    return 0;
  });
}

// ---- lattice <-> network grid -------------------------------------------------------------

namespace flgpu {

// network input of sample s: channel 0 = r / rms(r) on nodes (z,y,x) < n of
// the (n+1)^3 lattice, channel 1 = log(kappa) / 4 of the cell with that lower
// corner (0 without kappa).  rms from the dot in *ss: sqrt(ss / n^3).
template <class T>
__global__ void k_pack(int n, const T *__restrict__ r, const double *__restrict__ kappa,
                       const double *ss, float *x, long sample) {
  long nv = (long)n * n * n;
  long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  int xx = (int)(v % n), yy = (int)((v / n) % n), zz = (int)(v / ((long)n * n));
  long idx = ((long)zz * (n + 1) + yy) * (n + 1) + xx;
  double rms = sqrt(*ss / (double)nv);
  double inv = rms > 0.0 ? 1.0 / rms : 0.0;
  float *o = x + (sample * nv + v) * 2;
  o[0] = (float)((double)r[idx] * inv);
  o[1] = kappa ? (float)(log(kappa[v]) * 0.25) : 0.f;
}

// target of sample s: e / rms(e) on the network grid
__global__ void k_pack_target(int n, const double *__restrict__ e, const double *ss, float *t,
                              long sample) {
  long nv = (long)n * n * n;
  long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  int xx = (int)(v % n), yy = (int)((v / n) % n), zz = (int)(v / ((long)n * n));
  long idx = ((long)zz * (n + 1) + yy) * (n + 1) + xx;
  double rms = sqrt(*ss / (double)nv);
  t[sample * nv + v] = (float)(rms > 0.0 ? e[idx] / rms : 0.0);
}

// lattice vector from the network output (zero on Dirichlet nodes)
template <class T>
__global__ void k_unpack(int n, const float *__restrict__ y, T *q) {
  long M = n + 1;
  long N = M * M * M;
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  int xx = (int)(idx % M), yy = (int)((idx / M) % M), zz = (int)(idx / (M * M));
  bool interior = xx > 0 && yy > 0 && zz > 0 && xx < n && yy < n && zz < n;
  q[idx] = interior ? (T)y[((long)zz * n + yy) * n + xx] : (T)0;
}

// q += beta c with beta = (c . r1) / (c . A c), the energy-optimal step along c, when that step
// adds at least min_gain of the energy of the cycle's own step: (c . r1)^2 / (c . A c) >=
// min_gain (z1 . r).  d = {c . r1, c . A c, z1 . r}; *flag = 1 when added, else 0.
template <class T>
__global__ void k_add_opt(long N, const double *__restrict__ d, double min_gain, double *flag,
                          const T *c, T *q) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const double num = d[0], den = d[1], zr = d[2];
  const bool add = den > 0.0 && zr > 0.0 && num * num / den >= min_gain * zr;
  if (i == 0) *flag = add ? 1.0 : 0.0;
  if (i >= N || !add) return;
  const T beta = (T)(num / den);
  q[i] += beta * c[i];
}

// The network's correction inside flexible CG.  After the V-cycle z1 = q, the residual
// r1 = r - A z1 is taken to the level the network runs on (the cycle's own restrictions, in
// the cycle's precision), the network maps it to an error estimate there, the cycle's
// prolongations bring that back (c), and q += beta c with the energy-optimal beta, when that
// step carries enough energy (k_add_opt).  On the top level this is z = z1 + beta N(r1).
struct UnetHook : LearnedHook {
  flgpu_ctx *ctx;
  flgpu_mg *mg;
  const flgpu_mat *A;
  flgpu_unet *u;
  const flgpu_vec *kappa;
  int n, lc, top, mprec;
  long N;
  int prec;
  void *r1 = nullptr, *c = nullptr, *Ac = nullptr;
  std::vector<void *> lv;  // levels lc..top in the cycle's precision (below the top)
  float *xin = nullptr;
  ~UnetHook() override {
    dev_free(r1);
    dev_free(c);
    dev_free(Ac);
    for (void *p : lv) dev_free(p);
    dev_free(xin);
  }
  void correct(void *r, void *q, double min_gain, int flag) override {
    hipStream_t s = ctx->stream;
    FLGPU_REQUIRE(prec == FLGPU_F64, "the learned correction runs inside an FP64 outer iteration");
    spmv(A, -1.0, q, 1.0, r, r1);
    long nv = (long)n * n * n;
    const double *kd = kappa ? (const double *)kappa->d : nullptr;
    unet_alloc_acts(u, 1, n, false);
    if (lc == top) {
      vdot_dev(ctx, N, prec, r1, r1, 10);
      k_pack<double><<<grid_for(nv, 256), 256, 0, s>>>(n, (const double *)r1, kd,
                                                       ctx->red_out + 10, xin, 0);
      unet_forward(u, 1, n, xin);
      k_unpack<double><<<grid_for(N, 256), 256, 0, s>>>(n, u->a[8], (double *)c);
    } else {
      vcopy(ctx, N, FLGPU_F64, r1, mprec, lv[top - lc]);
      for (int l = top; l > lc; l--) mg_restrict(mg, l, lv[l - lc], lv[l - 1 - lc]);
      long Nc = mg_level_rows(mg, lc);
      vdot_dev(ctx, Nc, mprec, lv[0], lv[0], 10);
      if (mprec == FLGPU_F64)
        k_pack<double><<<grid_for(nv, 256), 256, 0, s>>>(n, (const double *)lv[0], kd,
                                                         ctx->red_out + 10, xin, 0);
      else
        k_pack<float><<<grid_for(nv, 256), 256, 0, s>>>(n, (const float *)lv[0], kd,
                                                        ctx->red_out + 10, xin, 0);
      unet_forward(u, 1, n, xin);
      if (mprec == FLGPU_F64)
        k_unpack<double><<<grid_for(Nc, 256), 256, 0, s>>>(n, u->a[8], (double *)lv[0]);
      else
        k_unpack<float><<<grid_for(Nc, 256), 256, 0, s>>>(n, u->a[8], (float *)lv[0]);
      for (int l = lc + 1; l <= top; l++) mg_prolong(mg, l, lv[l - 1 - lc], lv[l - lc]);
      vcopy(ctx, N, mprec, lv[top - lc], FLGPU_F64, c);
    }
    spmv(A, 1.0, c, 0.0, Ac, Ac);
    {  // c . r1, c . A c and the cycle step's z1 . r, one pass
      const void *dx[3] = {c, c, q}, *dy[3] = {r1, Ac, r};
      const int slots[3] = {11, 12, 13};
      vdot_multi_dev(ctx, N, prec, 3, dx, dy, slots);
    }
    k_add_opt<double><<<grid_for(N, 256), 256, 0, s>>>(N, ctx->red_out + 11, min_gain,
                                                       ctx->red_out + flag, (const double *)c,
                                                       (double *)q);
    FLGPU_HIP(hipGetLastError());
  }
};

std::unique_ptr<LearnedHook> learned_hook(flgpu_mg *mg, const flgpu_mat *A) {
  FLGPU_REQUIRE(mg, "the learned correction needs a hierarchy");
  flgpu_unet *u;
  const flgpu_vec *kappa;
  int n;
  mg_unet(mg, &u, &kappa, &n);
  FLGPU_REQUIRE(u, "no learned network is attached to the hierarchy (flgpu_mg_attach_unet)");
  FLGPU_REQUIRE(n % 4 == 0, "the network grid needs n divisible by 4");
  std::unique_ptr<UnetHook> h(new UnetHook);
  h->ctx = A->ctx;
  h->mg = mg;
  h->A = A;
  h->u = u;
  h->kappa = kappa;
  h->n = n;
  h->N = A->nrows;
  h->prec = A->prec;
  h->lc = mg_level_cells(mg, n);
  h->top = mg_top_level(mg);
  h->mprec = mg_precision(mg);
  FLGPU_REQUIRE(h->lc >= 0, "the network's grid is no level of the hierarchy");
  FLGPU_REQUIRE(mg_level_rows(mg, h->lc) == (long)(n + 1) * (n + 1) * (n + 1),
                "the network's level is not a 3D Q1 lattice of n cells per axis");
  FLGPU_REQUIRE(mg_level_rows(mg, h->top) == h->N, "A does not match the hierarchy");
  size_t b = prec_bytes(A->prec) * h->N;
  h->r1 = dev_alloc(b);
  h->c = dev_alloc(b);
  h->Ac = dev_alloc(b);
  if (h->lc < h->top)
    for (int l = h->lc; l <= h->top; l++)
      h->lv.push_back(dev_alloc(prec_bytes(h->mprec) * mg_level_rows(mg, l)));
  h->xin = (float *)dev_alloc(sizeof(float) * 2 * (long)n * n * n);
  return std::unique_ptr<LearnedHook>(h.release());
}

}  // namespace flgpu

extern "C" int flgpu_unet_pack(const flgpu_vec *r, const flgpu_vec *kappa, int n, flgpu_vec *x,
                               long sample, int target) {
  return guard([&] {
    FLGPU_REQUIRE(r && x, "arguments must not be NULL");
    FLGPU_REQUIRE(r->prec == FLGPU_F64 && r->n == (long)(n + 1) * (n + 1) * (n + 1),
                  "r must be an FP64 lattice vector of (n+1)^3 values");
    FLGPU_REQUIRE(!kappa || (kappa->prec == FLGPU_F64 && kappa->n == (long)n * n * n),
                  "kappa must be an FP64 vector of n^3 cell values");
    long nv = (long)n * n * n;
    int ch = target ? 1 : 2;
    FLGPU_REQUIRE(x->prec == FLGPU_F32 && sample >= 0 && (sample + 1) * nv * ch <= x->n,
                  "x too small for this sample");
    flgpu_ctx *ctx = r->ctx;
    CTX_LOCK(ctx);
    vdot_dev(ctx, r->n, FLGPU_F64, r->d, r->d, 14);
    if (target)
      k_pack_target<<<grid_for(nv, 256), 256, 0, ctx->stream>>>(n, (const double *)r->d,
                                                                ctx->red_out + 14, (float *)x->d,
                                                                sample);
    else
      k_pack<double><<<grid_for(nv, 256), 256, 0, ctx->stream>>>(
          n, (const double *)r->d, kappa ? (const double *)kappa->d : nullptr, ctx->red_out + 14,
          (float *)x->d, sample);
    FLGPU_HIP(hipGetLastError());
    return 0;
  });
}

// dst[k] = src rows idx[k] (rows of row_len values), k = 0..count-1
extern "C" int flgpu_vec_gather_rows(const flgpu_vec *src, long row_len, const long *idx,
                                     int count, flgpu_vec *dst) {
  return guard([&] {
    FLGPU_REQUIRE(src && dst && idx && row_len > 0 && count >= 0, "bad arguments");
    FLGPU_REQUIRE(src->prec == dst->prec && src->ctx == dst->ctx, "precision/context mismatch");
    FLGPU_REQUIRE(dst->n >= (long)count * row_len, "dst too small");
    long rows = src->n / row_len;
    flgpu_ctx *ctx = src->ctx;
    CTX_LOCK(ctx);
    size_t eb = prec_bytes(src->prec);
    for (int k = 0; k < count; k++) {
      FLGPU_REQUIRE(idx[k] >= 0 && idx[k] < rows, "row index out of range");
      FLGPU_HIP(hipMemcpyAsync((char *)dst->d + (size_t)k * row_len * eb,
                               (const char *)src->d + (size_t)idx[k] * row_len * eb, row_len * eb,
                               hipMemcpyDeviceToDevice, ctx->stream));
    }
    return 0;
  });
}
