// Context, errors, device vectors and BLAS-1 kernels of libflgpu.so.
//
// Vector operations restate femlisp's blas on <sparse-vector>
// (src/matlisp/vector.lisp:224-326, src/matlisp/sparse-vector.lisp:248-280):
// fill!/x<-0, copy!, scal!, axpy!, dot, norm.  Dots are reduced in FP64 in a
// fixed order (fixed grid, fixed tree) so results are bit-reproducible.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "flgpu_internal.h"

namespace flgpu {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

void *dev_alloc(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) bytes = 8;
  FLGPU_HIP(hipMalloc(&p, bytes));
  return p;
}

void dev_free(void *p) {
  if (p) (void)hipFree(p);
}

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;
constexpr int kRedSlots = 64;

// Scratch of flgpu_vec_dot_noalloc: the per-block partials of one dot (kRedBlocks values)
// and its result, 8.2 KB on every device the library runs on, held until the process ends.
// Host code never passes this name to a kernel (on the host it is a stand-in, not a device
// address): it asks each device for the array's address once (scratch_on) and uses that.
static __device__ __align__(256) double scratch[kRedBlocks + 32];

// ---- kernels ------------------------------------------------------------------

template <class T>
__global__ void k_fill(long n, T *x, T v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] = v;
}

template <class S, class D>
__global__ void k_copy(long n, const S *x, D *y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = (D)x[i];
}

template <class T>
__global__ void k_scal(long n, T a, T *x) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] = a * x[i];
}

template <class T>
__global__ void k_axpy(long n, T a, const T *x, T *y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = a * x[i] + y[i];
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// per-block partial of sum x_i y_i in FP64
template <class T>
__global__ __launch_bounds__(kRedThreads) void k_dot_partial(long n, const T *x, const T *y,
                                                              double *partial) {
  __shared__ double s[kRedThreads / 64];
  double acc = 0.0;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  // four strides' loads in flight together, summed in stride order
  for (; i + 3 * stride < n; i += 4 * stride) {
    T xv[4], yv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      xv[u] = x[i + u * stride];
      yv[u] = y[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) acc += (double)xv[u] * (double)yv[u];
  }
  for (; i < n; i += stride) acc += (double)x[i] * (double)y[i];
  acc = wave_sum(acc);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) s[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kRedThreads / 64; k++) t += s[k];
    partial[blockIdx.x] = t;
  }
}

// up to three dots in one pass over the vectors: dot k's per-block partial is formed exactly
// as k_dot_partial forms it (same grid, same order), into partial[k * gridDim.x + block]
template <class T, int K>
__global__ __launch_bounds__(kRedThreads) void k_dot_partial_multi(long n, const T *x0, const T *y0,
                                                                    const T *x1, const T *y1,
                                                                    const T *x2, const T *y2,
                                                                    double *partial) {
  __shared__ double s[K][kRedThreads / 64];
  const T *xs[3] = {x0, x1, x2}, *ys[3] = {y0, y1, y2};
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0.0;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  // four strides' loads in flight together; each dot still sums in stride order
  for (; i + 3 * stride < n; i += 4 * stride) {
    T xv[4][K], yv[4][K];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int k = 0; k < K; k++) {
        xv[u][k] = xs[k][i + u * stride];
        yv[u][k] = ys[k][i + u * stride];
      }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int k = 0; k < K; k++) acc[k] += (double)xv[u][k] * (double)yv[u][k];
  }
  for (; i < n; i += stride)
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] += (double)xs[k][i] * (double)ys[k][i];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; k++) {
    double v = wave_sum(acc[k]);
    if (lane == 0) s[k][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double t = 0.0;
    for (int j = 0; j < kRedThreads / 64; j++) t += s[threadIdx.x][j];
    partial[(long)threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// The CG/FCG step's dots q.r, r.r (and q.a for K = 3) with each vector loaded once: the same
// products summed in the same order as k_dot_partial_multi over (q, r, q) x (r, r, a)
template <typename T, int K, typename TQ = T>
__global__ __launch_bounds__(kRedThreads) void k_dot_partial_qr(long n, const TQ *q, const T *r,
                                                                const T *a, double *partial) {
  __shared__ double s[K][kRedThreads / 64];
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0.0;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    TQ qv[4];
    T rv[4], av[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      qv[u] = q[i + u * stride];
      rv[u] = r[i + u * stride];
      if (K == 3) av[u] = a[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      acc[0] += (double)qv[u] * (double)rv[u];
      acc[1] += (double)rv[u] * (double)rv[u];
      if (K == 3) acc[K - 1] += (double)qv[u] * (double)av[u];
    }
  }
  for (; i < n; i += stride) {
    double qd = (double)q[i], rd = (double)r[i];
    acc[0] += qd * rd;
    acc[1] += rd * rd;
    if (K == 3) acc[K - 1] += qd * (double)a[i];
  }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; k++) {
    double v = wave_sum(acc[k]);
    if (lane == 0) s[k][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double t = 0.0;
    for (int j = 0; j < kRedThreads / 64; j++) t += s[threadIdx.x][j];
    partial[(long)threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// block k reduces dot k's partials exactly as k_dot_final does, into out[slot[k]]
__global__ __launch_bounds__(kRedBlocks) void k_dot_final_multi(const double *partial, int np,
                                                                 double *out, int s0, int s1,
                                                                 int s2) {
  __shared__ double s[kRedBlocks / 64];
  const double *pk = partial + (long)blockIdx.x * np;
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += pk[i];
  acc = wave_sum(acc);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) s[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x / 64); k++) t += s[k];
    int slot = blockIdx.x == 0 ? s0 : (blockIdx.x == 1 ? s1 : s2);
    out[slot] = t;
  }
}

__global__ __launch_bounds__(kRedBlocks) void k_dot_final(const double *partial, int np,
                                                           double *out) {
  __shared__ double s[kRedBlocks / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += partial[i];
  acc = wave_sum(acc);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) s[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x / 64); k++) t += s[k];
    *out = t;
  }
}

// ---- host wrappers --------------------------------------------------------------

static unsigned ew_grid(long n) {
  long g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  return (unsigned)(g < 1 ? 1 : g);
}

void vfill(flgpu_ctx *ctx, long n, int prec, void *x, double v) {
  if (n <= 0) return;
  if (prec == FLGPU_F64)
    k_fill<double><<<ew_grid(n), 256, 0, ctx->stream>>>(n, (double *)x, v);
  else
    k_fill<float><<<ew_grid(n), 256, 0, ctx->stream>>>(n, (float *)x, (float)v);
  FLGPU_HIP(hipGetLastError());
}

void vcopy(flgpu_ctx *ctx, long n, int sp, const void *src, int dp, void *dst) {
  if (n <= 0 || (src == dst && sp == dp)) return;
  unsigned g = ew_grid(n);
  hipStream_t s = ctx->stream;
  if (sp == FLGPU_F64 && dp == FLGPU_F64)
    k_copy<double, double><<<g, 256, 0, s>>>(n, (const double *)src, (double *)dst);
  else if (sp == FLGPU_F64)
    k_copy<double, float><<<g, 256, 0, s>>>(n, (const double *)src, (float *)dst);
  else if (dp == FLGPU_F64)
    k_copy<float, double><<<g, 256, 0, s>>>(n, (const float *)src, (double *)dst);
  else
    k_copy<float, float><<<g, 256, 0, s>>>(n, (const float *)src, (float *)dst);
  FLGPU_HIP(hipGetLastError());
}

void vscal(flgpu_ctx *ctx, long n, int prec, double a, void *x) {
  if (n <= 0) return;
  if (prec == FLGPU_F64)
    k_scal<double><<<ew_grid(n), 256, 0, ctx->stream>>>(n, a, (double *)x);
  else
    k_scal<float><<<ew_grid(n), 256, 0, ctx->stream>>>(n, (float)a, (float *)x);
  FLGPU_HIP(hipGetLastError());
}

void vaxpy(flgpu_ctx *ctx, long n, int prec, double a, const void *x, void *y) {
  if (n <= 0) return;
  if (prec == FLGPU_F64)
    k_axpy<double><<<ew_grid(n), 256, 0, ctx->stream>>>(n, a, (const double *)x, (double *)y);
  else
    k_axpy<float><<<ew_grid(n), 256, 0, ctx->stream>>>(n, (float)a, (const float *)x, (float *)y);
  FLGPU_HIP(hipGetLastError());
}

void vdot_dev(flgpu_ctx *ctx, long n, int prec, const void *x, const void *y, int slot) {
  FLGPU_REQUIRE(slot >= 0 && slot < kRedSlots, "reduction slot out of range");
  if (prec == FLGPU_F64)
    k_dot_partial<double><<<kRedBlocks, kRedThreads, 0, ctx->stream>>>(
        n, (const double *)x, (const double *)y, ctx->red_partial);
  else
    k_dot_partial<float><<<kRedBlocks, kRedThreads, 0, ctx->stream>>>(
        n, (const float *)x, (const float *)y, ctx->red_partial);
  k_dot_final<<<1, kRedBlocks, 0, ctx->stream>>>(ctx->red_partial, kRedBlocks,
                                                 ctx->red_out + slot);
  FLGPU_HIP(hipGetLastError());
}

// the sum of np partials (in k_dot_final's order) into reduction slot `slot`
void dot_final(flgpu_ctx *ctx, const double *partial, int np, int slot) {
  FLGPU_REQUIRE(slot >= 0 && slot < kRedSlots, "reduction slot out of range");
  k_dot_final<<<1, kRedBlocks, 0, ctx->stream>>>(partial, np, ctx->red_out + slot);
  FLGPU_HIP(hipGetLastError());
}

// dots (x[k], y[k]) into slots[k], k < count <= 3, in one pass; each equals vdot_dev's result
void vdot_multi_dev(flgpu_ctx *ctx, long n, int prec, int count, const void *const *x,
                    const void *const *y, const int *slots) {
  FLGPU_REQUIRE(count >= 1 && count <= 3, "vdot_multi_dev takes 1 to 3 dots");
  for (int k = 0; k < count; k++)
    FLGPU_REQUIRE(slots[k] >= 0 && slots[k] < kRedSlots, "reduction slot out of range");
  const void *xx[3] = {x[0], count > 1 ? x[1] : x[0], count > 2 ? x[2] : x[0]};
  const void *yy[3] = {y[0], count > 1 ? y[1] : y[0], count > 2 ? y[2] : y[0]};
  hipStream_t st = ctx->stream;
  // (q, r, q) x (r, r, a): the CG/FCG step's dots, each vector loaded once
  const bool qr = count >= 2 && xx[1] == yy[0] && yy[1] == yy[0] && (count < 3 || xx[2] == xx[0]);
#define FLGPU_DOTS(T, K)                                                                      \
  k_dot_partial_multi<T, K><<<kRedBlocks, kRedThreads, 0, st>>>(                              \
      n, (const T *)xx[0], (const T *)yy[0], (const T *)xx[1], (const T *)yy[1],              \
      (const T *)xx[2], (const T *)yy[2], ctx->red_partial)
#define FLGPU_QR(T, K)                                                                        \
  k_dot_partial_qr<T, K><<<kRedBlocks, kRedThreads, 0, st>>>(                                 \
      n, (const T *)xx[0], (const T *)yy[0], (const T *)yy[2], ctx->red_partial)
  if (qr) {
    if (prec == FLGPU_F64) {
      if (count == 2) FLGPU_QR(double, 2);
      else FLGPU_QR(double, 3);
    } else {
      if (count == 2) FLGPU_QR(float, 2);
      else FLGPU_QR(float, 3);
    }
  } else if (prec == FLGPU_F64) {
    if (count == 1) FLGPU_DOTS(double, 1);
    else if (count == 2) FLGPU_DOTS(double, 2);
    else FLGPU_DOTS(double, 3);
  } else {
    if (count == 1) FLGPU_DOTS(float, 1);
    else if (count == 2) FLGPU_DOTS(float, 2);
    else FLGPU_DOTS(float, 3);
  }
#undef FLGPU_DOTS
#undef FLGPU_QR
  k_dot_final_multi<<<count, kRedBlocks, 0, st>>>(ctx->red_partial, kRedBlocks, ctx->red_out,
                                                  slots[0], count > 1 ? slots[1] : slots[0],
                                                  count > 2 ? slots[2] : slots[0]);
  FLGPU_HIP(hipGetLastError());
}

// the FCG step's dots z.r, r.r, z.a with z in FP32 (an FP32 cycle's output) and r, a in FP64:
// the same products, summed in the same order, as vdot_multi_dev over the FP64 copy of z
void vdot_qr_f32(flgpu_ctx *ctx, long n, const float *z, const double *r, const double *a,
                 const int *slots) {
  for (int k = 0; k < 3; k++)
    FLGPU_REQUIRE(slots[k] >= 0 && slots[k] < kRedSlots, "reduction slot out of range");
  k_dot_partial_qr<double, 3, float><<<kRedBlocks, kRedThreads, 0, ctx->stream>>>(
      n, z, r, a, ctx->red_partial);
  k_dot_final_multi<<<3, kRedBlocks, 0, ctx->stream>>>(ctx->red_partial, kRedBlocks, ctx->red_out,
                                                       slots[0], slots[1], slots[2]);
  FLGPU_HIP(hipGetLastError());
}

void fetch_slots(flgpu_ctx *ctx, int nslots) {
  FLGPU_HIP(hipMemcpyAsync(ctx->host_out, ctx->red_out, sizeof(double) * nslots,
                           hipMemcpyDeviceToHost, ctx->stream));
  FLGPU_HIP(hipStreamSynchronize(ctx->stream));
}

void fetch_slots_async(flgpu_ctx *ctx, int nslots) {
  FLGPU_HIP(hipMemcpyAsync(ctx->host_out, ctx->red_out, sizeof(double) * nslots,
                           hipMemcpyDeviceToHost, ctx->stream));
  FLGPU_HIP(hipEventRecord(ctx->ev_fetch, ctx->stream));
}

void fetch_wait(flgpu_ctx *ctx) { FLGPU_HIP(hipEventSynchronize(ctx->ev_fetch)); }

// The address of `scratch` on `device`, which must be current: hipGetSymbolAddress on the
// first call for that device, then the kept value (each device holds its own copy).
static double *scratch_on(int device) {
  static std::mutex mu;
  static std::vector<double *> ptr;
  std::lock_guard<std::mutex> g(mu);
  FLGPU_REQUIRE(device >= 0, "bad device index");
  if ((int)ptr.size() <= device) ptr.resize(device + 1, nullptr);
  if (!ptr[device]) {
    void *p = nullptr;
    FLGPU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    FLGPU_REQUIRE(p != nullptr, "no device address for the scratch array");
    ptr[device] = (double *)p;
  }
  return ptr[device];
}

double vdot_host(flgpu_ctx *ctx, long n, int prec, const void *x, const void *y) {
  vdot_dev(ctx, n, prec, x, y, 0);
  fetch_slots(ctx, 1);
  return ctx->host_out[0];
}

flgpu_vec *vec_new(flgpu_ctx *ctx, long n, int prec) {
  FLGPU_REQUIRE(n >= 0, "vector length must be >= 0");
  FLGPU_REQUIRE(prec == FLGPU_F64 || prec == FLGPU_F32, "precision must be FLGPU_F64 or FLGPU_F32");
  flgpu_vec *v = new flgpu_vec;
  v->ctx = ctx;
  v->n = n;
  v->prec = prec;
  v->d = dev_alloc(n * prec_bytes(prec));
  vfill(ctx, n, prec, v->d, 0.0);
  return v;
}

void vec_free(flgpu_vec *v) {
  if (!v) return;
  dev_free(v->d);
  delete v;
}

}  // namespace flgpu

using namespace flgpu;

// ---- C-ABI ---------------------------------------------------------------------

extern "C" int flgpu_last_error(char *buf, int len) {
  if (!buf || len <= 0) return -1;
  std::snprintf(buf, (size_t)len, "%s", g_last_error.c_str());
  return 0;
}

extern "C" int flgpu_ctx_create(int device, flgpu_ctx **out) {
  return guard([&] {
    FLGPU_REQUIRE(out != nullptr, "out must not be NULL");
    int count = 0;
    FLGPU_HIP(hipGetDeviceCount(&count));
    FLGPU_REQUIRE(device >= 0 && device < count, "no such device");
    FLGPU_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    FLGPU_HIP(hipGetDeviceProperties(&prop, device));
    flgpu_ctx *c = new flgpu_ctx;
    c->device = device;
    c->cus = prop.multiProcessorCount;
    c->name = prop.gcnArchName;
    FLGPU_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->red_partial = (double *)dev_alloc(sizeof(double) * 3 * kRedBlocks);
    c->red_out = (double *)dev_alloc(sizeof(double) * kRedSlots);
    FLGPU_HIP(hipHostMalloc((void **)&c->host_out, sizeof(double) * kRedSlots, hipHostMallocDefault));
    FLGPU_HIP(hipEventCreate(&c->ev_start));
    FLGPU_HIP(hipEventCreate(&c->ev_stop));
    FLGPU_HIP(hipEventCreateWithFlags(&c->ev_fetch, hipEventDisableTiming));
    *out = c;
    return 0;
  });
}

extern "C" int flgpu_ctx_destroy(flgpu_ctx *ctx) {
  return guard([&] {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    dev_free(ctx->red_partial);
    dev_free(ctx->red_out);
    dev_free(ctx->solve_buf);
    if (ctx->host_out) (void)hipHostFree(ctx->host_out);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    if (ctx->ev_fetch) (void)hipEventDestroy(ctx->ev_fetch);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
  });
}

extern "C" int flgpu_ctx_synchronize(flgpu_ctx *ctx) {
  return guard([&] {
    FLGPU_REQUIRE(ctx, "ctx must not be NULL");
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    FLGPU_HIP(hipSetDevice(ctx->device));
    FLGPU_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
  });
}

extern "C" int flgpu_ctx_info(flgpu_ctx *ctx, char *name, int len, int *cus, long *mem) {
  return guard([&] {
    FLGPU_REQUIRE(ctx, "ctx must not be NULL");
    if (name && len > 0) std::snprintf(name, (size_t)len, "%s", ctx->name.c_str());
    if (cus) *cus = ctx->cus;
    if (mem) {
      size_t fr = 0, tot = 0;
      FLGPU_HIP(hipSetDevice(ctx->device));
      FLGPU_HIP(hipMemGetInfo(&fr, &tot));
      *mem = (long)tot;
    }
    return 0;
  });
}

extern "C" int flgpu_ctx_timer_start(flgpu_ctx *ctx) {
  return guard([&] {
    FLGPU_REQUIRE(ctx, "ctx must not be NULL");
    CTX_LOCK(ctx);
    FLGPU_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    return 0;
  });
}

extern "C" int flgpu_ctx_timer_stop(flgpu_ctx *ctx, double *ms) {
  return guard([&] {
    FLGPU_REQUIRE(ctx && ms, "arguments must not be NULL");
    CTX_LOCK(ctx);
    FLGPU_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
    FLGPU_HIP(hipEventSynchronize(ctx->ev_stop));
    float t = 0.f;
    FLGPU_HIP(hipEventElapsedTime(&t, ctx->ev_start, ctx->ev_stop));
    *ms = t;
    return 0;
  });
}

extern "C" int flgpu_vec_create(flgpu_ctx *ctx, long n, int prec, flgpu_vec **out) {
  return guard([&] {
    FLGPU_REQUIRE(ctx && out, "ctx and out must not be NULL");
    CTX_LOCK(ctx);
    *out = vec_new(ctx, n, prec);
    return 0;
  });
}

extern "C" int flgpu_vec_destroy(flgpu_vec *v) {
  return guard([&] {
    if (!v) return 0;
    CTX_LOCK(v->ctx);
    FLGPU_HIP(hipStreamSynchronize(v->ctx->stream));
    vec_free(v);
    return 0;
  });
}

extern "C" int flgpu_vec_info(const flgpu_vec *v, long *n, int *prec) {
  return guard([&] {
    FLGPU_REQUIRE(v, "vector must not be NULL");
    if (n) *n = v->n;
    if (prec) *prec = v->prec;
    return 0;
  });
}

extern "C" int flgpu_vec_upload(flgpu_vec *v, const double *host) {
  return guard([&] {
    FLGPU_REQUIRE(v && (host || v->n == 0), "vector and host array must not be NULL");
    CTX_LOCK(v->ctx);
    if (v->n == 0) return 0;
    if (v->prec == FLGPU_F64) {
      FLGPU_HIP(hipMemcpyAsync(v->d, host, sizeof(double) * v->n, hipMemcpyHostToDevice,
                               v->ctx->stream));
      FLGPU_HIP(hipStreamSynchronize(v->ctx->stream));
    } else {
      double *tmp = (double *)dev_alloc(sizeof(double) * v->n);
      FLGPU_HIP(hipMemcpyAsync(tmp, host, sizeof(double) * v->n, hipMemcpyHostToDevice,
                               v->ctx->stream));
      vcopy(v->ctx, v->n, FLGPU_F64, tmp, FLGPU_F32, v->d);
      FLGPU_HIP(hipStreamSynchronize(v->ctx->stream));
      dev_free(tmp);
    }
    return 0;
  });
}

extern "C" int flgpu_vec_download(const flgpu_vec *v, double *host) {
  return guard([&] {
    FLGPU_REQUIRE(v && (host || v->n == 0), "vector and host array must not be NULL");
    CTX_LOCK(v->ctx);
    if (v->n == 0) return 0;
    if (v->prec == FLGPU_F64) {
      FLGPU_HIP(hipMemcpyAsync(host, v->d, sizeof(double) * v->n, hipMemcpyDeviceToHost,
                               v->ctx->stream));
      FLGPU_HIP(hipStreamSynchronize(v->ctx->stream));
    } else {
      double *tmp = (double *)dev_alloc(sizeof(double) * v->n);
      vcopy(v->ctx, v->n, FLGPU_F32, v->d, FLGPU_F64, tmp);
      FLGPU_HIP(hipMemcpyAsync(host, tmp, sizeof(double) * v->n, hipMemcpyDeviceToHost,
                               v->ctx->stream));
      FLGPU_HIP(hipStreamSynchronize(v->ctx->stream));
      dev_free(tmp);
    }
    return 0;
  });
}

extern "C" int flgpu_vec_fill(flgpu_vec *v, double value) {
  return guard([&] {
    FLGPU_REQUIRE(v, "vector must not be NULL");
    CTX_LOCK(v->ctx);
    vfill(v->ctx, v->n, v->prec, v->d, value);
    return 0;
  });
}

extern "C" int flgpu_vec_copy(const flgpu_vec *x, flgpu_vec *y) {
  return guard([&] {
    FLGPU_REQUIRE(x && y, "vectors must not be NULL");
    FLGPU_REQUIRE(x->n == y->n, "copy!: length mismatch");
    FLGPU_REQUIRE(x->ctx == y->ctx, "vectors belong to different contexts");
    CTX_LOCK(x->ctx);
    vcopy(x->ctx, x->n, x->prec, x->d, y->prec, y->d);
    return 0;
  });
}

extern "C" int flgpu_vec_scal(double a, flgpu_vec *x) {
  return guard([&] {
    FLGPU_REQUIRE(x, "vector must not be NULL");
    CTX_LOCK(x->ctx);
    vscal(x->ctx, x->n, x->prec, a, x->d);
    return 0;
  });
}

extern "C" int flgpu_vec_axpy(double a, const flgpu_vec *x, flgpu_vec *y) {
  return guard([&] {
    FLGPU_REQUIRE(x && y, "vectors must not be NULL");
    FLGPU_REQUIRE(x->n == y->n, "axpy!: length mismatch");
    FLGPU_REQUIRE(x->prec == y->prec, "axpy!: precision mismatch");
    FLGPU_REQUIRE(x->ctx == y->ctx, "vectors belong to different contexts");
    CTX_LOCK(x->ctx);
    vaxpy(x->ctx, x->n, x->prec, a, x->d, y->d);
    return 0;
  });
}

extern "C" int flgpu_vec_dot(const flgpu_vec *x, const flgpu_vec *y, double *out) {
  return guard([&] {
    FLGPU_REQUIRE(x && y && out, "arguments must not be NULL");
    FLGPU_REQUIRE(x->n == y->n, "dot: length mismatch");
    FLGPU_REQUIRE(x->prec == y->prec, "dot: precision mismatch");
    FLGPU_REQUIRE(x->ctx == y->ctx, "vectors belong to different contexts");
    CTX_LOCK(x->ctx);
    *out = vdot_host(x->ctx, x->n, x->prec, x->d, y->d);
    return 0;
  });
}

extern "C" int flgpu_vec_dot_noalloc(const flgpu_vec *x, const flgpu_vec *y, double *out) {
  return guard([&] {
    FLGPU_REQUIRE(x && y && out, "arguments must not be NULL");
    FLGPU_REQUIRE(x->n == y->n, "dot: length mismatch");
    FLGPU_REQUIRE(x->prec == y->prec, "dot: precision mismatch");
    FLGPU_REQUIRE(x->ctx == y->ctx, "vectors belong to different contexts");
    CTX_LOCK(x->ctx);
    double *sp = scratch_on(x->ctx->device);
    hipStream_t s = x->ctx->stream;
    // the kernels and grid of flgpu_vec_dot: every block writes its partial and the final
    // pass writes the result, so nothing in the scratch needs clearing between calls
    if (x->prec == FLGPU_F64)
      k_dot_partial<double><<<kRedBlocks, kRedThreads, 0, s>>>(
          x->n, (const double *)x->d, (const double *)y->d, sp);
    else
      k_dot_partial<float><<<kRedBlocks, kRedThreads, 0, s>>>(
          x->n, (const float *)x->d, (const float *)y->d, sp);
    k_dot_final<<<1, kRedBlocks, 0, s>>>(sp, kRedBlocks, sp + kRedBlocks);
    FLGPU_HIP(hipGetLastError());
    FLGPU_HIP(hipMemcpyAsync(out, sp + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, s));
    FLGPU_HIP(hipStreamSynchronize(s));
    return 0;
  });
}

extern "C" int flgpu_vec_norm2(const flgpu_vec *x, double *out) {
  return guard([&] {
    FLGPU_REQUIRE(x && out, "arguments must not be NULL");
    CTX_LOCK(x->ctx);
    double s = vdot_host(x->ctx, x->n, x->prec, x->d, x->d);
    *out = std::sqrt(s);
    return 0;
  });
}
