/*
 * conv_oracle.c -- CPU reference of the learned coarse correction network.
 *
 * TEST INFRASTRUCTURE ONLY (checker of the GPU training and inference
 * kernels).  femlisp has no neural network; this is the plain definition of
 * the network the GPU implements: a 3-level 3D U-Net on an n^3 grid
 * (channels-last tensors [batch][z][y][x][c], zero padding, 3x3x3 kernels
 * W[tap][ci][co] with tap = (dz+1)*9 + (dy+1)*3 + (dx+1)):
 *
 *   L0a conv(cin->16) relu, L0b conv(16->16) relu = s0         (n)
 *   avg-pool 2
 *   L1a conv(16->32) relu, L1b conv(32->32) relu = s1          (n/2)
 *   avg-pool 2
 *   L2a conv(32->64) relu, L2b conv(64->64) relu               (n/4)
 *   nearest upsample, concat [up, s1] (96 ch)
 *   L3  conv(96->32) relu                                       (n/2)
 *   nearest upsample, concat [up, s0] (48 ch)
 *   L4  conv(48->16) relu, L5 conv(16->1)                       (n)
 *
 * Loss: mean over all output values of (y - t)^2.  Double precision
 * throughout; `bf16` = 1 rounds the operands that the GPU feeds to its BF16
 * matrix cores (activations entering a convolution, weights, the incoming
 * gradient of the input-gradient convolution, and both factors of the weight
 * gradient's products) to bfloat16 first.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NL 9
static const int CIN[NL] = {-1, 16, 16, 32, 32, 64, 96, 48, 16};
static const int COUT[NL] = {16, 16, 32, 32, 64, 64, 32, 16, 1};
static const int LEVEL[NL] = {0, 0, 1, 1, 2, 2, 1, 0, 0};

static double bf16r(double x) {
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return f;
}

long flo_unet_nparams(int cin) {
  long p = 0;
  for (int l = 0; l < NL; l++) {
    int ci = l == 0 ? cin : CIN[l];
    p += 27L * ci * COUT[l] + COUT[l];
  }
  return p;
}

/* parameter blocks in order: for each layer W[27][ci][co] then bias[co] */
static void layer_ptrs(const double *theta, int cin, const double **W, const double **B) {
  long off = 0;
  for (int l = 0; l < NL; l++) {
    int ci = l == 0 ? cin : CIN[l];
    W[l] = theta + off;
    off += 27L * ci * COUT[l];
    B[l] = theta + off;
    off += COUT[l];
  }
}

static void conv_fwd(int batch, int n, int ci, int co, const double *in, const double *W,
                     const double *b, double *out, int bf16) {
  long nv = (long)n * n * n;
  for (int bb = 0; bb < batch; bb++)
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
          double *o = out + ((bb * nv) + ((long)z * n + y) * n + x) * co;
          for (int k = 0; k < co; k++) o[k] = b[k];
          for (int t = 0; t < 27; t++) {
            int zz = z + t / 9 - 1, yy = y + (t / 3) % 3 - 1, xx = x + t % 3 - 1;
            if (zz < 0 || zz >= n || yy < 0 || yy >= n || xx < 0 || xx >= n) continue;
            const double *ip = in + ((bb * nv) + ((long)zz * n + yy) * n + xx) * ci;
            for (int c = 0; c < ci; c++) {
              double a = bf16 ? bf16r(ip[c]) : ip[c];
              const double *wr = W + ((long)t * ci + c) * co;
              for (int k = 0; k < co; k++) o[k] += a * (bf16 ? bf16r(wr[k]) : wr[k]);
            }
          }
        }
}

/* gin += conv^T(gout); gW += in (x) gout; gb += sum gout */
static void conv_bwd(int batch, int n, int ci, int co, const double *in, const double *W,
                     const double *gout, double *gin, double *gW, double *gb, int bf16) {
  long nv = (long)n * n * n;
  for (int bb = 0; bb < batch; bb++)
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
          const double *g = gout + ((bb * nv) + ((long)z * n + y) * n + x) * co;
          for (int k = 0; k < co; k++) gb[k] += g[k];
          for (int t = 0; t < 27; t++) {
            int zz = z + t / 9 - 1, yy = y + (t / 3) % 3 - 1, xx = x + t % 3 - 1;
            if (zz < 0 || zz >= n || yy < 0 || yy >= n || xx < 0 || xx >= n) continue;
            long vi = (bb * nv) + ((long)zz * n + yy) * n + xx;
            const double *ip = in + vi * ci;
            double *gi = gin ? gin + vi * ci : NULL;
            for (int c = 0; c < ci; c++) {
              const double *wr = W + ((long)t * ci + c) * co;
              double *gw = gW + ((long)t * ci + c) * co;
              double s = 0.0;
              for (int k = 0; k < co; k++) {
                gw[k] += bf16 ? bf16r(ip[c]) * bf16r(g[k]) : ip[c] * g[k];
                if (gi) s += (bf16 ? bf16r(g[k]) : g[k]) * (bf16 ? bf16r(wr[k]) : wr[k]);
              }
              if (gi) gi[c] += s;
            }
          }
        }
}

static void relu(long m, double *a) {
  for (long i = 0; i < m; i++)
    if (a[i] < 0) a[i] = 0;
}

static void pool(int batch, int n, int c, const double *in, double *out) {
  int h = n / 2;
  memset(out, 0, sizeof(double) * batch * (long)h * h * h * c);
  for (int b = 0; b < batch; b++)
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
          const double *ip = in + (((long)b * n + z) * n * n + (long)y * n + x) * c;
          double *op = out + (((long)b * h + z / 2) * h * h + (long)(y / 2) * h + x / 2) * c;
          for (int k = 0; k < c; k++) op[k] += 0.125 * ip[k];
        }
}

/* up (n/2 -> n, channels cu) and skip (n, channels cs) into cat (n, cu+cs) */
static void upcat(int batch, int n, int cu, int cs, const double *up, const double *skip,
                  double *cat) {
  int h = n / 2;
  for (int b = 0; b < batch; b++)
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
          long v = ((long)b * n + z) * n * n + (long)y * n + x;
          const double *u = up + (((long)b * h + z / 2) * h * h + (long)(y / 2) * h + x / 2) * cu;
          double *o = cat + v * (cu + cs);
          for (int k = 0; k < cu; k++) o[k] = u[k];
          for (int k = 0; k < cs; k++) o[cu + k] = skip[v * cs + k];
        }
}

typedef struct {
  double *a[NL];  /* post-activation outputs of each layer */
  double *p0, *p1, *c1, *c0;
} acts_t;

static long vox(int batch, int n) { return (long)batch * n * n * n; }

static void forward(int batch, int n, int cin, const double *theta, const double *x, acts_t *A,
                    int bf16) {
  const double *W[NL], *B[NL];
  layer_ptrs(theta, cin, W, B);
  int nl[3] = {n, n / 2, n / 4};
  for (int l = 0; l < NL; l++) A->a[l] = (double *)malloc(sizeof(double) * vox(batch, nl[LEVEL[l]]) * COUT[l]);
  A->p0 = (double *)malloc(sizeof(double) * vox(batch, nl[1]) * 16);
  A->p1 = (double *)malloc(sizeof(double) * vox(batch, nl[2]) * 32);
  A->c1 = (double *)malloc(sizeof(double) * vox(batch, nl[1]) * 96);
  A->c0 = (double *)malloc(sizeof(double) * vox(batch, nl[0]) * 48);
  conv_fwd(batch, n, cin, 16, x, W[0], B[0], A->a[0], bf16);
  relu(vox(batch, n) * 16, A->a[0]);
  conv_fwd(batch, n, 16, 16, A->a[0], W[1], B[1], A->a[1], bf16);
  relu(vox(batch, n) * 16, A->a[1]);
  pool(batch, n, 16, A->a[1], A->p0);
  conv_fwd(batch, nl[1], 16, 32, A->p0, W[2], B[2], A->a[2], bf16);
  relu(vox(batch, nl[1]) * 32, A->a[2]);
  conv_fwd(batch, nl[1], 32, 32, A->a[2], W[3], B[3], A->a[3], bf16);
  relu(vox(batch, nl[1]) * 32, A->a[3]);
  pool(batch, nl[1], 32, A->a[3], A->p1);
  conv_fwd(batch, nl[2], 32, 64, A->p1, W[4], B[4], A->a[4], bf16);
  relu(vox(batch, nl[2]) * 64, A->a[4]);
  conv_fwd(batch, nl[2], 64, 64, A->a[4], W[5], B[5], A->a[5], bf16);
  relu(vox(batch, nl[2]) * 64, A->a[5]);
  upcat(batch, nl[1], 64, 32, A->a[5], A->a[3], A->c1);
  conv_fwd(batch, nl[1], 96, 32, A->c1, W[6], B[6], A->a[6], bf16);
  relu(vox(batch, nl[1]) * 32, A->a[6]);
  upcat(batch, n, 32, 16, A->a[6], A->a[1], A->c0);
  conv_fwd(batch, n, 48, 16, A->c0, W[7], B[7], A->a[7], bf16);
  relu(vox(batch, n) * 16, A->a[7]);
  conv_fwd(batch, n, 16, 1, A->a[7], W[8], B[8], A->a[8], bf16);
}

static void free_acts(acts_t *A) {
  for (int l = 0; l < NL; l++) free(A->a[l]);
  free(A->p0);
  free(A->p1);
  free(A->c1);
  free(A->c0);
}

/* forward only: y[batch n^3] */
int flo_unet_forward(int batch, int n, int cin, const double *theta, const double *x,
                     double *y, int bf16) {
  if (n % 4 != 0) return -1;
  acts_t A;
  forward(batch, n, cin, theta, x, &A, bf16);
  memcpy(y, A.a[8], sizeof(double) * vox(batch, n));
  free_acts(&A);
  return 0;
}

static void relu_bwd(long m, const double *act, double *g) {
  for (long i = 0; i < m; i++)
    if (act[i] <= 0) g[i] = 0;
}

/* loss and its gradient with respect to all parameters (grad[nparams]) */
double flo_unet_loss_grad(int batch, int n, int cin, const double *theta, const double *x,
                          const double *t, double *grad, int bf16) {
  acts_t A;
  forward(batch, n, cin, theta, x, &A, bf16);
  const double *W[NL], *B[NL];
  layer_ptrs(theta, cin, W, B);
  double *GW[NL], *GB[NL];
  memset(grad, 0, sizeof(double) * flo_unet_nparams(cin));
  layer_ptrs(grad, cin, (const double **)GW, (const double **)GB);
  int nl[3] = {n, n / 2, n / 4};
  long m = vox(batch, n);
  double loss = 0.0;
  double *g8 = (double *)malloc(sizeof(double) * m);
  for (long i = 0; i < m; i++) {
    double d = A.a[8][i] - t[i];
    loss += d * d;
    g8[i] = 2.0 * d / m;
  }
  loss /= m;
  long v0 = vox(batch, nl[0]), v1 = vox(batch, nl[1]), v2 = vox(batch, nl[2]);
  double *g7 = (double *)calloc(v0 * 16, sizeof(double));
  conv_bwd(batch, n, 16, 1, A.a[7], W[8], g8, g7, GW[8], GB[8], bf16);
  relu_bwd(v0 * 16, A.a[7], g7);
  double *gc0 = (double *)calloc(v0 * 48, sizeof(double));
  conv_bwd(batch, n, 48, 16, A.c0, W[7], g7, gc0, GW[7], GB[7], bf16);
  /* split concat: up part -> a6 (sum over children), skip part -> a1 */
  double *g6 = (double *)calloc(v1 * 32, sizeof(double));
  double *g1 = (double *)calloc(v0 * 16, sizeof(double));
  for (int b = 0; b < batch; b++)
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
          long v = ((long)b * n + z) * n * n + (long)y * n + x;
          int h = n / 2;
          long pv = ((long)b * h + z / 2) * h * h + (long)(y / 2) * h + x / 2;
          for (int k = 0; k < 32; k++) g6[pv * 32 + k] += gc0[v * 48 + k];
          for (int k = 0; k < 16; k++) g1[v * 16 + k] += gc0[v * 48 + 32 + k];
        }
  relu_bwd(v1 * 32, A.a[6], g6);
  double *gc1 = (double *)calloc(v1 * 96, sizeof(double));
  conv_bwd(batch, nl[1], 96, 32, A.c1, W[6], g6, gc1, GW[6], GB[6], bf16);
  double *g5 = (double *)calloc(v2 * 64, sizeof(double));
  double *g3 = (double *)calloc(v1 * 32, sizeof(double));
  {
    int nn = nl[1], h = nl[2];
    for (int b = 0; b < batch; b++)
      for (int z = 0; z < nn; z++)
        for (int y = 0; y < nn; y++)
          for (int x = 0; x < nn; x++) {
            long v = ((long)b * nn + z) * nn * nn + (long)y * nn + x;
            long pv = ((long)b * h + z / 2) * h * h + (long)(y / 2) * h + x / 2;
            for (int k = 0; k < 64; k++) g5[pv * 64 + k] += gc1[v * 96 + k];
            for (int k = 0; k < 32; k++) g3[v * 32 + k] += gc1[v * 96 + 64 + k];
          }
  }
  relu_bwd(v2 * 64, A.a[5], g5);
  double *g4 = (double *)calloc(v2 * 64, sizeof(double));
  conv_bwd(batch, nl[2], 64, 64, A.a[4], W[5], g5, g4, GW[5], GB[5], bf16);
  relu_bwd(v2 * 64, A.a[4], g4);
  double *gp1 = (double *)calloc(v2 * 32, sizeof(double));
  conv_bwd(batch, nl[2], 32, 64, A.p1, W[4], g4, gp1, GW[4], GB[4], bf16);
  /* pool backward into a3 (adds to the skip gradient) */
  {
    int nn = nl[1], h = nl[2];
    for (int b = 0; b < batch; b++)
      for (int z = 0; z < nn; z++)
        for (int y = 0; y < nn; y++)
          for (int x = 0; x < nn; x++) {
            long v = ((long)b * nn + z) * nn * nn + (long)y * nn + x;
            long pv = ((long)b * h + z / 2) * h * h + (long)(y / 2) * h + x / 2;
            for (int k = 0; k < 32; k++) g3[v * 32 + k] += 0.125 * gp1[pv * 32 + k];
          }
  }
  relu_bwd(v1 * 32, A.a[3], g3);
  double *g2 = (double *)calloc(v1 * 32, sizeof(double));
  conv_bwd(batch, nl[1], 32, 32, A.a[2], W[3], g3, g2, GW[3], GB[3], bf16);
  relu_bwd(v1 * 32, A.a[2], g2);
  double *gp0 = (double *)calloc(v1 * 16, sizeof(double));
  conv_bwd(batch, nl[1], 16, 32, A.p0, W[2], g2, gp0, GW[2], GB[2], bf16);
  for (int b = 0; b < batch; b++)
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
          long v = ((long)b * n + z) * n * n + (long)y * n + x;
          int h = n / 2;
          long pv = ((long)b * h + z / 2) * h * h + (long)(y / 2) * h + x / 2;
          for (int k = 0; k < 16; k++) g1[v * 16 + k] += 0.125 * gp0[pv * 16 + k];
        }
  relu_bwd(v0 * 16, A.a[1], g1);
  double *g0 = (double *)calloc(v0 * 16, sizeof(double));
  conv_bwd(batch, n, 16, 16, A.a[0], W[1], g1, g0, GW[1], GB[1], bf16);
  relu_bwd(v0 * 16, A.a[0], g0);
  conv_bwd(batch, n, cin, 16, x, W[0], g0, NULL, GW[0], GB[0], bf16);
  free(g8); free(g7); free(gc0); free(g6); free(g1); free(gc1); free(g5); free(g3);
  free(g4); free(gp1); free(g2); free(gp0); free(g0);
  free_acts(&A);
  return loss;
}

/* one Adam step (Kingma & Ba 2015), bias-corrected, step t >= 1 */
void flo_adam(long np, double *theta, const double *g, double *m, double *v, int t, double lr,
              double b1, double b2, double eps) {
  double c1 = 1.0 - pow(b1, t), c2 = 1.0 - pow(b2, t);
  for (long i = 0; i < np; i++) {
    m[i] = b1 * m[i] + (1.0 - b1) * g[i];
    v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
    theta[i] -= lr * (m[i] / c1) / (sqrt(v[i] / c2) + eps);
  }
}
