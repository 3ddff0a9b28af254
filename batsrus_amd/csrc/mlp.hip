// mlp.hip -- training and inference of the face-flux surrogate network on
// gfx950 matrix cores (see mlp.hpp for the network and operand layouts).
//
// Training step on a minibatch of B rows (all kernels hand-written):
//   forward   Z_l = A_{l-1} W_l + b_l, A_l = bf16(silu(Z_l))   (k_gemm FWD)
//   loss      dY = bf16(2 (Y - T) / (8 B)), per-row squared error (k_loss)
//   backward  dZ_{l-1} = bf16((dZ_l W_l^T) * silu'(Z_{l-1}))     (k_gemm BWD)
//             partial dW_l = A_{l-1}^T dZ_l over row chunks        (k_gemm WGRAD)
//             partial db_l = column sums of dZ_l over row chunks   (k_colsum)
//   update    sum the chunk partials in a fixed order, Adam on the FP32
//             master weights, refresh the BF16 copies           (k_adam)
// Every reduction has a fixed order, so a training run is deterministic.
// Inference (k_infer) keeps all 58 KiB of fragment-ordered BF16 weights in
// LDS and pushes 16-row tiles through the four layers per wavefront.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/bats_hip.h"
#include "mlp.hpp"
#include "mlp_host.hpp"

using namespace bats::mlp;

namespace {

enum { EPI_FWD = 0, EPI_OUT = 1, EPI_BWD = 2, EPI_WGRAD = 3 };

struct GemmArgs {
  int M, N, K;
  const __bf16 *A;  // A(m,k) = A[m*sAm + k*sAk]
  long sAm, sAk;
  const __bf16 *B;  // B(k,n) = B[k*sBk + n*sBn]
  long sBk, sBn;
  const float *bias;     // FWD/OUT: [N]
  float *Cf;             // FWD: Z, OUT: Y, WGRAD: partial (ld = ldc)
  long ldc;
  __bf16 *Cb;            // FWD: A = silu(Z), BWD: dZ (ld = ldc)
  const float *Zprev;    // BWD: pre-activation of the layer below (ld = ldc)
  int kChunk;            // WGRAD: rows per split (blockIdx.y)
  long partialStride;    // WGRAD: floats between splits
};

// one 16x16 output tile per wavefront, K in steps of 32
template <int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  const int lane = threadIdx.x & 63;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nTn = (g.N + 15) / 16;
  const int m0 = (tile / nTn) * 16, n0 = (tile % nTn) * 16;
  if (m0 >= g.M) return;  // wave-uniform
  int kBeg = 0, kEnd = g.K;
  if (EPI == EPI_WGRAD) {
    kBeg = blockIdx.y * g.kChunk;
    kEnd = min(g.K, kBeg + g.kChunk);
  }
  const int r = lane & 15, kq = 8 * (lane >> 4);
  const int am = m0 + r, bn = n0 + r;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = kBeg; k0 < kEnd; k0 += 32) {
    bf16x8 a, b;
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int k = k0 + kq + t;
      const bool ok = k < kEnd;
      a[t] = (ok && am < g.M) ? g.A[am * g.sAm + k * g.sAk] : (__bf16)0.0f;
      b[t] = (ok && bn < g.N) ? g.B[k * g.sBk + bn * g.sBn] : (__bf16)0.0f;
    }
    acc = mfma(a, b, acc);
  }
  const int n = n0 + r;
  if (n >= g.N) return;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int m = m0 + 4 * (lane >> 4) + i;
    if (m >= g.M) continue;
    const float v = acc[i];
    if (EPI == EPI_FWD) {
      const float z = v + g.bias[n];
      g.Cf[m * g.ldc + n] = z;
      g.Cb[m * g.ldc + n] = (__bf16)silu(z);
    } else if (EPI == EPI_OUT) {
      g.Cf[m * g.ldc + n] = v + g.bias[n];
    } else if (EPI == EPI_BWD) {
      g.Cb[m * g.ldc + n] = (__bf16)(v * dsilu(g.Zprev[m * g.ldc + n]));
    } else {
      g.Cf[blockIdx.y * g.partialStride + m * g.ldc + n] = v;
    }
  }
}

// dY = bf16(2 (Y - T) / (8 B)); per-row squared error sums
// loss (1/(8B)) sum_r W_r |Y_r - T_r|^2 (W null: all one), its gradient
// dY and the per-row squared errors W_r |Y_r - T_r|^2
__global__ void k_loss(const float *Y, const float *T, const float *W, int B,
                       __bf16 *dY, float *rowErr) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= B) return;
  float s = 0.f;
  const float wr = W ? W[row] : 1.0f;
  const float sc = 2.0f * wr / (8.0f * (float)B);
  for (int j = 0; j < 8; j++) {
    const float d = Y[row * 8 + j] - T[row * 8 + j];
    s += d * d;
    dY[row * 8 + j] = (__bf16)(sc * d);
  }
  rowErr[row] = wr * s;
}

// partial column sums of a BF16 [rows x N] matrix over row chunks
__global__ void k_colsum(const __bf16 *X, int rows, int N, int chunk,
                         float *partial, long partialStride, int off) {
  const int n = threadIdx.x;  // N <= 256
  if (n >= N) return;
  const int r0 = blockIdx.x * chunk, r1 = min(rows, r0 + chunk);
  float s = 0.f;
  for (int r = r0; r < r1; r++) s += (float)X[(long)r * N + n];
  partial[blockIdx.x * partialStride + off + n] = s;
}

// gradient = sum of the row-chunk partials, in chunk order
__global__ void k_sum_partials(const float *partial, int nSplit, float *grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float g = 0.f;
  for (int s = 0; s < nSplit; s++) g += partial[(long)s * NPARAM + p];
  grad[p] = g;
}

struct AdamArgs {
  const float *grad;
  float *w, *m, *v;
  __bf16 *wb;            // BF16 copy, flat layout
  float lr, b1, b2, eps, bc1, bc2;  // bc = 1 - beta^t
  // with a schedule (graph-captured training): lr, bc1, bc2 of step
  // *step are sched[3 * *step + 0..2]
  const float *sched;
  const long *step;
};

// Adam (Kingma & Ba 2015) on the FP32 master weights
__global__ void k_adam(AdamArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float lr = a.lr, bc1 = a.bc1, bc2 = a.bc2;
  if (a.sched) {
    const float *q = a.sched + 3 * *a.step;
    lr = q[0]; bc1 = q[1]; bc2 = q[2];
  }
  const float g = a.grad[p];
  const float m = a.b1 * a.m[p] + (1.f - a.b1) * g;
  const float v = a.b2 * a.v[p] + (1.f - a.b2) * g * g;
  a.m[p] = m;
  a.v[p] = v;
  const float w = a.w[p] - lr * (m / bc1) / (sqrtf(v / bc2) + a.eps);
  a.w[p] = w;
  a.wb[p] = (__bf16)w;
}

// per-block partial sums of squares of the 8 target columns (fixed order)
__global__ void k_col_sumsq(const float *T, long n, double *partial) {
  __shared__ double s[256][8];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long r = blockIdx.x * 256L + threadIdx.x; r < n; r += gridDim.x * 256L)
    for (int j = 0; j < 8; j++) acc[j] += (double)T[r * 8 + j] * T[r * 8 + j];
  for (int j = 0; j < 8; j++) s[threadIdx.x][j] = acc[j];
  __syncthreads();
  if (threadIdx.x < 8) {
    double t = 0.0;
    for (int q = 0; q < 256; q++) t += s[q][threadIdx.x];
    partial[blockIdx.x * 8 + threadIdx.x] = t;
  }
}

// last layer times (mul) or over (div) the per-output scale
__global__ void k_scale_last(float *w, const float *scale, int div) {
  const int t = threadIdx.x;  // 64*8 + 8 entries
  if (t < 64 * 8) {
    const int n = t % 8;
    w[OW4 + t] = div ? w[OW4 + t] / scale[n] : w[OW4 + t] * scale[n];
  } else if (t < 64 * 8 + 8) {
    const int n = t - 64 * 8;
    w[OB4 + n] = div ? w[OB4 + n] / scale[n] : w[OB4 + n] * scale[n];
  }
}

__global__ void k_to_bf16(const float *x, __bf16 *y, long n) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < n) y[t] = (__bf16)x[t];
}

// fragment-ordered, zero-padded BF16 weights for k_infer
__global__ void k_pack_frag(const float *w, __bf16 *frag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NFRAG) return;
  const int tile = t / TILE, lane = (t % TILE) / 8, e = t % 8;
  int l = 0;
  while (l + 1 < NL && tile >= FT[l + 1]) l++;
  const int lt = tile - FT[l], nKb = KP[l] / 32;
  const int nb = lt / nKb, kb = lt % nKb;
  const int k = frag_k(l, kb, lane >> 4, e), n = nb * 16 + (lane & 15);
  const float v = (k < DIM[l] && n < DIM[l + 1])
                      ? w[OW[l] + k * DIM[l + 1] + n] : 0.0f;
  frag[t] = (__bf16)v;
}

}  // namespace

namespace {

__global__ void __launch_bounds__(kInferThreads, 8)
k_infer(const __bf16 *frag, const float *w, const float *X, float *Y,
        int rows) {
  __shared__ __attribute__((aligned(16))) __bf16 sW[NFRAG];
  __shared__ float sBias[NBIAS];
  for (int t = threadIdx.x; t < NFRAG / 8; t += kInferThreads)
    reinterpret_cast<bf16x8 *>(sW)[t] = reinterpret_cast<const bf16x8 *>(frag)[t];
  load_biases(w, sBias);
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  const int nTile = (rows + 15) / 16;
  auto out = [&](int row, int col, float v) { Y[(long)row * 8 + col] = v; };
  for (int tile = blockIdx.x * kInferWaves + wave; tile < nTile;
       tile += gridDim.x * kInferWaves)
    infer_tile(sW, sBias, X, rows, tile * 16, out);
}

// minibatch gather: rows idx[0..B) of the dataset into X (BF16) and the
// standardized targets T / scale
// a minibatch: inputs (BF16), targets divided by scale, and (W non-null)
// the loss weights dataJ / meanJ.  With step non-null the rows are
// idx[(*step % perEpoch) * B + t] (graph-captured training).
__global__ void k_gather_rows(const float *dataX, const float *dataT,
                              const int *idx, int B, __bf16 *X, float *T,
                              const float *scale, const float *dataJ,
                              float invMeanJ, float *W, const long *step,
                              long perEpoch) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B) return;
  const long s = step ? idx[(*step % perEpoch) * B + t] : idx[t];
  for (int j = 0; j < 16; j++) X[t * 16 + j] = (__bf16)dataX[s * 16 + j];
  for (int j = 0; j < 8; j++) T[t * 8 + j] = dataT[s * 8 + j] / scale[j];
  if (W) W[t] = dataJ[s] * invMeanJ;
}
__global__ void k_gather_rows_f32(const float *dataX, const float *dataT,
                                  const float *dataB, const float *dataJ,
                                  const int *idx, int B, float *X, float *T,
                                  float *Bo, float *Jo) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B) return;
  const long s = idx[t];
  for (int j = 0; j < 16; j++) X[(long)t * 16 + j] = dataX[s * 16 + j];
  for (int j = 0; j < 8; j++) T[(long)t * 8 + j] = dataT[s * 8 + j];
  for (int j = 0; j < 8; j++) Bo[(long)t * 8 + j] = dataB[s * 8 + j];
  Jo[t] = dataJ[s];
}

}  // namespace

// =====================================================================
// host side: the model object
// =====================================================================
struct bats_mlp {
  hipStream_t stream;
  int dev;                   // device the model (and its stream) lives on
  float *w, *m, *v, *grad;   // FP32 master weights, Adam moments, gradient
  __bf16 *wb;                // BF16 copy of w (flat layout)
  __bf16 *frag;              // fragment-ordered BF16 weights for inference
  int step;                  // Adam steps taken
  int fragStale;             // frag lags w (repacked before inference)
  // training buffers for a minibatch of cap rows
  int cap;
  __bf16 *X, *A1, *A2, *A3, *dZ1, *dZ2, *dZ3, *dY;
  float *Z1, *Z2, *Z3, *Y, *T, *rowErr, *partial;
  float *rowW;               // loss weights of the minibatch rows or null
  int nSplit;
  // dataset (device): inputs (n x 16), targets (n x 8) and the baseline
  // flux the target corrects (n x 8, zero unless harvested), FP32
  float *dataX, *dataT, *dataB, *dataJ;  // + weight of each sample's target
  long nData, capData;
};

#define MCHK(x)                                                          \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      fprintf(stderr, "bats_mlp: %s failed: %s (%s:%d)\n", #x,           \
              hipGetErrorString(e_), __FILE__, __LINE__);                \
      return e_ == hipErrorOutOfMemory ? BATS_ENOMEM : BATS_EHIP;        \
    }                                                                    \
  } while (0)

static const int kChunkRows = 256;  // rows per weight-gradient split

// ---- staging memory of bats_mlp_forward_noalloc -----------------------
// bats_mlp_forward allocates device buffers for the rows it stages in and
// out on every call; the no-allocation variant stages them through this
// fixed array instead: the inputs (BATS_MLP_NOALLOC_ROWS x 16 floats), then
// the outputs (BATS_MLP_NOALLOC_ROWS x 8), 6 MiB in all.  Each device that
// runs the library holds its own copy until the process ends.  In host code
// the array's name is only a stand-in; its device address on each device
// comes from hipGetSymbolAddress, taken once per device in bats_mlp_create
// with that device current, and only that address is passed to the copies
// and kernels.
static __device__ __align__(256) float
    scratch[(size_t)BATS_MLP_NOALLOC_ROWS * (BATS_MLP_NIN + BATS_MLP_NOUT)];
static const int kMaxDevices = 64;
static std::mutex g_scratchLock;
static float *g_scratchDev[kMaxDevices];  // scratch's address per device

// record the address of this device's copy of scratch (the current device
// must be dev)
static int noalloc_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return BATS_ENODEV;
  std::lock_guard<std::mutex> lock(g_scratchLock);
  if (g_scratchDev[dev]) return BATS_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    fprintf(stderr, "bats_mlp: hipGetSymbolAddress(scratch) failed on "
                    "device %d: %s\n", dev, hipGetErrorString(e));
    return BATS_EHIP;
  }
  g_scratchDev[dev] = static_cast<float *>(p);
  return BATS_OK;
}

static float *noalloc_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratchLock);
  return g_scratchDev[dev];
}

static int repack(bats_mlp *m) {
  if (!m->fragStale) return BATS_OK;
  hipLaunchKernelGGL(k_pack_frag, dim3((NFRAG + 255) / 256), dim3(256), 0,
                     m->stream, m->w, m->frag);
  MCHK(hipGetLastError());
  m->fragStale = 0;
  return BATS_OK;
}

static int refresh_copies(bats_mlp *m) {
  hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0,
                     m->stream, m->w, m->wb, (long)NPARAM);
  hipLaunchKernelGGL(k_pack_frag, dim3((NFRAG + 255) / 256), dim3(256), 0,
                     m->stream, m->w, m->frag);
  MCHK(hipGetLastError());
  return BATS_OK;
}

extern "C" int bats_mlp_create(bats_mlp **out) {
  if (!out) return BATS_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return BATS_ENODEV;
  bats_mlp *m = new bats_mlp();
  memset(m, 0, sizeof *m);
  MCHK(hipGetDevice(&m->dev));
  const int st = noalloc_setup(m->dev);
  if (st) {
    delete m;
    return st;
  }
  MCHK(hipStreamCreate(&m->stream));
  MCHK(hipMalloc(&m->w, sizeof(float) * NPARAM));
  MCHK(hipMalloc(&m->m, sizeof(float) * NPARAM));
  MCHK(hipMalloc(&m->v, sizeof(float) * NPARAM));
  MCHK(hipMalloc(&m->grad, sizeof(float) * NPARAM));
  MCHK(hipMalloc(&m->wb, sizeof(__bf16) * NPARAM));
  MCHK(hipMalloc(&m->frag, sizeof(__bf16) * NFRAG));
  MCHK(hipMemset(m->w, 0, sizeof(float) * NPARAM));
  MCHK(hipMemset(m->m, 0, sizeof(float) * NPARAM));
  MCHK(hipMemset(m->v, 0, sizeof(float) * NPARAM));
  MCHK(hipMemset(m->grad, 0, sizeof(float) * NPARAM));
  *out = m;
  return refresh_copies(m);
}

static void free_batch(bats_mlp *m) {
  void *p[] = {m->X, m->A1, m->A2, m->A3, m->dZ1, m->dZ2, m->dZ3, m->dY,
               m->Z1, m->Z2, m->Z3, m->Y, m->T, m->rowErr, m->partial,
               m->rowW};
  for (void *q : p)
    if (q) (void)hipFree(q);
  m->X = m->A1 = m->A2 = m->A3 = m->dZ1 = m->dZ2 = m->dZ3 = m->dY = nullptr;
  m->Z1 = m->Z2 = m->Z3 = m->Y = m->T = m->rowErr = m->partial = nullptr;
  m->rowW = nullptr;
  m->cap = 0;
}

extern "C" int bats_mlp_free(bats_mlp *m) {
  if (!m) return BATS_EINVAL;
  (void)hipStreamSynchronize(m->stream);
  free_batch(m);
  void *p[] = {m->w, m->m, m->v, m->grad, m->wb, m->frag, m->dataX, m->dataT,
               m->dataB, m->dataJ};
  for (void *q : p)
    if (q) (void)hipFree(q);
  (void)hipStreamDestroy(m->stream);
  delete m;
  return BATS_OK;
}

extern "C" int bats_mlp_set_params(bats_mlp *m, const float *w) {
  if (!m || !w) return BATS_EINVAL;
  MCHK(hipMemcpyAsync(m->w, w, sizeof(float) * NPARAM, hipMemcpyHostToDevice,
                      m->stream));
  MCHK(hipMemsetAsync(m->m, 0, sizeof(float) * NPARAM, m->stream));
  MCHK(hipMemsetAsync(m->v, 0, sizeof(float) * NPARAM, m->stream));
  m->step = 0;
  int st = refresh_copies(m);
  if (st) return st;
  MCHK(hipStreamSynchronize(m->stream));
  return BATS_OK;
}

extern "C" int bats_mlp_get_params(bats_mlp *m, float *w, float *grad) {
  if (!m) return BATS_EINVAL;
  MCHK(hipStreamSynchronize(m->stream));
  if (w) MCHK(hipMemcpy(w, m->w, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  if (grad)
    MCHK(hipMemcpy(grad, m->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  return BATS_OK;
}

static int ensure_batch(bats_mlp *m, int B) {
  if (B <= m->cap) return BATS_OK;
  free_batch(m);
  const size_t b = (size_t)B;
  MCHK(hipMalloc(&m->X, sizeof(__bf16) * b * 16));
  MCHK(hipMalloc(&m->A1, sizeof(__bf16) * b * 128));
  MCHK(hipMalloc(&m->A2, sizeof(__bf16) * b * 128));
  MCHK(hipMalloc(&m->A3, sizeof(__bf16) * b * 64));
  MCHK(hipMalloc(&m->dZ1, sizeof(__bf16) * b * 128));
  MCHK(hipMalloc(&m->dZ2, sizeof(__bf16) * b * 128));
  MCHK(hipMalloc(&m->dZ3, sizeof(__bf16) * b * 64));
  MCHK(hipMalloc(&m->dY, sizeof(__bf16) * b * 8));
  MCHK(hipMalloc(&m->Z1, sizeof(float) * b * 128));
  MCHK(hipMalloc(&m->Z2, sizeof(float) * b * 128));
  MCHK(hipMalloc(&m->Z3, sizeof(float) * b * 64));
  MCHK(hipMalloc(&m->Y, sizeof(float) * b * 8));
  MCHK(hipMalloc(&m->T, sizeof(float) * b * 8));
  MCHK(hipMalloc(&m->rowErr, sizeof(float) * b));
  MCHK(hipMalloc(&m->rowW, sizeof(float) * b));
  m->nSplit = (B + kChunkRows - 1) / kChunkRows;
  MCHK(hipMalloc(&m->partial, sizeof(float) * (size_t)m->nSplit * NPARAM));
  m->cap = B;
  return BATS_OK;
}

static GemmArgs gemm0() {
  GemmArgs g;
  memset(&g, 0, sizeof g);
  return g;
}

static void launch_gemm(int epi, const GemmArgs &g, hipStream_t s, int splits) {
  const int tiles = ((g.M + 15) / 16) * ((g.N + 15) / 16);
  const dim3 grid((tiles + 3) / 4, splits), blk(256);
  switch (epi) {
  case EPI_FWD: hipLaunchKernelGGL(k_gemm<EPI_FWD>, grid, blk, 0, s, g); break;
  case EPI_OUT: hipLaunchKernelGGL(k_gemm<EPI_OUT>, grid, blk, 0, s, g); break;
  case EPI_BWD: hipLaunchKernelGGL(k_gemm<EPI_BWD>, grid, blk, 0, s, g); break;
  default: hipLaunchKernelGGL(k_gemm<EPI_WGRAD>, grid, blk, 0, s, g); break;
  }
}

// forward of the minibatch in m->X (B rows) into Z1..Z3, A1..A3, Y
static void forward_batch(bats_mlp *m, int B) {
  const __bf16 *in[4] = {m->X, m->A1, m->A2, m->A3};
  float *Z[4] = {m->Z1, m->Z2, m->Z3, m->Y};
  __bf16 *Aout[3] = {m->A1, m->A2, m->A3};
  for (int l = 0; l < NL; l++) {
    GemmArgs g = gemm0();
    g.M = B; g.N = DIM[l + 1]; g.K = DIM[l];
    g.A = in[l]; g.sAm = DIM[l]; g.sAk = 1;
    g.B = m->wb + OW[l]; g.sBk = DIM[l + 1]; g.sBn = 1;
    g.bias = m->w + OB[l];
    g.Cf = Z[l]; g.ldc = DIM[l + 1];
    if (l < NL - 1) {
      g.Cb = Aout[l];
      launch_gemm(EPI_FWD, g, m->stream, 1);
    } else {
      launch_gemm(EPI_OUT, g, m->stream, 1);
    }
  }
}

// gradient of the mean-square loss of the minibatch (X, T already on the
// device) into m->grad and, if lr > 0, one Adam step; rowErr filled
// the step counter of graph-captured training, advanced after each step
__global__ void k_step_inc(long *step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

static int grad_step(bats_mlp *m, int B, float lr, bool update,
                     const float *W = nullptr, const float *sched = nullptr,
                     const long *dStep = nullptr) {
  forward_batch(m, B);
  hipLaunchKernelGGL(k_loss, dim3((B + 255) / 256), dim3(256), 0, m->stream,
                     m->Y, m->T, W, B, m->dY, m->rowErr);
  const __bf16 *Ain[4] = {m->X, m->A1, m->A2, m->A3};
  __bf16 *dZ[4] = {m->dZ1, m->dZ2, m->dZ3, m->dY};
  const float *Zs[3] = {m->Z1, m->Z2, m->Z3};
  const int S = (B + kChunkRows - 1) / kChunkRows;
  for (int l = NL - 1; l >= 0; l--) {
    // weight gradient partials: dW_l = A_{l-1}^T dZ_l over row chunks
    GemmArgs g = gemm0();
    g.M = DIM[l]; g.N = DIM[l + 1]; g.K = B;
    g.A = Ain[l]; g.sAm = 1; g.sAk = DIM[l];
    g.B = dZ[l]; g.sBk = DIM[l + 1]; g.sBn = 1;
    g.Cf = m->partial + OW[l]; g.ldc = DIM[l + 1];
    g.kChunk = kChunkRows; g.partialStride = NPARAM;
    launch_gemm(EPI_WGRAD, g, m->stream, S);
    hipLaunchKernelGGL(k_colsum, dim3(S), dim3(256), 0, m->stream, dZ[l], B,
                       DIM[l + 1], kChunkRows, m->partial, (long)NPARAM, OB[l]);
    if (l > 0) {
      // dZ_{l-1} = (dZ_l W_l^T) * silu'(Z_{l-1})
      GemmArgs d = gemm0();
      d.M = B; d.N = DIM[l]; d.K = DIM[l + 1];
      d.A = dZ[l]; d.sAm = DIM[l + 1]; d.sAk = 1;
      d.B = m->wb + OW[l]; d.sBk = 1; d.sBn = DIM[l + 1];
      d.Cb = dZ[l - 1]; d.ldc = DIM[l]; d.Zprev = Zs[l - 1];
      launch_gemm(EPI_BWD, d, m->stream, 1);
    }
  }
  hipLaunchKernelGGL(k_sum_partials, dim3((NPARAM + 255) / 256), dim3(256), 0,
                     m->stream, m->partial, S, m->grad);
  if (update) {
    AdamArgs a;
    a.grad = m->grad; a.w = m->w; a.m = m->m; a.v = m->v; a.wb = m->wb;
    a.b1 = 0.9f; a.b2 = 0.999f; a.eps = 1e-8f;
    a.sched = sched;
    a.step = dStep;
    if (!sched) {
      m->step++;
      a.lr = lr;
      a.bc1 = 1.f - powf(a.b1, (float)m->step);
      a.bc2 = 1.f - powf(a.b2, (float)m->step);
    } else {
      a.lr = a.bc1 = a.bc2 = 0.f;
    }
    hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0,
                       m->stream, a);
    m->fragStale = 1;
  }
  MCHK(hipGetLastError());
  return BATS_OK;
}


extern "C" int bats_mlp_loss_grad(bats_mlp *m, const float *X, const float *T,
                                  int B, double *Loss, float *grad) {
  if (!m || !X || !T || B <= 0) return BATS_EINVAL;
  int st = ensure_batch(m, B);
  if (st) return st;
  std::vector<__bf16> xb((size_t)B * 16);
  for (size_t t = 0; t < xb.size(); t++) xb[t] = (__bf16)X[t];
  MCHK(hipMemcpyAsync(m->X, xb.data(), sizeof(__bf16) * xb.size(),
                      hipMemcpyHostToDevice, m->stream));
  MCHK(hipMemcpyAsync(m->T, T, sizeof(float) * (size_t)B * 8,
                      hipMemcpyHostToDevice, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  st = grad_step(m, B, 0.f, false);
  if (st) return st;
  std::vector<float> e(B);
  MCHK(hipMemcpyAsync(e.data(), m->rowErr, sizeof(float) * B,
                      hipMemcpyDeviceToHost, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  double s = 0.0;
  for (int r = 0; r < B; r++) s += e[r];
  if (Loss) *Loss = s / (8.0 * B);
  if (grad)
    MCHK(hipMemcpy(grad, m->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  return BATS_OK;
}

extern "C" int bats_mlp_forward(bats_mlp *m, const float *X, float *Y, int rows) {
  if (!m || !X || !Y || rows <= 0) return BATS_EINVAL;
  float *dX = nullptr, *dY = nullptr;
  MCHK(hipMalloc(&dX, sizeof(float) * (size_t)rows * 16));
  MCHK(hipMalloc(&dY, sizeof(float) * (size_t)rows * 8));
  MCHK(hipMemcpyAsync(dX, X, sizeof(float) * (size_t)rows * 16,
                      hipMemcpyHostToDevice, m->stream));
  int st = bats::mlp::forward_device(m, dX, dY, rows, m->stream);
  if (st == BATS_OK)
    MCHK(hipMemcpyAsync(Y, dY, sizeof(float) * (size_t)rows * 8,
                        hipMemcpyDeviceToHost, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  (void)hipFree(dX);
  (void)hipFree(dY);
  return st;
}

extern "C" int bats_mlp_forward_noalloc(bats_mlp *m, const float *X,
                                        float *Y, int rows) {
  if (!m || !X || !Y || rows <= 0) return BATS_EINVAL;
  if (rows > BATS_MLP_NOALLOC_ROWS) {
    fprintf(stderr, "bats_mlp_forward_noalloc: %d rows exceed the staging "
                    "limit of %d rows\n", rows, (int)BATS_MLP_NOALLOC_ROWS);
    return BATS_ELIMIT;
  }
  float *dX = noalloc_scratch(m->dev);
  if (!dX) return BATS_EHIP;
  float *dY = dX + (size_t)BATS_MLP_NOALLOC_ROWS * BATS_MLP_NIN;
  // every row written by the copy in is read, every row read by the copy
  // out is written by the kernel: nothing in scratch needs clearing
  MCHK(hipMemcpyAsync(dX, X, sizeof(float) * (size_t)rows * BATS_MLP_NIN,
                      hipMemcpyHostToDevice, m->stream));
  const int st = bats::mlp::forward_device(m, dX, dY, rows, m->stream);
  if (st == BATS_OK)
    MCHK(hipMemcpyAsync(Y, dY, sizeof(float) * (size_t)rows * BATS_MLP_NOUT,
                        hipMemcpyDeviceToHost, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  return st;
}

int bats::mlp::forward_device(bats_mlp *m, const float *dX, float *dY,
                              int rows, hipStream_t stream) {
  if (!m || rows <= 0) return BATS_EINVAL;
  // the weights must be current on this stream
  int st = repack(m);
  if (st) return st;
  MCHK(hipStreamSynchronize(m->stream));
  const int nTile = (rows + 15) / 16;
  int grid = (nTile + kInferWaves - 1) / kInferWaves;
  if (grid > kInferGrid) grid = kInferGrid;
  hipLaunchKernelGGL(k_infer, dim3(grid), dim3(kInferThreads), 0, stream, m->frag, m->w,
                     dX, dY, rows);
  MCHK(hipGetLastError());
  return BATS_OK;
}

// ---- dataset and training loop ----------------------------------------
static int ensure_data(bats_mlp *m, long n) {
  if (n <= m->capData) return BATS_OK;
  float *x = nullptr, *t = nullptr, *b = nullptr, *jw = nullptr;
  MCHK(hipMalloc(&x, sizeof(float) * (size_t)n * 16));
  MCHK(hipMalloc(&t, sizeof(float) * (size_t)n * 8));
  MCHK(hipMalloc(&b, sizeof(float) * (size_t)n * 8));
  MCHK(hipMalloc(&jw, sizeof(float) * (size_t)n));
  if (m->nData > 0) {
    MCHK(hipMemcpyAsync(x, m->dataX, sizeof(float) * (size_t)m->nData * 16,
                        hipMemcpyDeviceToDevice, m->stream));
    MCHK(hipMemcpyAsync(t, m->dataT, sizeof(float) * (size_t)m->nData * 8,
                        hipMemcpyDeviceToDevice, m->stream));
    MCHK(hipMemcpyAsync(b, m->dataB, sizeof(float) * (size_t)m->nData * 8,
                        hipMemcpyDeviceToDevice, m->stream));
    MCHK(hipMemcpyAsync(jw, m->dataJ, sizeof(float) * (size_t)m->nData,
                        hipMemcpyDeviceToDevice, m->stream));
  }
  MCHK(hipStreamSynchronize(m->stream));
  if (m->dataX) (void)hipFree(m->dataX);
  if (m->dataT) (void)hipFree(m->dataT);
  if (m->dataB) (void)hipFree(m->dataB);
  if (m->dataJ) (void)hipFree(m->dataJ);
  m->dataX = x;
  m->dataT = t;
  m->dataB = b;
  m->dataJ = jw;
  m->capData = n;
  return BATS_OK;
}

int bats::mlp::reserve_data(bats_mlp *m, long nMore, float **X, float **T,
                            float **B, float **J) {
  int st = ensure_data(m, m->nData + nMore);
  if (st) return st;
  *X = m->dataX + (size_t)m->nData * 16;
  *T = m->dataT + (size_t)m->nData * 8;
  *B = m->dataB + (size_t)m->nData * 8;
  *J = m->dataJ + (size_t)m->nData;
  return BATS_OK;
}

__global__ void k_fill(float *x, long n, float v) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < n) x[t] = v;
}

void bats::mlp::commit_data(bats_mlp *m, long nAdded) { m->nData += nAdded; }

bats::mlp::Weights bats::mlp::weights(bats_mlp *m) {
  (void)repack(m);
  (void)hipStreamSynchronize(m->stream);  // weights current
  return Weights{m->frag, m->w};
}

extern "C" int bats_mlp_set_data(bats_mlp *m, const float *X, const float *T,
                                 long n) {
  if (!m || n < 0 || (n > 0 && (!X || !T))) return BATS_EINVAL;
  m->nData = 0;
  int st = ensure_data(m, n);
  if (st) return st;
  if (n > 0) {
    MCHK(hipMemcpy(m->dataX, X, sizeof(float) * (size_t)n * 16,
                   hipMemcpyHostToDevice));
    MCHK(hipMemcpy(m->dataT, T, sizeof(float) * (size_t)n * 8,
                   hipMemcpyHostToDevice));
    MCHK(hipMemset(m->dataB, 0, sizeof(float) * (size_t)n * 8));
    hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, m->stream,
                       m->dataJ, n, 1.0f);
    MCHK(hipStreamSynchronize(m->stream));
  }
  m->nData = n;
  return BATS_OK;
}

extern "C" int bats_mlp_get_data(bats_mlp *m, long *n, float *X, float *T,
                                 float *B) {
  if (!m) return BATS_EINVAL;
  MCHK(hipStreamSynchronize(m->stream));
  if (n) *n = m->nData;
  if (X && m->nData)
    MCHK(hipMemcpy(X, m->dataX, sizeof(float) * (size_t)m->nData * 16,
                   hipMemcpyDeviceToHost));
  if (T && m->nData)
    MCHK(hipMemcpy(T, m->dataT, sizeof(float) * (size_t)m->nData * 8,
                   hipMemcpyDeviceToHost));
  if (B && m->nData)
    MCHK(hipMemcpy(B, m->dataB, sizeof(float) * (size_t)m->nData * 8,
                   hipMemcpyDeviceToHost));
  return BATS_OK;
}

// splitmix64: the seeded shuffles of the training loop
static uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static void shuffle(std::vector<int> &v, uint64_t &s) {
  for (size_t i = v.size(); i > 1; i--) {
    const size_t j = splitmix64(s) % i;
    std::swap(v[i - 1], v[j]);
  }
}

extern "C" int bats_mlp_train(bats_mlp *m, int nEpoch, int Batch, double Lr0,
                              double HoldFrac, int Flags,
                              unsigned long long Seed, double *TrainLoss,
                              double *HeldOutRelErr_V) {
  const bool Standardize = (Flags & BATS_MLP_STANDARDIZE) != 0;
  const bool JumpWeight = (Flags & BATS_MLP_JUMP_WEIGHT) != 0;
  const bool UseGraph = (Flags & BATS_MLP_NO_GRAPH) == 0;
  if (Flags & ~(BATS_MLP_STANDARDIZE | BATS_MLP_JUMP_WEIGHT | BATS_MLP_NO_GRAPH))
    return BATS_EINVAL;
  if (!m || nEpoch < 0 || Batch <= 0 || !(Lr0 >= 0.0) || HoldFrac < 0.0 ||
      HoldFrac >= 1.0 || m->nData < 2)
    return BATS_EINVAL;
  const long n = m->nData;
  if (n > 0x7fffffffL) return BATS_EINVAL;
  const long nHold = (long)(HoldFrac * (double)n);
  const long nTrain = n - nHold;
  if (nTrain < 1) return BATS_EINVAL;
  const int B = (int)(Batch < nTrain ? Batch : nTrain);
  int st = ensure_batch(m, B);
  if (st) return st;
  // seeded split: the first nTrain of a permutation train, the rest are held
  uint64_t s = Seed;
  std::vector<int> perm(n);
  for (long i = 0; i < n; i++) perm[i] = (int)i;
  shuffle(perm, s);
  std::vector<int> train(perm.begin(), perm.begin() + nTrain);
  std::vector<int> hold(perm.begin() + nTrain, perm.end());
  int *dIdx = nullptr;
  MCHK(hipMalloc(&dIdx, sizeof(int) * (size_t)(nTrain > nHold ? nTrain : nHold)));
  // optionally standardize the targets: each output divided by its RMS
  // over the samples while training; the last layer carries the scale in
  // between calls, so the network always outputs physical targets
  float *dScale = nullptr;
  MCHK(hipMalloc(&dScale, sizeof(float) * 8));
  {
    const float one[8] = {1, 1, 1, 1, 1, 1, 1, 1};
    MCHK(hipMemcpy(dScale, one, sizeof one, hipMemcpyHostToDevice));
  }
  if (Standardize) {
    const int nb = 256;
    double *dPart = nullptr;
    MCHK(hipMalloc(&dPart, sizeof(double) * nb * 8));
    hipLaunchKernelGGL(k_col_sumsq, dim3(nb), dim3(256), 0, m->stream,
                       m->dataT, n, dPart);
    std::vector<double> part(nb * 8);
    MCHK(hipMemcpyAsync(part.data(), dPart, sizeof(double) * part.size(),
                        hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    (void)hipFree(dPart);
    float sc[8];
    for (int j = 0; j < 8; j++) {
      double t = 0.0;
      for (int q = 0; q < nb; q++) t += part[q * 8 + j];
      const double rms = sqrt(t / (double)n);
      sc[j] = rms > 0.0 ? (float)rms : 1.0f;
    }
    MCHK(hipMemcpyAsync(dScale, sc, sizeof sc, hipMemcpyHostToDevice,
                        m->stream));
    hipLaunchKernelGGL(k_scale_last, dim3(1), dim3(64 * 8 + 8), 0, m->stream,
                       m->w, dScale, 1);
    hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0,
                       m->stream, m->w, m->wb, (long)NPARAM);
  }
  // jump weights: each sample's squared error weighted by its jump norm
  // over the mean jump norm of the training samples
  float invMeanJ = 1.0f;
  if (JumpWeight) {
    std::vector<float> hj(n);
    MCHK(hipStreamSynchronize(m->stream));
    MCHK(hipMemcpy(hj.data(), m->dataJ, sizeof(float) * n,
                   hipMemcpyDeviceToHost));
    double sj = 0.0;
    for (long i = 0; i < nTrain; i++) sj += hj[train[i]];
    invMeanJ = sj > 0.0 ? (float)((double)nTrain / sj) : 1.0f;
  }
  const long stepsPerEpoch = nTrain / B;
  const long total = stepsPerEpoch * nEpoch;
  double lastLoss = 0.0;
  std::vector<float> e(B);
  if (total > 0) {
    // the learning rate (cosine decay) and Adam's bias corrections of every
    // step, read on the device: one training step is captured once in a
    // hipGraph and replayed (the minibatch of step q is rows q % perEpoch of
    // the epoch's permutation; a device counter advances after each step)
    std::vector<float> sched(3 * (size_t)total);
    for (long q = 0; q < total; q++) {
      const double lr = Lr0 * 0.5 * (1.0 + cos(M_PI * (double)q / (double)total));
      const float t = (float)(m->step + q + 1);
      sched[3 * q] = (float)lr;
      sched[3 * q + 1] = 1.f - powf(0.9f, t);
      sched[3 * q + 2] = 1.f - powf(0.999f, t);
    }
    float *dSched = nullptr;
    long *dStep = nullptr;
    MCHK(hipMalloc(&dSched, sizeof(float) * sched.size()));
    MCHK(hipMalloc(&dStep, sizeof(long)));
    MCHK(hipMemcpyAsync(dSched, sched.data(), sizeof(float) * sched.size(),
                        hipMemcpyHostToDevice, m->stream));
    MCHK(hipMemsetAsync(dStep, 0, sizeof(long), m->stream));
    float *W = JumpWeight ? m->rowW : nullptr;
    // one training step: the same kernels whether captured or launched
    auto one_step = [&]() {
      hipLaunchKernelGGL(k_gather_rows, dim3((B + 255) / 256), dim3(256), 0,
                         m->stream, m->dataX, m->dataT, dIdx, B, m->X, m->T,
                         dScale, m->dataJ, invMeanJ, W, dStep, stepsPerEpoch);
      const int r = grad_step(m, B, 0.f, true, W, dSched, dStep);
      hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, m->stream, dStep);
      return r;
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (UseGraph) {
      MCHK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
      st = one_step();
      const hipError_t ec = hipStreamEndCapture(m->stream, &graph);
      if (st == BATS_OK && ec == hipSuccess)
        st = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess
                 ? BATS_OK : BATS_EHIP;
      else if (st == BATS_OK)
        st = BATS_EHIP;
      if (graph) (void)hipGraphDestroy(graph);
    }
    for (int ep = 0; st == BATS_OK && ep < nEpoch; ep++) {
      shuffle(train, s);
      if (hipMemcpyAsync(dIdx, train.data(), sizeof(int) * train.size(),
                         hipMemcpyHostToDevice, m->stream) != hipSuccess) {
        st = BATS_EHIP;
        break;
      }
      for (long b = 0; st == BATS_OK && b < stepsPerEpoch; b++) {
        if (UseGraph)
          st = hipGraphLaunch(exec, m->stream) == hipSuccess ? BATS_OK : BATS_EHIP;
        else
          st = one_step();
      }
      if (st != BATS_OK) break;
      // loss of the epoch's last minibatch
      if (hipMemcpyAsync(e.data(), m->rowErr, sizeof(float) * B,
                         hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
          hipStreamSynchronize(m->stream) != hipSuccess) {
        st = BATS_EHIP;
        break;
      }
      double epLoss = 0.0;
      for (int r = 0; r < B; r++) epLoss += e[r];
      lastLoss = epLoss / (8.0 * B);
    }
    (void)hipStreamSynchronize(m->stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    (void)hipFree(dSched);
    (void)hipFree(dStep);
    if (st) {
      (void)hipFree(dIdx);
      (void)hipFree(dScale);
      return st;
    }
    m->step += total;
  }
  // back to physical targets (a no-op scale of one without Standardize)
  hipLaunchKernelGGL(k_scale_last, dim3(1), dim3(64 * 8 + 8), 0, m->stream,
                     m->w, dScale, 0);
  hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0,
                     m->stream, m->w, m->wb, (long)NPARAM);
  m->fragStale = 1;
  MCHK(hipStreamSynchronize(m->stream));
  (void)hipFree(dScale);
  if (TrainLoss) *TrainLoss = lastLoss;
  // held-out relative error: ||Y - T|| / ||T|| overall and per output
  if (HeldOutRelErr_V && nHold > 0) {
    float *hx = nullptr, *ht = nullptr, *hy = nullptr, *hb = nullptr,
          *hj = nullptr;
    MCHK(hipMalloc(&hx, sizeof(float) * (size_t)nHold * 16));
    MCHK(hipMalloc(&ht, sizeof(float) * (size_t)nHold * 8));
    MCHK(hipMalloc(&hy, sizeof(float) * (size_t)nHold * 8));
    MCHK(hipMalloc(&hb, sizeof(float) * (size_t)nHold * 8));
    MCHK(hipMalloc(&hj, sizeof(float) * (size_t)nHold));
    MCHK(hipMemcpyAsync(dIdx, hold.data(), sizeof(int) * hold.size(),
                        hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_gather_rows_f32, dim3((nHold + 255) / 256), dim3(256),
                       0, m->stream, m->dataX, m->dataT, m->dataB, m->dataJ,
                       dIdx, (int)nHold, hx, ht, hb, hj);
    MCHK(hipStreamSynchronize(m->stream));
    st = bats::mlp::forward_device(m, hx, hy, (int)nHold, m->stream);
    std::vector<float> Y((size_t)nHold * 8), T((size_t)nHold * 8),
        Bv((size_t)nHold * 8), Jv((size_t)nHold);
    MCHK(hipMemcpyAsync(Y.data(), hy, sizeof(float) * Y.size(),
                        hipMemcpyDeviceToHost, m->stream));
    MCHK(hipMemcpyAsync(T.data(), ht, sizeof(float) * T.size(),
                        hipMemcpyDeviceToHost, m->stream));
    MCHK(hipMemcpyAsync(Bv.data(), hb, sizeof(float) * Bv.size(),
                        hipMemcpyDeviceToHost, m->stream));
    MCHK(hipMemcpyAsync(Jv.data(), hj, sizeof(float) * Jv.size(),
                        hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    double num[9] = {0}, den[9] = {0}, denT[9] = {0};
    for (long r = 0; r < nHold; r++)
      for (int j = 0; j < 8; j++) {
        // predicted flux base + w Y against the full flux base + w T,
        // w the sample's target weight (the jump norm for harvested faces)
        const double w = Jv[r];
        const double d = w * ((double)Y[r * 8 + j] - (double)T[r * 8 + j]);
        const double t = w * (double)T[r * 8 + j] + (double)Bv[r * 8 + j];
        const double tt = w * (double)T[r * 8 + j];
        num[j + 1] += d * d;
        den[j + 1] += t * t;
        denT[j + 1] += tt * tt;
        num[0] += d * d;
        den[0] += t * t;
        denT[0] += tt * tt;
      }
    for (int j = 0; j < 9; j++) {
      HeldOutRelErr_V[j] = den[j] > 0 ? sqrt(num[j] / den[j]) : sqrt(num[j]);
      HeldOutRelErr_V[9 + j] =
          denT[j] > 0 ? sqrt(num[j] / denT[j]) : sqrt(num[j]);
    }
    (void)hipFree(hx);
    (void)hipFree(ht);
    (void)hipFree(hy);
    (void)hipFree(hb);
    (void)hipFree(hj);
  }
  (void)hipFree(dIdx);
  return st;
}
