// mlp.hpp -- the face-flux surrogate network on gfx950 matrix cores.
//
// MLP 16 -> 128 -> 128 -> 64 -> 8 with SiLU between layers, y = x W + b.
// FP32 master weights; BF16 weights and activations feed
// v_mfma_f32_16x16x32_bf16, which accumulates in FP32.  The parameter
// vector is flat: W1 (16x128, row-major [in][out]), b1, W2, b2, W3, b3, W4,
// b4 -- 27,464 values.
//
// MFMA 16x16x32 BF16 operand layout (lane l = 0..63 of a wave):
//   A (16 x 32, rows x k): lane l holds row l%16, k = 8*(l/16) .. +7
//   B (32 x 16, k x cols): lane l holds col l%16, k = 8*(l/16) .. +7
//   D (16 x 16):           lane l holds col l%16, rows 4*(l/16) .. +3
#pragma once
#include <hip/hip_runtime.h>

namespace bats {
namespace mlp {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int NL = 4;                       // layers
constexpr int DIM[NL + 1] = {16, 128, 128, 64, 8};
// offsets of W_l and b_l in the flat parameter vector
constexpr int OW1 = 0, OB1 = OW1 + 16 * 128;
constexpr int OW2 = OB1 + 128, OB2 = OW2 + 128 * 128;
constexpr int OW3 = OB2 + 128, OB3 = OW3 + 128 * 64;
constexpr int OW4 = OB3 + 64, OB4 = OW4 + 64 * 8;
constexpr int NPARAM = OB4 + 8;             // 27464
constexpr int OW[NL] = {OW1, OW2, OW3, OW4};
constexpr int OB[NL] = {OB1, OB2, OB3, OB4};

// fragment-ordered BF16 weights for the inference kernel.  The network runs
// transposed, Z^T = W^T A^T: a weight tile is the A operand (16 outputs n x
// 32 inputs k) and the activations are the B operand, taken straight from
// the previous layer's accumulators.  Lane l of a layer's output block nb
// holds outputs n = 16 nb + 4 (l/16) + i, i = 0..3, of activation row l%16;
// slot t of its B fragment for input block kb therefore carries input
// k = frag_k(layer, kb, l/16, t): the two output blocks 2 kb and 2 kb + 1
// of the previous layer, four values each (layer 1 reads its inputs as
// k = 8 (l/16) + t).  Layer l is padded to KP[l] x NP[l] and stored tile by
// tile (output block major, then input block), 64 lanes x 8 values each.
constexpr int KP[NL] = {32, 128, 128, 64};
constexpr int NP[NL] = {128, 128, 64, 16};
constexpr int TILE = 64 * 8;                // bf16 values per tile
constexpr int FT0 = 0;
constexpr int FT1 = FT0 + (KP[0] / 32) * (NP[0] / 16);
constexpr int FT2 = FT1 + (KP[1] / 32) * (NP[1] / 16);
constexpr int FT3 = FT2 + (KP[2] / 32) * (NP[2] / 16);
constexpr int NTILE = FT3 + (KP[3] / 32) * (NP[3] / 16);   // 58
constexpr int FT[NL] = {FT0, FT1, FT2, FT3};
constexpr int NFRAG = NTILE * TILE;         // 29696 bf16 = 58 KiB
__host__ __device__ inline int frag_k(int layer, int kb, int g, int t) {
  if (layer == 0) return kb * 32 + 8 * g + t;
  return 32 * kb + (t < 4 ? 4 * g + t : 16 + 4 * g + (t - 4));
}
// the biases packed for the inference kernel's LDS copy
constexpr int CB1 = 0, CB2 = 128, CB3 = 256, CB4 = 320, NBIAS = 328;

// copy the four bias vectors of the flat parameter vector w into sBias
__device__ __forceinline__ void load_biases(const float *w, float *sBias) {
  for (int t = threadIdx.x; t < NBIAS; t += blockDim.x) {
    const int o = t < CB2 ? OB1 + t
                : t < CB3 ? OB2 + (t - CB2)
                : t < CB4 ? OB3 + (t - CB3)
                          : OB4 + (t - CB4);
    sBias[t] = w[o];
  }
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// v_exp_f32 and v_rcp_f32 (about 1 ulp each) instead of the IEEE division
// sequence: the activation is the inference kernel's VALU cost
__device__ __forceinline__ float sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

// =====================================================================
// batched inference: Y[rows x 8] = MLP(X[rows x 16]) with fragment-ordered
// BF16 weights (flat FP32 biases); a workgroup walks 64-row groups
// =====================================================================
// inference workgroups: 16 waves share one LDS copy of the weights (59 KiB);
// two workgroups per CU, 8 waves per SIMD (at most 64 VGPRs)
constexpr int kInferThreads = 1024, kInferWaves = kInferThreads / 64;
constexpr int kInferGrid = 2 * 256;

// one 16-row tile through the network; X rows are FP32, Y rows FP32.  sW:
// fragments in LDS, bias: the packed biases (load_biases) in LDS.  The
// activations never leave the registers: each layer's accumulators, through
// SiLU and BF16, are the next layer's B fragments (see frag_k).
// X rows are FP32 (rounded to BF16 here) or already BF16.
template <typename XT, typename OUT>
__device__ __forceinline__ void infer_tile(const __bf16 *sW, const float *bias,
                                           const XT *X, int rows, int r0,
                                           OUT out) {
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const bf16x8 *W = reinterpret_cast<const bf16x8 *>(sW);
  // layer 1: B fragment straight from X (inputs k >= 16 are padding)
  bf16x8 x;
  const int row0 = r0 + r;
  if (sizeof(XT) == 2) {
    // 16-byte load of inputs 8g .. 8g+7 of the row
    x = (g < 2 && row0 < rows)
            ? *reinterpret_cast<const bf16x8 *>(X + (long)row0 * 16 + 8 * g)
            : bf16x8{};
  } else {
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int k = 8 * g + t;
      x[t] = (k < 16 && row0 < rows) ? (__bf16)X[(long)row0 * 16 + k]
                                     : (__bf16)0.f;
    }
  }
  // accumulators start at the bias of their outputs 16 nb + 4 g + i
  auto bias4 = [&](const float *b, int nb) {
    const float *p = b + nb * 16 + 4 * g;
    return f32x4{p[0], p[1], p[2], p[3]};
  };
  // SiLU of output blocks 2 kb and 2 kb + 1 -> B fragment kb
  auto act = [&](const f32x4 *a, int kb) {
    bf16x8 v;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      v[t] = (__bf16)silu(a[2 * kb][t]);
      v[4 + t] = (__bf16)silu(a[2 * kb + 1][t]);
    }
    return v;
  };
  f32x4 acc[8];
#pragma unroll
  for (int nb = 0; nb < 8; nb++) {
    acc[nb] = mfma(W[(FT0 + nb) * 64 + lane], x, bias4(bias + CB1, nb));
    __builtin_amdgcn_sched_barrier(0);
  }
  bf16x8 b4[4];
#pragma unroll
  for (int kb = 0; kb < 4; kb++) b4[kb] = act(acc, kb);
  // layer 2: 128 -> 128
#pragma unroll
  for (int nb = 0; nb < 8; nb++) {
    acc[nb] = bias4(bias + CB2, nb);
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
      acc[nb] = mfma(W[(FT1 + nb * 4 + kb) * 64 + lane], b4[kb], acc[nb]);
    // keep the weight fragments of one output block live at a time
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int kb = 0; kb < 4; kb++) b4[kb] = act(acc, kb);
  // layer 3: 128 -> 64
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    acc[nb] = bias4(bias + CB3, nb);
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
      acc[nb] = mfma(W[(FT2 + nb * 4 + kb) * 64 + lane], b4[kb], acc[nb]);
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int kb = 0; kb < 2; kb++) b4[kb] = act(acc, kb);
  // layer 4: 64 -> 8 (padded to 16 outputs; lanes with g < 2 hold 0..7)
  f32x4 y = {0.f, 0.f, 0.f, 0.f};
  y = mfma(W[(FT3 + 0) * 64 + lane], b4[0], y);
  y = mfma(W[(FT3 + 1) * 64 + lane], b4[1], y);
  if (g < 2) {
    const int row = r0 + r;
    if (row < rows) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int n = 4 * g + i;
        out(row, n, y[i] + bias[CB4 + n]);
      }
    }
  }
}

}  // namespace mlp
}  // namespace bats
