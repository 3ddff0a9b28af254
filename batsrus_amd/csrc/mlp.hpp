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

// fragment-ordered BF16 weights for the inference kernel: layer l is padded
// to KP[l] x NP[l] (multiples of 32 x 16) and stored tile by tile
// (k-block major, then n-block), 64 lanes x 8 values per 32x16 tile
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
constexpr int ACT_LD = 128 + 8;  // bf16 row stride of the activation tile
// inference workgroups: 16 waves share one LDS copy of the weights (58 KiB
// + 16 activation tiles of 4.3 KiB), 4 waves per SIMD
constexpr int kInferThreads = 1024, kInferWaves = kInferThreads / 64;

// one 16-row tile through the network; X rows are FP32, Y rows FP32.  sW:
// fragments in LDS, bias: the packed biases (load_biases) in LDS, sAct:
// this wave's 16 x ACT_LD activation tile
template <typename OUT>
__device__ __forceinline__ void infer_tile(const __bf16 *sW, const float *bias,
                                           __bf16 *sAct, const float *X,
                                           int rows, int r0, OUT out) {
  const int lane = threadIdx.x & 63, r = lane & 15, kq = 8 * (lane >> 4);
  const bf16x8 *W = reinterpret_cast<const bf16x8 *>(sW);
  // layer 1: A fragment straight from X (k >= 16 is padding)
  bf16x8 x;
#pragma unroll
  for (int t = 0; t < 8; t++) {
    const int k = kq + t, row = r0 + r;
    x[t] = (k < 16 && row < rows) ? (__bf16)X[(long)row * 16 + k] : (__bf16)0.f;
  }
  // accumulators start at the bias of their column (lane l holds column
  // l%16 of each 16-column block)
  auto bias4 = [&](const float *b, int nb) {
    const float v = b[nb * 16 + r];
    return f32x4{v, v, v, v};
  };
  f32x4 acc[8];
#pragma unroll
  for (int nb = 0; nb < 8; nb++)
    acc[nb] = mfma(x, W[(FT0 + nb) * 64 + lane], bias4(bias + CB1, nb));
  auto store_act = [&](f32x4 *a, int nNb) {
#pragma unroll
    for (int nb = 0; nb < 8; nb++) {
      if (nb >= nNb) break;
      const int n = nb * 16 + r;
#pragma unroll
      for (int i = 0; i < 4; i++)
        sAct[(4 * (lane >> 4) + i) * ACT_LD + n] = (__bf16)silu(a[nb][i]);
    }
  };
  auto load_act = [&](int kb) {
    const bf16x8 *p =
        reinterpret_cast<const bf16x8 *>(sAct + r * ACT_LD + kb * 32 + kq);
    return *p;
  };
  store_act(acc, 8);
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): LDS writes done
  __builtin_amdgcn_wave_barrier();
  // layer 2: 128 -> 128
  bf16x8 a4[4];
#pragma unroll
  for (int kb = 0; kb < 4; kb++) a4[kb] = load_act(kb);
#pragma unroll
  for (int nb = 0; nb < 8; nb++) {
    acc[nb] = bias4(bias + CB2, nb);
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
      acc[nb] = mfma(a4[kb], W[(FT1 + kb * 8 + nb) * 64 + lane], acc[nb]);
  }
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
  store_act(acc, 8);
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
  // layer 3: 128 -> 64
#pragma unroll
  for (int kb = 0; kb < 4; kb++) a4[kb] = load_act(kb);
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    acc[nb] = bias4(bias + CB3, nb);
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
      acc[nb] = mfma(a4[kb], W[(FT2 + kb * 4 + nb) * 64 + lane], acc[nb]);
  }
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
  store_act(acc, 4);
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
  // layer 4: 64 -> 8 (padded to 16 columns)
  a4[0] = load_act(0);
  a4[1] = load_act(1);
  f32x4 y = {0.f, 0.f, 0.f, 0.f};
  y = mfma(a4[0], W[(FT3 + 0) * 64 + lane], y);
  y = mfma(a4[1], W[(FT3 + 1) * 64 + lane], y);
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
  if (r < 8) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = r0 + 4 * (lane >> 4) + i;
      if (row < rows) out(row, r, y[i] + bias[CB4 + r]);
    }
  }
}

}  // namespace mlp
}  // namespace bats
