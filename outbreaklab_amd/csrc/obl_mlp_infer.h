// Inference of the microphysics surrogate network on the MFMA units (gfx950,
// v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation).
//
// Each layer is computed transposed, Y^T = W^T X^T: the weights are the A operand (rows = output
// features), the activations of 16 sample rows the B operand (columns = rows). The accumulator of
// an output tile leaves lane (r, g) = (lane % 16, lane / 16) holding features 16 t + 4 g + e,
// e = 0..3, of row r. Two adjacent output tiles therefore give each lane 8 features of its row,
// which is exactly a B operand of the next layer if that layer's reduction index is permuted the
// same way: within each 32-wide k block, slot s of lane group g stands for feature
// 4 g + s (s < 4) or 16 + 4 g + (s - 4) (s >= 4). The weight fragments are packed on the device in
// that order (k_mlp_pack), so activations never leave registers between layers. The first layer
// reads features in natural order (slot s of group g = feature 8 g + s; 16 features, the upper 16
// slots zero).
//
// Quantisation matches the training forward pass with quant on: features, weights and hidden
// activations rounded to BF16 (RNE), biases and accumulation FP32. GELU is the tanh form,
// evaluated as z / (1 + exp(-2u)) with the hardware exp2 and reciprocal.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "obl_mlp.h"

namespace obl {

typedef __bf16 mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mlp_v4f __attribute__((ext_vector_type(4)));

// packed fragments: [layer][mt][kb][64 lanes][8 bf16]; layer 1: 8 x 1, layers 2-3: 8 x 4, layer 4: 1 x 4
constexpr int MLP_FRAG = 64 * 8;
constexpr int MLP_NFRAG = 8 + 32 + 32 + 4;
constexpr int MLP_FRAG_L1 = 0, MLP_FRAG_L2 = 8, MLP_FRAG_L3 = 40, MLP_FRAG_L4 = 72;
constexpr int MLP_PACK_HALVES = MLP_NFRAG * MLP_FRAG;  // bf16 values
constexpr int MLP_BIAS_FLOATS = 3 * MLP_H + MLP_OUT;   // b1 b2 b3 b4

__device__ __forceinline__ int mlp_slot_feature(int layer, int g, int s) {
  return layer == 0 ? 8 * g + s : (s < 4 ? 4 * g + s : 16 + 4 * g + (s - 4));
}

__device__ __forceinline__ uint32_t mlp_bf16_bits(float x) {
  uint32_t u = __float_as_uint(x);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ uint32_t mlp_pack2(float a, float b) { return mlp_bf16_bits(a) | (mlp_bf16_bits(b) << 16); }

__device__ __forceinline__ mlp_bf16x8 mlp_as_bf16x8(uint4 v) {
  union {
    uint4 u;
    mlp_bf16x8 b;
  } c;
  c.u = v;
  return c.b;
}

// tanh-form GELU: 0.5 z (1 + tanh u) = z / (1 + exp(-2 u)), u = sqrt(2/pi) (z + 0.044715 z^3)
__device__ __forceinline__ float mlp_gelu(float z) {
  const float c = -2.0f * 0.7978845608f * 1.4426950409f;  // -2 sqrt(2/pi) log2(e)
  float e = __builtin_amdgcn_exp2f(c * z * (1.0f + 0.044715f * z * z));
  return z * __builtin_amdgcn_rcpf(1.0f + e);
}

// One wave computes the network for 64 rows (4 column tiles of 16). `feat(nt, f)` fills f[0..15]
// with the features of row 16 nt + (lane % 16) of this wave (zeros for rows past the end);
// the outputs (FP32, before output scaling) go to y[64][MLP_OUT + 1] in LDS.
template <class Feat>
__device__ __forceinline__ void mlp_wave64(const uint16_t* frag, const float* bias, const Feat& feat, float* y) {
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15, g = lane >> 4;
  uint4 B[4][4];  // [kb][nt] activations as B operands
#pragma unroll
  for (int kb = 0; kb < 4; kb++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) B[kb][nt] = make_uint4(0u, 0u, 0u, 0u);
  // ---- layer 1 input: features 8 g .. 8 g + 7 of the lane's row (groups 2, 3: zero pad) ----
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    float f[MLP_IN];
    feat(nt, f);
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (g == 0) v = make_uint4(mlp_pack2(f[0], f[1]), mlp_pack2(f[2], f[3]), mlp_pack2(f[4], f[5]), mlp_pack2(f[6], f[7]));
    if (g == 1)
      v = make_uint4(mlp_pack2(f[8], f[9]), mlp_pack2(f[10], f[11]), mlp_pack2(f[12], f[13]), mlp_pack2(f[14], f[15]));
    B[0][nt] = v;
  }
  // ---- hidden layers ----
#pragma unroll 1
  for (int L = 0; L < 3; L++) {
    const int nkb = L == 0 ? 1 : 4;
    const uint16_t* fr = frag + (size_t)(L == 0 ? MLP_FRAG_L1 : (L == 1 ? MLP_FRAG_L2 : MLP_FRAG_L3)) * MLP_FRAG;
    const float* b = bias + L * MLP_H;
    uint4 Bn[4][4];
#pragma unroll
    for (int mt = 0; mt < 8; mt++) {
      mlp_v4f acc[4];
#pragma unroll
      for (int nt = 0; nt < 4; nt++) acc[nt] = mlp_v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; kb++) {
        if (kb < nkb) {
          uint4 a = *reinterpret_cast<const uint4*>(fr + ((size_t)(mt * nkb + kb) * 64 + lane) * 8);
          mlp_bf16x8 av = mlp_as_bf16x8(a);
#pragma unroll
          for (int nt = 0; nt < 4; nt++)
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, mlp_as_bf16x8(B[kb][nt]), acc[nt], 0, 0, 0);
        }
      }
      // bias, GELU, BF16: features 16 mt + 4 g + e -> slots (mt % 2) * 4 + e of k block mt / 2
      float b0 = b[16 * mt + 4 * g], b1 = b[16 * mt + 4 * g + 1], b2 = b[16 * mt + 4 * g + 2], b3 = b[16 * mt + 4 * g + 3];
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        uint32_t p0 = mlp_pack2(mlp_gelu(acc[nt][0] + b0), mlp_gelu(acc[nt][1] + b1));
        uint32_t p1 = mlp_pack2(mlp_gelu(acc[nt][2] + b2), mlp_gelu(acc[nt][3] + b3));
        if ((mt & 1) == 0) {
          Bn[mt >> 1][nt].x = p0;
          Bn[mt >> 1][nt].y = p1;
        } else {
          Bn[mt >> 1][nt].z = p0;
          Bn[mt >> 1][nt].w = p1;
        }
      }
    }
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++) B[kb][nt] = Bn[kb][nt];
  }
  // ---- output layer: 16 outputs ----
  {
    const uint16_t* fr = frag + (size_t)MLP_FRAG_L4 * MLP_FRAG;
    mlp_v4f acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[nt] = mlp_v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      mlp_bf16x8 av = mlp_as_bf16x8(*reinterpret_cast<const uint4*>(fr + ((size_t)kb * 64 + lane) * 8));
#pragma unroll
      for (int nt = 0; nt < 4; nt++)
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, mlp_as_bf16x8(B[kb][nt]), acc[nt], 0, 0, 0);
    }
    const float* b = bias + 3 * MLP_H;
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int e = 0; e < 4; e++) y[(16 * nt + lr) * (MLP_OUT + 1) + 4 * g + e] = acc[nt][e] + b[4 * g + e];
  }
}

void launch_mlp_pack(hipStream_t s, const float* P, uint16_t* frag, float* bias);

}  // namespace obl
