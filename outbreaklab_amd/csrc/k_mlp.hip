// Microphysics surrogate: a 16 -> 128 -> 128 -> 128 -> 16 MLP (GELU) standing in for the
// microphysics block (model.js:775-899). No reference counterpart.
//
// Training (this file): one block takes 16 sample rows through the forward pass and back,
// in FP32 arithmetic; with `quant` on, weights, features and hidden activations are rounded
// to BF16 exactly where the inference kernel feeds them to BF16 MFMA (straight-through
// gradients). Each block writes its own gradient and loss partials; k_mlp_reduce sums them in block
// order, so training is reproducible run to run; k_adam applies Adam to the FP32 master parameters.
//
// Parameter layout (FP32): W1[16][128] b1[128] W2[128][128] b2[128] W3[128][128] b3[128]
// W4[128][16] b4[16] (W[k][c]: input k, output c).
#include <hip/hip_runtime.h>

#include "obl_device.h"
#include "obl_kernels.h"
#include "obl_mlp.h"
#include "obl_mlp_infer.h"

namespace obl {

__device__ __forceinline__ float bf16_round(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xffff0000u);
}

constexpr int MR = 16;  // rows per training block

__device__ __forceinline__ float qz(float x, bool q) { return q ? bf16_round(x) : x; }

__global__ void __launch_bounds__(256) k_mlp_grad(MlpTrainArgs a) {
  __shared__ float f[MR][MLP_IN], tg[MR][MLP_OUT], yv[MR][MLP_OUT];
  __shared__ float z1[MR][MLP_H], z2[MR][MLP_H], z3[MR][MLP_H], hb[MR][MLP_H], db[MR][MLP_H];
  __shared__ float sq[MR][MLP_OUT];
  const int tid = threadIdx.x;
  const int row0 = blockIdx.x * MR;
  const bool q = a.quant != 0;
  const float* P = a.P;
  if (tid < MR) {
    int r = row0 + tid;
    bool ok = r < a.batch;
    size_t src = 0;
    if (ok) src = a.idx ? (size_t)a.idx[r] : (size_t)(hash3i(r, (int)a.seed, 0x5bd1e995) % (uint32_t)a.nrows);
    const float* row = a.rows + src * (size_t)a.stride;
    float x[MLP_NIN];
#pragma unroll
    for (int i = 0; i < MLP_NIN; i++) x[i] = ok ? row[i] : 0.0f;
    float ff[MLP_IN];
    mlp_features(x, ff);
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) f[tid][i] = ok ? qz(ff[i], q) : 0.0f;
#pragma unroll
    for (int j = 0; j < MLP_OUT; j++) tg[tid][j] = (ok && j < MLP_NOUT) ? row[MLP_NIN + j] * a.out_inv_scale[j] : 0.0f;
  }
  __syncthreads();
  const int c = tid & 127, rg = tid >> 7;  // column, row group (8 rows)
  // ---- forward ----
  {
    float acc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) acc[i] = P[MLP_B1 + c];
    for (int k = 0; k < MLP_IN; k++) {
      float w = qz(P[MLP_W1 + k * MLP_H + c], q);
#pragma unroll
      for (int i = 0; i < 8; i++) acc[i] += f[rg * 8 + i][k] * w;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) z1[rg * 8 + i][c] = acc[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 8; i++) hb[rg * 8 + i][c] = qz(gelu_f(z1[rg * 8 + i][c]), q);
  __syncthreads();
  {
    float acc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) acc[i] = P[MLP_B2 + c];
    for (int k = 0; k < MLP_H; k++) {
      float w = qz(P[MLP_W2 + k * MLP_H + c], q);
#pragma unroll
      for (int i = 0; i < 8; i++) acc[i] += hb[rg * 8 + i][k] * w;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) z2[rg * 8 + i][c] = acc[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 8; i++) hb[rg * 8 + i][c] = qz(gelu_f(z2[rg * 8 + i][c]), q);
  __syncthreads();
  {
    float acc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) acc[i] = P[MLP_B3 + c];
    for (int k = 0; k < MLP_H; k++) {
      float w = qz(P[MLP_W3 + k * MLP_H + c], q);
#pragma unroll
      for (int i = 0; i < 8; i++) acc[i] += hb[rg * 8 + i][k] * w;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) z3[rg * 8 + i][c] = acc[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 8; i++) hb[rg * 8 + i][c] = qz(gelu_f(z3[rg * 8 + i][c]), q);
  __syncthreads();
  {
    // output layer: one (row, output) per thread
    int r = tid >> 4, j = tid & 15;
    float acc = P[MLP_B4 + j];
    for (int k = 0; k < MLP_H; k++) acc += hb[r][k] * qz(P[MLP_W4 + k * MLP_OUT + j], q);
    yv[r][j] = acc;
    bool ok = row0 + r < a.batch && j < MLP_NOUT;
    float dd = ok ? acc - tg[r][j] : 0.0f;
    if (a.y_out && row0 + r < a.batch) a.y_out[(size_t)(row0 + r) * MLP_OUT + j] = acc;
    sq[r][j] = dd * dd;
    // dy, kept in yv
    __syncthreads();
    yv[r][j] = 2.0f * dd * a.loss_scale;
  }
  __syncthreads();
  if (tid == 0 && a.lpart) {
    float l = 0.0f;
    for (int r = 0; r < MR; r++)
      for (int j = 0; j < MLP_OUT; j++) l += sq[r][j];
    a.lpart[blockIdx.x] = l;
  }
  if (!a.G) return;
  float* G = a.G + (size_t)blockIdx.x * MLP_NPARAM;  // this block's partial gradient
  // ---- backward ----
  // dW4[k][j] = sum_r h3[r][k] dy[r][j]; db4[j] = sum_r dy[r][j]
  for (int e = tid; e < MLP_H * MLP_OUT; e += 256) {
    int k = e >> 4, j = e & 15;
    float s = 0.0f;
#pragma unroll
    for (int r = 0; r < MR; r++) s += hb[r][k] * yv[r][j];
    G[MLP_W4 + e] = s;
  }
  if (tid < MLP_OUT) {
    float s = 0.0f;
#pragma unroll
    for (int r = 0; r < MR; r++) s += yv[r][tid];
    G[MLP_B4 + tid] = s;
  }
  // dh3[r][k] = sum_j dy[r][j] W4[k][j]; dz3 = dh3 * gelu'(z3)
#pragma unroll
  for (int i = 0; i < 8; i++) {
    int r = rg * 8 + i;
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < MLP_OUT; j++) s += yv[r][j] * qz(P[MLP_W4 + c * MLP_OUT + j], q);
    db[r][c] = s * gelu_grad_f(z3[r][c]);
  }
  __syncthreads();
  // layers 3, 2: dW[k][c] = sum_r h_in[r][k] dz[r][c]; dh_in[r][k] = sum_c dz[r][c] W[k][c]
  const float* zin[2] = {&z2[0][0], &z1[0][0]};
  const int woff[2] = {MLP_W3, MLP_W2}, boff[2] = {MLP_B3, MLP_B2};
  for (int L = 0; L < 2; L++) {
    // h_in = gelu(z_in) (quantised as in the forward pass)
#pragma unroll
    for (int i = 0; i < 8; i++) hb[rg * 8 + i][c] = qz(gelu_f(zin[L][(rg * 8 + i) * MLP_H + c]), q);
    __syncthreads();
    for (int kk = 0; kk < 64; kk++) {
      int k = rg * 64 + kk;
      float s = 0.0f;
#pragma unroll
      for (int r = 0; r < MR; r++) s += hb[r][k] * db[r][c];
      G[woff[L] + k * MLP_H + c] = s;
    }
    if (tid < MLP_H) {
      float s = 0.0f;
#pragma unroll
      for (int r = 0; r < MR; r++) s += db[r][tid];
      G[boff[L] + tid] = s;
    }
    float nd[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      int r = rg * 8 + i;
      float s = 0.0f;
      for (int cc = 0; cc < MLP_H; cc++) s += db[r][cc] * qz(P[woff[L] + c * MLP_H + cc], q);
      nd[i] = s * gelu_grad_f(zin[L][r * MLP_H + c]);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) db[rg * 8 + i][c] = nd[i];
    __syncthreads();
  }
  // layer 1: dW1[k][c] = sum_r f[r][k] dz1[r][c]
  for (int e = tid; e < MLP_IN * MLP_H; e += 256) {
    int k = e >> 7, cc = e & 127;
    float s = 0.0f;
#pragma unroll
    for (int r = 0; r < MR; r++) s += f[r][k] * db[r][cc];
    G[MLP_W1 + e] = s;
  }
  if (tid < MLP_H) {
    float s = 0.0f;
#pragma unroll
    for (int r = 0; r < MR; r++) s += db[r][tid];
    G[MLP_B1 + tid] = s;
  }
}

// G[p] = sum over blocks b of Gpart[b][p], in block order; loss = sum of the blocks' loss partials
__global__ void k_mlp_reduce(const float* __restrict__ Gpart, const float* __restrict__ lpart, int nblk,
                             float* __restrict__ G, float* __restrict__ loss) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < MLP_NPARAM && Gpart) {
    float s = 0.0f;
    for (int b = 0; b < nblk; b++) s += Gpart[(size_t)b * MLP_NPARAM + p];
    G[p] = s;
  }
  if (p == 0 && loss) {
    float l = 0.0f;
    for (int b = 0; b < nblk; b++) l += lpart[b];
    *loss += l;
  }
}

__global__ void k_adam(float* P, float* G, float* m, float* v, int n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float g = G[i];
  float mi = b1 * m[i] + (1.0f - b1) * g;
  float vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  G[i] = 0.0f;
}

// ---------------- inference-form packing and evaluation ----------------
// BF16 weight fragments in the order of obl_mlp_infer.h, and the FP32 biases, from the master
// parameters (one thread per fragment element).
__global__ void k_mlp_pack(const float* __restrict__ P, uint16_t* __restrict__ F, float* __restrict__ bias) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < MLP_BIAS_FLOATS) {
    int L = t / MLP_H, o = t % MLP_H;
    const int boff[4] = {MLP_B1, MLP_B2, MLP_B3, MLP_B4};
    bias[t] = P[boff[L] + o];
  }
  if (t >= MLP_PACK_HALVES) return;
  int fi = t / MLP_FRAG, w = t % MLP_FRAG, lane = w / 8, s = w % 8;
  int layer, mt, kb;
  if (fi < MLP_FRAG_L2) { layer = 0; mt = fi; kb = 0; }
  else if (fi < MLP_FRAG_L3) { layer = 1; mt = (fi - MLP_FRAG_L2) / 4; kb = (fi - MLP_FRAG_L2) % 4; }
  else if (fi < MLP_FRAG_L4) { layer = 2; mt = (fi - MLP_FRAG_L3) / 4; kb = (fi - MLP_FRAG_L3) % 4; }
  else { layer = 3; mt = 0; kb = fi - MLP_FRAG_L4; }
  int lr = lane & 15, g = lane >> 4;
  int k = 32 * kb + mlp_slot_feature(layer, g, s), o = 16 * mt + lr;
  const int kin = layer == 0 ? MLP_IN : MLP_H, nout = layer == 3 ? MLP_OUT : MLP_H;
  const int woff[4] = {MLP_W1, MLP_W2, MLP_W3, MLP_W4};
  float v = k < kin ? P[woff[layer] + k * nout + o] : 0.0f;
  F[t] = (uint16_t)mlp_bf16_bits(v);
}

// Raw outputs of the inference form for sample rows (14 inputs each, `stride` floats apart):
// y[row][MLP_OUT]. 256 threads, 64 rows per wave.
__global__ void __launch_bounds__(256) k_mlp_eval(const uint16_t* __restrict__ F, const float* __restrict__ bias,
                                                  const float* __restrict__ rows, int stride, int n,
                                                  float* __restrict__ y) {
  __shared__ float ys[4][64 * (MLP_OUT + 1)];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15;
  const int row0 = (blockIdx.x * 4 + wave) * 64;
  auto feat = [&](int nt, float* f) {
    int r = row0 + 16 * nt + lr;
    float x[MLP_NIN];
#pragma unroll
    for (int i = 0; i < MLP_NIN; i++) x[i] = r < n ? rows[(size_t)r * stride + i] : 0.0f;
    mlp_features(x, f);
    if (r >= n)
#pragma unroll
      for (int i = 0; i < MLP_IN; i++) f[i] = 0.0f;
  };
  mlp_wave64(F, bias, feat, ys[wave]);
  __syncthreads();
  int r = row0 + lane;
  if (r < n)
#pragma unroll
    for (int j = 0; j < MLP_OUT; j++) y[(size_t)r * MLP_OUT + j] = ys[wave][lane * (MLP_OUT + 1) + j];
}

void launch_mlp_pack(hipStream_t s, const float* P, uint16_t* frag, float* bias) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_PACK_HALVES + 255) / 256), dim3(256), 0, s, P, frag, bias);
}
void launch_mlp_eval(hipStream_t s, const uint16_t* frag, const float* bias, const float* rows, int stride, int n,
                     float* y) {
  hipLaunchKernelGGL(k_mlp_eval, dim3((n + 255) / 256), dim3(256), 0, s, frag, bias, rows, stride, n, y);
}

int mlp_grad_blocks(int batch) { return (batch + MR - 1) / MR; }
void launch_mlp_grad(hipStream_t s, const MlpTrainArgs& a, float* G, float* loss) {
  const int nblk = mlp_grad_blocks(a.batch);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, a.G, a.lpart, nblk, G, loss);
}
void launch_adam(hipStream_t s, float* P, float* G, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                 int t) {
  float c1 = 1.0f / (1.0f - powf(b1, (float)t)), c2 = 1.0f / (1.0f - powf(b2, (float)t));
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, P, G, m, v, n, lr, b1, b2, eps, c1, c2);
}

}  // namespace obl
