// The microphysics surrogate network: sizes, parameter layout, features, activation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace obl {

constexpr int MLP_NIN = 14;   // T, p0, rho0, qv, qc, qr, qi, qs, qg, Nr, Ni, w, dz, dt
constexpr int MLP_IN = 16;    // features (two zero pads)
constexpr int MLP_H = 128;
constexpr int MLP_OUT = 16;   // outputs (9 used)
constexpr int MLP_NOUT = 9;   // dT, dqv, dqc, dqi, dqr, dqs, dqg, dNr, dNi
constexpr int MLP_W1 = 0;
constexpr int MLP_B1 = MLP_W1 + MLP_IN * MLP_H;
constexpr int MLP_W2 = MLP_B1 + MLP_H;
constexpr int MLP_B2 = MLP_W2 + MLP_H * MLP_H;
constexpr int MLP_W3 = MLP_B2 + MLP_H;
constexpr int MLP_B3 = MLP_W3 + MLP_H * MLP_H;
constexpr int MLP_W4 = MLP_B3 + MLP_H;
constexpr int MLP_B4 = MLP_W4 + MLP_H * MLP_OUT;
constexpr int MLP_NPARAM = MLP_B4 + MLP_OUT;

// Input features (kept identical in the CPU oracle): temperature and pressure centred,
// mixing ratios and number concentrations on log scales, the rest scaled to O(1).
__device__ __forceinline__ void mlp_features(const float* x, float* f) {
  f[0] = (x[0] - 250.0f) / 30.0f;
  f[1] = x[1] / 1e5f - 0.6f;
  f[2] = x[2] - 0.7f;
#pragma unroll
  for (int j = 0; j < 6; j++) f[3 + j] = (logf(fmaxf(x[3 + j], 0.0f) + 1e-8f) + 12.0f) / 4.0f;
  f[9] = (logf(fmaxf(x[9], 0.0f) + 1.0f) - 5.0f) / 5.0f;
  f[10] = (logf(fmaxf(x[10], 0.0f) + 1.0f) - 5.0f) / 5.0f;
  f[11] = x[11] / 10.0f;
  f[12] = x[12] / 500.0f - 1.0f;
  f[13] = x[13] / 5.0f - 1.0f;
  f[14] = 0.0f;
  f[15] = 0.0f;
}

// GELU, tanh form
__device__ __forceinline__ float gelu_f(float z) {
  return 0.5f * z * (1.0f + tanhf(0.7978845608f * (z + 0.044715f * z * z * z)));
}
__device__ __forceinline__ float gelu_grad_f(float z) {
  float t = tanhf(0.7978845608f * (z + 0.044715f * z * z * z));
  float du = 0.7978845608f * (1.0f + 3.0f * 0.044715f * z * z);
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * du;
}

struct MlpTrainArgs {
  const float* P;       // parameters
  float* G;             // per-block gradient partials [blocks][MLP_NPARAM] (nullptr: forward / loss only)
  float* lpart;         // per-block loss partials [blocks]
  const float* rows;    // sample rows (14 inputs, 9 outputs, ...)
  int stride;           // floats per row
  size_t nrows;
  const int* idx;       // batch row indices (nullptr: hash-sampled with seed)
  unsigned seed;
  int batch;
  float out_inv_scale[MLP_OUT];
  float loss_scale;     // 1 / (batch * MLP_NOUT)
  float* y_out;         // optional [batch][MLP_OUT] network outputs
  int quant;            // round weights / activations to BF16 as inference does
};

// what the model step needs of a network: its inference form, output scales and envelope
struct MlpView {
  const uint16_t* frag;
  const float* bias;
  float out_scale[MLP_NOUT];
  float lo[MLP_NIN], hi[MLP_NIN];
  float dead_frac;  // dead zone of absent species, as a fraction of each output scale
  float hand_min;   // least total condensate (kg/kg) of a point handed to the network
};

int mlp_grad_blocks(int batch);
// the blocks' partials of a.G / a.lpart, then their sums in block order into G (if a.G) and
// added to *loss (if loss)
void launch_mlp_grad(hipStream_t s, const MlpTrainArgs& a, float* G, float* loss);
void launch_mlp_eval(hipStream_t s, const uint16_t* frag, const float* bias, const float* rows, int stride, int n,
                     float* y);
void launch_adam(hipStream_t s, float* P, float* G, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                 int t);

}  // namespace obl

struct obl_mlp;
namespace obl {
int mlp_view(obl_mlp* m, MlpView* v);
}
