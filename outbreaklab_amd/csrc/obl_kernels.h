// Kernel launchers shared between the kernel files and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include "obl_device.h"

namespace obl {

// ---- model step (k_step.hip) ----
void launch_lsf(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float* BD, float4* lsfA,
                float4* lsfB, float* phT);
void launch_ci(hipStream_t s, const MU& u, const Dims& d, const float* CI, float4* ciMap);
// VX: the initial vortex's table, nz x nr float4 (tangential wind, theta', 0, 0), or null
void launch_init(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* lsfA, const float4* lsfB,
                 const float4* VX, float4* V, float4* Q1, float4* Q2, int seed);
void launch_bubble(hipStream_t s, const MU& u, const Dims& d, float4* V0, float4* V1, double x, double y, double z,
                   double rh, double rz, double dth);
// sfc: the sea surface map (nx * ny float4: land fraction, SST K, 0, 0); null outside hurricane mode
void launch_nut(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, const float4* Q1,
                const float4* sfc, float2* aux);
// outbreak mode, in one launch (k_adv_t): the turbulent viscosity and the wind advection's predictor and corrector
// (V, Q1 -> Vout, aux), and the moisture advection's (V, Q1, Q2 -> Q1o, Q2o)
void launch_adv(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, const float4* Q1,
                const float4* Q2, float4* Vout, float2* aux, float4* Q1o, float4* Q2o);
void launch_phys(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* Vin, const float4* Q1in,
                 const float4* Q2in, const float2* aux, const float4* lsfA, const float4* lsfB, const float* phT,
                 const float4* ciMap, float4* Vout, float4* Q1o, float4* Q2o);
// teacher pairs of the microphysics block: rows of OBL_DUMP_STRIDE floats =
// 14 inputs (T, p0, rho0, qv, qc, qr, qi, qs, qg, Nr, Ni, w, dz, dt),
// 9 outputs (dT, dqv, dqc, dqi, dqr, dqs, dqg, dNr, dNi), then the level index
#define OBL_DUMP_STRIDE 24
struct MicroDump {
  float* buf;
  unsigned* count;
  unsigned cap;
  unsigned salt;
  float p_active;
  float p_clear;
};
void launch_phys_dump(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* Vin, const float4* Q1in,
                      const float4* Q2in, const float2* aux, const float4* lsfA, const float4* lsfB, const float* phT,
                      const float4* ciMap, float4* Vout, float4* Q1o, float4* Q2o, const MicroDump& dump);
// ---- microphysics surrogate in the step (k_step.hip physics surrogate mode, k_surr.hip) ----
// one record per point handed to the surrogate: cell index (bits), T, p0, rho0 | qv, qc, qi, qr |
// qs, qg, Nr, w (cell centre) | sponge keep factor, pi, Q2 spare channel, 0
constexpr int SURR_REC = 16;
struct SurrPhys {
  float* rec;
  unsigned* count;
  unsigned cap;
  float lo[14], hi[14];  // training envelope of the 14 inputs
  float qmin;            // least total condensate of a handed point
};
void launch_phys_surr(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* Vin, const float4* Q1in,
                      const float4* Q2in, const float2* aux, const float4* lsfA, const float4* lsfB, const float* phT,
                      const float4* ciMap, float4* Vout, float4* Q1o, float4* Q2o, const SurrPhys& sp);
struct SurrApply {
  const float* rec;
  const unsigned* count;
  unsigned cap;
  const uint16_t* frag;  // packed BF16 weights (obl_mlp_infer.h)
  const float* bias;
  float out_scale[9];
  float dead[9];  // creation thresholds of absent species (same units as the tendencies)
  float4* Vout;
  float4* Q1o;
  float4* Q2o;
  unsigned* nfallback;
};
void launch_surr_apply(hipStream_t s, const MU& u, const SurrApply& a, int blocks);
void launch_surr_tally(hipStream_t s, unsigned* count, unsigned cap, unsigned long long* tally);
constexpr int JUDGE_NSUM = 20;  // d2[8], r2[8] (qv qc qi qr qs qg Nr theta), cc d2, cc r2, dBZ d2, dBZ columns
constexpr int JUDGE_NMAX = 3;   // max |d column condensate|, max column condensate, max |d dBZ|
struct JudgeArgs {
  const float4* BS;
  const float4 *Vs, *Q1s, *Q2s;  // surrogate step
  const float4 *Vp, *Q1p, *Q2p;  // plain step from the same inputs
  double* prs;                   // accumulated surface precipitation, nx * ny (kg m^-2)
  double* prp;
  double* sums;
  unsigned long long* maxes;     // non-negative doubles as bit patterns
};
void launch_judge_cols(hipStream_t s, const MU& u, const Dims& d, const JudgeArgs& a);
void launch_judge_precip(hipStream_t s, const double* prs, const double* prp, size_t n, double* out);

// ---- hurricane mode (model.js:88-119 and the hurOn() / cn.w branches, WGSL_M_PREP, WGSL_M_FILT) ----
struct HurArgs {
  const float4* sfc;  // sea surface map, nx * ny float4 (land fraction, SST K, 0, 0)
  const float* phi;   // the last solve's padded phi (the sea fluxes' surface pressure)
  const float4* V;    // the physics kernel's input winds (lowest-level wind speed)
};
void launch_phys_hur(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* Vin, const float4* Q1in,
                     const float4* Q2in, const float2* aux, const float4* lsfA, const float4* lsfB, const float* phT,
                     const float4* ciMap, float4* Vout, float4* Q1o, float4* Q2o, const HurArgs& hu);
void launch_advv1_hur(hipStream_t s, const MU& u, const Dims& d, const float4* V, const float4* Vval, float4* Vt);
void launch_advv2_hur(hipStream_t s, const MU& u, const Dims& d, const float4* V, const float4* Vval, const float4* Vt,
                      float4* Vout);
void launch_advq1_hur(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, const float4* Q1,
                      const float4* Q2, float4* Q1t, float4* Q2t);
void launch_advq2_hur(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, const float4* Q1,
                      const float4* Q2, const float4* Q1t, const float4* Q2t, float4* Q1o, float4* Q2o);
void launch_prep(hipStream_t s, const MU& u, const Dims& d, const float4* V, const float* phi, float4* Vp);
void launch_filt(hipStream_t s, const MU& u, const Dims& d, const float4* V, const float4* Q1, const float4* Q2,
                 float4* Vo, float4* Q1o, float4* Q2o);

void launch_div(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, float* rhs);
// the divergence and the fold-and-split of the multi-piece solve's x forward operand in one kernel (padded Nx and Ny
// multiples of 8: div_fold_fits)
bool div_fold_fits(int Nx, int Ny);
void launch_div_fold(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, float* rhs, int P,
                     uint16_t* out, int RB, int KB);
void launch_proj(hipStream_t s, const MU& u, const Dims& d, const float4* V, const float* phi, float4* Vout);
// 2-D maps of the current state (model.js:1028-1100): out = 4 planar maps of nx*ny float4
void launch_maps(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, const float4* Q1,
                 const float4* Q2, const float* phi, const float4* sfc, float4* out);

// ---- pressure solve (k_solve.hip) ----
struct DstArgs {
  const float* in;
  float* out;
  const float* W0;  // even-mode matrix for this pass, [Npad][Kp] (zero-padded)
  const float* W1;  // odd-mode matrix
  int N;            // transform length (even)
  int Nh;           // N / 2
  int rows;         // X passes: rows (nyp * nz); Y passes: planes (nz)
  int inner;        // Y passes: row length (nxp); X passes: N
  int Kp;           // row stride of W0/W1
};
enum { DST_XF = 0, DST_XI = 1, DST_YF = 2, DST_YI = 3 };
void launch_dst(hipStream_t s, int mode, const DstArgs& a);

// Multi-piece BF16 passes (operands pre-split in the tiled layout of k_solve.hip). Pass chain:
//   fold_x: rhs -> A_xf;  XF: A_xf -> A_yf;  YF: A_yf -> spectral [z][sx][sy];  Thomas -> A_yi;
//   YI: A_yi -> A_xi;  XI: A_xi -> phi
constexpr int MX_MF = 4;  // 32-row fragments per wave: a block covers 32 MX_MF rows of A
struct MxArgs {
  const uint16_t* A;  // data operand, [2 parities][P pieces][aRB][KB + 1] tiles (one pitch tile unused)
  const uint16_t* B;  // sine matrices, [2][P][bRB][KB + 1] tiles
  int aRB, bRB, KB;   // 32-row blocks of A and of B per piece plane; 16-wide k-blocks (even)
  int Mvalid;         // rows of A that carry data (XF: (z, y) pairs)
  int tilesN;         // column tiles (64 NF wide) per parity
  int Nx, Ny, nz;     // padded grid
  uint16_t* outT;     // tiled output (XF, YI)
  int oRB, oKB;
  float* outF;        // FP32 output (YF: spectral array; XI: phi)
};
void launch_dst_mx(hipStream_t s, int mode, int P, int NF, const MxArgs& a);
void launch_mx_fold_x(hipStream_t s, int P, const float* rhs, uint16_t* out, int Nx, int Ny, int nz, int RB, int KB);

struct ThomasArgs {
  float* hat;
  float* cp32;
  double* cp64;
  double* dp64;
  const float* lam32;   // permuted: [nx] x-modes then [ny] y-modes
  const double* lam64;
  const float4* zop32;  // (lo, up, diag, rho_c) per level
  const double* zop64;  // the same in FP64, 4 per level
  int nx, ny, nz;
  int mode;             // 0 FP32 (reference order), 1 FP64 all, 2 FP64 where near-singular
  double shift_lo;      // mode 2: FP64 when |shift| < shift_lo
  int layout;           // 0: hat[z][sy][sx]; 1: hat[z][sx][sy] (the multi-piece BF16 chain)
  uint16_t* outT;       // non-null: the solution goes here as the y inverse pass's tiled split operand
  int P, oKB;           //   its pieces and k-blocks (rows z nx + sx, parity sy >= ny / 2, k = sy mod ny / 2)
  const int2* modes64;  // modes 1 and 2 with nz <= 104: the (sx, sy) modes solved in FP64 (mode 1: every mode),
  int n64;              //   in a kernel of their own; the FP32 kernel leaves them alone
};
enum { THOMAS_FP32 = 0, THOMAS_FP64 = 1, THOMAS_MIXED = 2 };
void launch_thomas(hipStream_t s, const ThomasArgs& a);

}  // namespace obl
