// Microphysics surrogate in the model step (no reference counterpart; it stands in for the
// microphysics block, model.js:775-899, at the points the physics kernel hands to it).
//
// The physics kernel, in its surrogate mode, appends every point that holds more than a trace of
// condensate and whose inputs lie inside the network's training envelope to a compact list of 16-float records; every
// other point runs the plain block inline.
// k_surr_apply then runs the network over the list on the MFMA units (obl_mlp_infer.h). The network
// stands in for the block's process part (model.js:775-874: warm rain, ice processes, melting,
// evaporation, sublimation); the block's closing saturation adjustment (model.js:876-889), a few
// exact algebraic iterations, is applied to its result as in the plain block, so cloud edges stay in
// thermodynamic equilibrium. On the network's tendencies, before the adjustment:
//   - a species absent before the block (below the dead zone, a fraction of that output's scale)
//     keeps its amount unless its predicted tendency exceeds the dead zone: a regression network
//     never predicts an exact zero, and clamped noise would otherwise seed thin condensate
//     everywhere;
//   - positivity: every species max(q + dq, 0);
//   - total water: vapour takes the remainder of the point's total water (qv + condensate) before
//     the block, which the block itself conserves; condensate is scaled down if it exceeds it;
//   - temperature from the latent heat of the phase changes, Lv (vapour -> liquid) and Lf
//     (liquid -> ice), so that the step's moist enthalpy is consistent with the water changes;
//   - Thompson Nr from the block's own drop-number rules applied to the network's mass changes
//     (nr_from_mass);
//   - a non-finite prediction falls back to the plain block for that point;
// followed by the physics kernel's own tail (theta = T / pi, the sponge's condensate removal,
// the final clamps).
//
// The judge (k_judge_cols) compares a surrogate step with the plain step from the same inputs,
// column by column: squared differences per channel, column condensate, composite reflectivity
// (zRain / zGrau / snow as model.js:94-106, 1042-1050) and the accumulated surface precipitation.
#include <hip/hip_runtime.h>

#include "obl_device.h"
#include "obl_kernels.h"
#include "obl_refl.h"
#include "obl_micro.h"
#include "obl_mlp_infer.h"

namespace obl {

constexpr int APPLY_THREADS = 512;

__device__ __forceinline__ bool finite4(float a, float b, float c, float d) {
  return isfinite(a) && isfinite(b) && isfinite(c) && isfinite(d);
}

// Thompson Nr after the network's mass changes, by the block's own drop-number rules (model.js:783-866): the
// autoconversion of the point's cloud water brings new drops of 2.68e-10 kg and accretion adds mass without drops
// (both exact, from the state before the block); above 0 C the snow and graupel the network melts bring their
// drops; any other change of the rain mass scales the drops in proportion (a chain of the block's proportional
// losses is one ratio); then self-collection / breakup and the block's bounds. Given the block's own mass
// changes this is the block's Nr; the network's own Nr output is not used (its per-step noise, large against the
// few-per-cent changes of self-collection, random-walks Nr off over a storm's life).
__device__ __forceinline__ float nr_from_mass(const MU& u, float dt, float rho, float T0, float qc0, float qr0,
                                              float qs0, float qg0, float nr0, float qr, float qs, float qg) {
  const float da = fminf(u.v[U_MICRO].x * fmaxf(qc0 - u.v[U_MICRO].y, 0.0f) * dt, qc0);
  float accr = 2.2f * qc0 * pw(qr0, 0.875f);
  if (qr0 > 1e-9f) accr = 3100.0f * rho * nr0 * pw(lamR(qr0, nr0), -2.8f) * qc0 * sqrtf(u.v[U_STEP].z / rho);
  const float dc = fminf(accr * dt, qc0 - da);
  float q1 = qr0 + da + dc, n = nr0 + da / 2.68e-10f;
  if (T0 - OBL_T0K >= 0.0f) {
    const float m1 = fmaxf(qs0 - qs, 0.0f), m2 = fmaxf(qg0 - qg, 0.0f);
    if (qg0 > 1e-9f) n += m2 / qg0 * n0G(qg0, rho) / (rho * lamG(qg0, rho));
    if (qs0 > 1e-9f) n += m1 / qs0 * 2e6f / (rho * pw(3.14159f * 100.0f * 2e6f / (rho * qs0), 0.25f));
    q1 += m1 + m2;
  }
  if (q1 > 1e-12f && qr < q1) n *= qr / q1;
  if (!(qr > 1e-9f)) return 0.0f;
  const float dm = 1.817f / lamR(qr, n);
  const float fb = clampf_(1.0f - ex(2300.0f * (dm - 1.1e-3f)), -2.0f, 1.0f);
  n *= ex(-7.12f * rho * qr * fb * dt);
  return clampf_(n, qr * 3.183e5f, qr * 3.183e11f);
}

__global__ void __launch_bounds__(APPLY_THREADS) k_surr_apply(MU u, SurrApply a) {
  __shared__ __attribute__((aligned(16))) uint16_t F[MLP_PACK_HALVES];
  __shared__ float bias[MLP_BIAS_FLOATS];
  __shared__ float ys[APPLY_THREADS / 64][64 * (MLP_OUT + 1)];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15;
  for (int i = tid; i < MLP_PACK_HALVES / 8; i += APPLY_THREADS)
    reinterpret_cast<uint4*>(F)[i] = reinterpret_cast<const uint4*>(a.frag)[i];
  for (int i = tid; i < MLP_BIAS_FLOATS; i += APPLY_THREADS) bias[i] = a.bias[i];
  __syncthreads();
  Dims d = dims(u);
  const bool thm = u.v[U_MICRO2].z > 0.5f;
  const unsigned n = min(*a.count, a.cap);
  for (unsigned base = blockIdx.x * APPLY_THREADS; base < n; base += gridDim.x * APPLY_THREADS) {
    const unsigned row0 = base + wave * 64;
    auto feat = [&](int nt, float* f) {
      unsigned r = row0 + 16 * nt + lr;
      if (r >= n) {
#pragma unroll
        for (int i = 0; i < MLP_IN; i++) f[i] = 0.0f;
        return;
      }
      const float4* R = reinterpret_cast<const float4*>(a.rec + (size_t)r * SURR_REC);
      float4 r0 = R[0], r1 = R[1], r2 = R[2];
      // T, p0, rho0, qv, qc, qr, qi, qs, qg, Nr, Ni, w, dz, dt
      float x[MLP_NIN] = {r0.y, r0.z, r0.w, r1.x, r1.y, r1.w, r1.z, r2.x, r2.y, r2.z, 0.0f, r2.w, d.dz, d.dt};
      mlp_features(x, f);
    };
    mlp_wave64(F, bias, feat, ys[wave]);
    __syncthreads();
    const unsigned r = row0 + lane;
    if (r < n) {
      const float4* R = reinterpret_cast<const float4*>(a.rec + (size_t)r * SURR_REC);
      float4 r0 = R[0], r1 = R[1], r2 = R[2], r3 = R[3];
      const unsigned c = __float_as_uint(r0.x);
      const float T0 = r0.y, p0 = r0.z, rho = r0.w;
      const float qv0 = r1.x, qc0 = r1.y, qi0 = r1.z, qr0 = r1.w;
      const float qs0 = r2.x, qg0 = r2.y, nr0 = r2.z;
      const float keep = r3.x, pi = r3.y, q2w = r3.z;
      const float* yv = ys[wave] + lane * (MLP_OUT + 1);
      float y[MLP_NOUT];
#pragma unroll
      for (int j = 0; j < MLP_NOUT; j++) y[j] = yv[j] * a.out_scale[j];
      bool ok = finite4(y[0], y[1], y[2], y[3]) && finite4(y[4], y[5], y[6], y[7]) && y[7] < 80.0f;
      float T, qv, qc, qi, qr, qs, qg, nr;
      if (ok) {
        // dead zone for absent species: outputs 2..6 = dqc, dqi, dqr, dqs, dqg
        const float q0s[5] = {qc0, qi0, qr0, qs0, qg0};
#pragma unroll
        for (int k = 0; k < 5; k++)
          if (q0s[k] < a.dead[2 + k] && y[2 + k] < a.dead[2 + k]) y[2 + k] = 0.0f;
        qc = fmaxf(qc0 + y[2], 0.0f);
        qi = fmaxf(qi0 + y[3], 0.0f);
        qr = fmaxf(qr0 + y[4], 0.0f);
        qs = fmaxf(qs0 + y[5], 0.0f);
        qg = fmaxf(qg0 + y[6], 0.0f);
        const float W = qv0 + qc0 + qi0 + qr0 + qs0 + qg0;
        float cond = qc + qi + qr + qs + qg;
        if (cond > W) {
          float f = W / cond;
          qc *= f; qi *= f; qr *= f; qs *= f; qg *= f;
          cond = qc + qi + qr + qs + qg;
        }
        qv = fmaxf(W - cond, 0.0f);
        T = T0 + OBL_LVC / OBL_CPD * (qv0 - qv) + OBL_LFC / OBL_CPD * ((qi + qs + qg) - (qi0 + qs0 + qg0));
        nr = nr0;
        if (thm) nr = nr_from_mass(u, d.dt, rho, T0, qc0, qr0, qs0, qg0, nr0, qr, qs, qg);
        // the block's closing saturation adjustment, exact
        micro_satadj(p0, T, qv, qc, qi);
      } else {
        T = T0; qv = qv0; qc = qc0; qi = qi0; qr = qr0; qs = qs0; qg = qg0; nr = nr0;
        micro_block(u, d.dt, p0, rho, thm, T, qv, qc, qi, qr, qs, qg, nr);
        atomicAdd(a.nfallback, 1u);
      }
      // the physics kernel's tail (model.js:894-905)
      const float th = T / pi;
      qc *= keep; qi *= keep; qr *= keep; qs *= keep; qg *= keep; nr *= keep;
      a.Q1o[c] = make_float4(fmaxf(qv, 0.0f), fmaxf(qc, 0.0f), fmaxf(qi, 0.0f), fmaxf(qr, 0.0f));
      a.Q2o[c] = make_float4(fmaxf(qs, 0.0f), fmaxf(qg, 0.0f), thm ? fmaxf(nr, 0.0f) : nr0, q2w);
      const int x = (int)(c % (unsigned)d.nx), t = (int)(c / (unsigned)d.nx);
      const int yy = t % d.ny, z = t / d.ny;
      reinterpret_cast<float*>(a.Vout)[4 * vidx(d, x, yy, z) + 3] = th;
    }
    __syncthreads();
  }
}

// per-step tallies of the surrogate, then the step's counters back to zero for the next step
__global__ void k_surr_tally(unsigned* count, unsigned cap, unsigned long long* tally) {
  unsigned c = count[0];
  tally[0] += 1ull;
  tally[1] += (unsigned long long)min(c, cap);
  tally[2] += (unsigned long long)count[1];
  tally[3] += (unsigned long long)(c > cap ? c - cap : 0u);
  count[0] = 0u;
  count[1] = 0u;
}
void launch_surr_tally(hipStream_t s, unsigned* count, unsigned cap, unsigned long long* tally) {
  hipLaunchKernelGGL(k_surr_tally, dim3(1), dim3(1), 0, s, count, cap, tally);
}

void launch_surr_apply(hipStream_t s, const MU& u, const SurrApply& a, int blocks) {
  hipLaunchKernelGGL(k_surr_apply, dim3(blocks), dim3(APPLY_THREADS), 0, s, u, a);
}

// ---------------- judge ----------------
// reflectivity helpers: obl_refl.h

// One thread per column (x fastest): per-channel squared differences against the plain step,
// column condensate, composite reflectivity and the surface precipitation rate, accumulated
// into the 2-D precipitation totals; block sums to the double statistics.
__global__ void __launch_bounds__(256) k_judge_cols(MU u, JudgeArgs a) {
  Dims d = dims(u);
  const bool thm = u.v[U_MICRO2].z > 0.5f;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  double acc[JUDGE_NSUM];
#pragma unroll
  for (int i = 0; i < JUDGE_NSUM; i++) acc[i] = 0.0;
  double mx[JUDGE_NMAX] = {0.0, 0.0, 0.0};
  if (x < d.nx) {
    float ccs = 0.0f, ccp = 0.0f, zs = -30.0f, zp = -30.0f;
    for (int z = 0; z < d.nz; z++) {
      size_t c = (size_t)x + (size_t)d.nx * ((size_t)y + (size_t)d.ny * z);
      float4 s1 = a.Q1s[c], s2 = a.Q2s[c], p1 = a.Q1p[c], p2 = a.Q2p[c];
      size_t vi = vidx(d, x, y, z);
      float ths = a.Vs[vi].w, thp = a.Vp[vi].w;
      float sv[8] = {s1.x, s1.y, s1.z, s1.w, s2.x, s2.y, s2.z, ths};
      float pv[8] = {p1.x, p1.y, p1.z, p1.w, p2.x, p2.y, p2.z, thp};
#pragma unroll
      for (int ch = 0; ch < 8; ch++) {
        double df = (double)sv[ch] - (double)pv[ch];
        acc[ch] += df * df;
        acc[8 + ch] += (double)pv[ch] * (double)pv[ch];
      }
      float4 A = bsA(a.BS, z, d.nz), Bz = bsB(a.BS, z, d.nz);
      float rho = A.z;
      ccs += rho * (s1.y + s1.z + s1.w + s2.x + s2.y) * d.dz;
      ccp += rho * (p1.y + p1.z + p1.w + p2.x + p2.y) * d.dz;
      zs = fmaxf(zs, dbz_of(thm, s1, s2, rho, ths * Bz.x));
      zp = fmaxf(zp, dbz_of(thm, p1, p2, rho, thp * Bz.x));
      if (z == 0) {
        // surface precipitation rate (kg m^-2 s^-1): the falling species' mass flux at the lowest level
        float rsfc = u.v[U_STEP].z;
        float rs = rho * (s1.w * vtRain(thm, s1.w, s2.z, rho, rsfc).x + s2.x * vtS(s2.x, rho, rsfc) +
                          s2.y * vtGrau(thm, s2.y, rho, rsfc));
        float rp = rho * (p1.w * vtRain(thm, p1.w, p2.z, rho, rsfc).x + p2.x * vtS(p2.x, rho, rsfc) +
                          p2.y * vtGrau(thm, p2.y, rho, rsfc));
        size_t col = (size_t)x + (size_t)d.nx * y;
        a.prs[col] += (double)rs * d.dt;
        a.prp[col] += (double)rp * d.dt;
      }
    }
    double dc = (double)ccs - (double)ccp;
    acc[16] += dc * dc;
    acc[17] += (double)ccp * (double)ccp;
    mx[0] = fabs(dc);
    mx[1] = (double)ccp;
    if (zs > 0.0f || zp > 0.0f) {
      // echo below 0 dBZ counts as none: a -30 dBZ "no echo" against a 2 dBZ trace is not a 32 dB error
      double dz_ = (double)fmaxf(zs, 0.0f) - (double)fmaxf(zp, 0.0f);
      acc[18] += dz_ * dz_;
      acc[19] += 1.0;
      mx[2] = fabs(dz_);
    }
  }
  // block reduction: wave shuffles, then one atomic per wave
#pragma unroll
  for (int i = 0; i < JUDGE_NSUM; i++) {
    double v = acc[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(&a.sums[i], v);
  }
#pragma unroll
  for (int i = 0; i < JUDGE_NMAX; i++) {
    double v = mx[i];
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o));
    if ((threadIdx.x & 63) == 0 && v > 0.0) atomicMax(&a.maxes[i], (unsigned long long)__double_as_longlong(v));
  }
}

// totals of the accumulated precipitation fields: sum (s - p)^2, sum p^2, sum s, sum p
__global__ void __launch_bounds__(256) k_judge_precip(const double* __restrict__ prs, const double* __restrict__ prp,
                                                      size_t n, double* __restrict__ out) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double s = prs[i], p = prp[i];
    v[0] += (s - p) * (s - p);
    v[1] += p * p;
    v[2] += s;
    v[3] += p;
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    double t = v[k];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o);
    if ((threadIdx.x & 63) == 0 && t != 0.0) atomicAdd(&out[k], t);
  }
}

void launch_judge_cols(hipStream_t s, const MU& u, const Dims& d, const JudgeArgs& a) {
  hipLaunchKernelGGL(k_judge_cols, dim3((d.nx + 255) / 256, d.ny), dim3(256), 0, s, u, a);
}
void launch_judge_precip(hipStream_t s, const double* prs, const double* prp, size_t n, double* out) {
  hipLaunchKernelGGL(k_judge_precip, dim3(256), dim3(256), 0, s, prs, prp, n, out);
}

}  // namespace obl
