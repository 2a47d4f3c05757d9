// The microphysics block of the physics kernel (model.js:775-899), shared by the physics
// kernel and by the surrogate's fallback path.
#pragma once
#include "obl_device.h"

namespace obl {

// The process part of the microphysics block (model.js:775-874): warm rain, ice-phase
// conversions, melting, rain evaporation, Thompson self-collection / breakup, sublimation.
// T, qv..qg, nr in and out.
__device__ __forceinline__ void micro_processes(const MU& u, float dt, float p0, float rho, bool thm, float& T, float& qv,
                                                float& qc, float& qi, float& qr, float& qs, float& qg, float& nr) {
  float rsfc = u.v[U_STEP].z;
  float rq = 1000.0f * rho;
  float rs = sqrtf(rsfc / rho);
  float pauto = u.v[U_MICRO].x * fmaxf(qc - u.v[U_MICRO].y, 0.0f);
  float accr = 2.2f * qc * pw(qr, 0.875f);
  if (thm && qr > 1e-9f) {
    float l = lamR(qr, nr);
    accr = 3100.0f * rho * nr * pw(l, -2.8f) * qc * rs;
  }
  float da = fminf(pauto * dt, qc);
  float dc = fminf(accr * dt, qc - da);
  qc -= da + dc;
  qr += da + dc;
  nr += da / 2.68e-10f;
  float dTc = T - OBL_T0K;
  if (dTc < 0.0f) {
    float e25 = ex(0.025f * dTc);
    float pqs = pw(fmaxf(rq * qs, 0.0f), 0.81f);
    float pqg = pw(fmaxf(rq * qg, 0.0f), 0.875f);
    float psaut = 1e-3f * e25 * fmaxf(qi - 6e-4f, 0.0f);
    float psacw = 5.4e-3f * qc * pqs * rs;
    float psaci = 5.4e-3f * qi * pqs * e25 * rs;
    float ewet = 1.0f / (1.0f + qg / 0.008f);
    float pgacw = 3.1e-3f * qc * pqg * rs * ewet;
    float pgaci = 3.0e-4f * qi * pqg * rs;
    float pgacr = 3.1e-3f * qr * pqg * rs * ewet;
    float pgfr = fminf(2e-4f * u.v[U_MICRO].z * (ex(0.66f * fminf(-dTc, 20.0f)) - 1.0f), 0.08f) * qr;
    float pgaut = 1e-3f * ex(0.09f * dTc) * fmaxf(qs - 6e-4f, 0.0f) * (thm ? 0.25f : 1.0f);
    float a1 = fminf((psaut + psaci + pgaci) * dt, qi);
    float fs1 = (psaut + psaci) / fmaxf(psaut + psaci + pgaci, 1e-20f);
    qi -= a1;
    qs += a1 * fs1;
    qg += a1 * (1.0f - fs1);
    float a2 = fminf((psacw + pgacw) * dt, qc);
    float fs2 = psacw / fmaxf(psacw + pgacw, 1e-20f);
    float rimeS = thm ? 0.95f : 0.7f;
    qc -= a2;
    qs += a2 * fs2 * rimeS;
    qg += a2 * (1.0f - fs2 * rimeS);
    float a3 = fminf((pgacr + pgfr) * dt, qr);
    if (T < 235.0f) a3 = qr;
    if (qr > 1e-12f) nr *= (qr - a3) / qr;
    qr -= a3;
    qg += a3;
    float a4 = fminf(pgaut * dt, qs);
    qs -= a4;
    qg += a4;
    T += OBL_LFC / OBL_CPD * (a2 + a3);
  } else {
    float m1 = fminf(0.012f * dTc * qs * dt, qs);
    float m2 = fminf(0.0022f * dTc * qg * dt, qg);
    if (thm) {
      if (qg > 1e-9f) {
        float l = lamG(qg, rho), n0 = n0G(qg, rho);
        float vent = 0.78f / (l * l) + 0.31f * 0.843f * 1.89f * 5430.0f * pw(rsfc / rho, 0.25f) * pw(l, -2.945f);
        float heat = 2.4e-2f * dTc - OBL_LVC * 2.26e-5f * rho * (qvsat(OBL_T0K, p0, 0.0f) - qv);
        m2 = clampf_(6.2832f / (rho * OBL_LFC) * heat * n0 * vent * dt, 0.0f, qg);
        nr += m2 / qg * n0 / (rho * l);
      }
      if (qs > 1e-9f) {
        float ls = pw(3.14159f * 100.0f * 2e6f / (rho * qs), 0.25f);
        nr += m1 / qs * 2e6f / (rho * ls);
      }
    }
    qs -= m1;
    qg -= m2;
    qr += m1 + m2;
    T -= OBL_LFC / OBL_CPD * (m1 + m2);
  }
  float qvsl = qvsat(T, p0, 0.0f);
  if (qr > 0.0f && qv < qvsl) {
    float rc = 0.001f * rho * qr;
    float ern = ((1.6f + 124.9f * pw(rc, 0.2046f)) * pw(rc, 0.525f) / (2.55e8f / (p0 * qvsl) + 5.4e5f)) *
                ((qvsl - qv) / (0.001f * rho * qvsl));
    float srt = u.v[U_MICRO2].y > 0.0f ? sqrtf(rc / (rc + 0.001f * rho * u.v[U_MICRO2].y)) : 1.0f;
    float e = ern * srt * dt * u.v[U_TURB].z;
    if (thm) {
      float l = lamR(qr, nr), n0 = rho * nr * l;
      float Ap = OBL_LVC * OBL_LVC / (2.4e-2f * OBL_RVC * T * T), Bp = 1.0f / (rho * qvsl * 2.26e-5f);
      float vent = 0.78f / (l * l) + 0.31f * 0.843f * 1.827f * 7492.0f * pw(rsfc / rho, 0.25f) * pw(l, -2.9f);
      e = 6.2832f * (1.0f - qv / qvsl) * n0 * vent / (rho * (Ap + Bp)) * dt * u.v[U_MICRO2].w;
    }
    e = clampf_(e, 0.0f, fminf(qr, qvsl - qv));
    if (thm && qr > 1e-12f) nr *= (qr - e) / qr;
    qr -= e;
    qv += e;
    T -= OBL_LVC / OBL_CPD * e;
  }
  if (thm) {
    if (qr > 1e-9f) {
      float dm = 1.817f / lamR(qr, nr);
      float fb = clampf_(1.0f - ex(2300.0f * (dm - 1.1e-3f)), -2.0f, 1.0f);
      nr *= ex(-7.12f * rho * qr * fb * dt);
      nr = clampf_(nr, qr * 3.183e5f, qr * 3.183e11f);
    } else {
      nr = 0.0f;
    }
  }
  float qvsi = qvsat(T, p0, 1.0f);
  if (dTc < 0.0f && qv < qvsi) {
    float sub = fmaxf(1.0f - qv / qvsi, 0.0f);
    float e = fminf((2e-3f * qs + 5e-4f * qg) * sub * dt, (qvsi - qv) * 0.5f);
    float fsn = qs / fmaxf(qs + qg, 1e-20f);
    e = fminf(e, qs + qg);
    qs -= e * fsn;
    qg -= e * (1.0f - fsn);
    qv += e;
    T -= OBL_LSC / OBL_CPD * e;
  }
}

// The block's closing step (model.js:876-889): the 3-iteration mixed-phase saturation adjustment
// (Hsie 1980) of vapour against cloud water and ice.
__device__ __forceinline__ void micro_satadj(float p0, float& T, float& qv, float& qc, float& qi) {
  float qi0 = qi, qcl = qc + qi, dsum = 0.0f;
#pragma unroll
  for (int it = 0; it < 3; it++) {
    float f = fIce(T);
    float qs_ = qvsat(T, p0, f);
    float L = mixf(OBL_LVC, OBL_LSC, f);
    float dd = (qv - qs_) / (1.0f + L * L * qs_ / (OBL_CPD * OBL_RVC * T * T));
    dd = fmaxf(dd, -qcl);
    qv -= dd;
    qcl += dd;
    T += L / OBL_CPD * dd;
    dsum += dd;
  }
  float f = fIce(T);
  qi = f * qcl;
  qc = qcl - qi;
  T += OBL_LFC / OBL_CPD * (qi - qi0 - f * dsum);
}

// The whole block (model.js:775-899).
__device__ __forceinline__ void micro_block(const MU& u, float dt, float p0, float rho, bool thm, float& T, float& qv,
                                            float& qc, float& qi, float& qr, float& qs, float& qg, float& nr) {
  micro_processes(u, dt, p0, rho, thm, T, qv, qc, qi, qr, qs, qg, nr);
  micro_satadj(p0, T, qv, qc, qi);
}

// Points handed to the surrogate: more than a trace of condensate (1e-9 kg/kg, the block's own
// rain threshold). Without condensate every process rate of the block is zero and only the
// saturation adjustment can act; traces (advection leaves 1e-14 .. 1e-36 behind) carry no
// information a network could use. The plain block handles both exactly at no real cost.
constexpr float MICRO_TRACE = 1e-9f;
__device__ __forceinline__ bool micro_active(float qc, float qi, float qr, float qs, float qg) {
  return qc + qi + qr + qs + qg > MICRO_TRACE;
}

}  // namespace obl
