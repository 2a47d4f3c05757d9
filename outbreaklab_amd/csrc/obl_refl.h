// Equivalent radar reflectivity of the model's hydrometeors (gfx950 device functions), shared by
// the surrogate's judge and the 2-D maps kernel: zRain / zGrau (model.js:94-106) and the dBZ of a
// point (model.js:1042-1050); `mar` selects hurricane mode's maritime graupel (model.js:101).
#pragma once
#include "obl_device.h"

namespace obl {

// rain (mm^6 m^-3): Lin power law, or the Thompson exponential size distribution's sixth moment
__device__ __forceinline__ float z_rain(bool thm, float qr, float nr, float rho) {
  if (!thm) return 3.63e9f * pw(fmaxf(rho * qr, 0.0f), 1.75f);
  if (qr < 1e-9f) return 0.0f;
  float l = lamR(qr, nr);
  float nb = qr * l * l * l / 3141.59f;  // the number within the size bounds
  return 720e18f * rho * nb / (l * l * l * l * l * l);
}
// graupel: wet above freezing (larger dielectric factor)
__device__ __forceinline__ float z_grau(bool thm, float qg, float rho, float T, bool mar = false) {
  if (mar) return (T < OBL_T0K ? 1.3e9f : 1.0e10f) * pw(fmaxf(rho * qg, 0.0f), 1.75f);
  if (!thm) return (T < OBL_T0K ? 1.2e10f : 4.33e10f) * pw(fmaxf(rho * qg, 0.0f), 1.75f);
  if (qg < 1e-9f) return 0.0f;
  float l = lamG(qg, rho), l2 = l * l;
  return 720e18f * n0G(qg, rho) / (l2 * l2 * l2 * l) * (T < OBL_T0K ? 0.0666f : 0.24f);
}
// dBZ of one point (rain + snow + graupel), -30 for no echo
__device__ __forceinline__ float dbz_of(bool thm, float4 q1, float4 q2, float rho, float T, bool mar = false) {
  float Z = z_rain(thm, q1.w, q2.z, rho) + 9.80e8f * pw(fmaxf(rho * q2.x, 0.0f), 1.75f) * 0.224f * 4.0f +
            z_grau(thm, q2.y, rho, T, mar);
  return Z > 1e-3f ? 10.0f * logf(Z) / 2.302585f : -30.0f;
}

}  // namespace obl
