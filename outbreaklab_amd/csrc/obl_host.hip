// Host-side builders shared by the Python and Node hosts (no device code): the initial vortex's table of
// hurricane mode (setVortex, model.js:1318-1365). Both hosts call this one routine, so a run set up from either
// host starts from the same table.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/outbreak_hip.h"

#pragma clang fp contract(off)

namespace {

double smooth01(double a, double b, double x) {
  double t = (x - a) / (b - a);
  t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  return t * t * (3.0 - 2.0 * t);
}

// vertical structure F(z) of the tangential wind: the deep form 1 - (z/H - sin(2 pi z/H)/(2 pi)) up to H = ztop,
// or a mid-level vortex exp(-((z - 5 km) / 2.5 km)^2)
double vertical_shape(double z, double ztop, int mid) {
  if (mid) {
    double s = (z - 5000.0) / 2500.0;
    return exp(-(s * s));
  }
  if (z >= ztop) return 0.0;
  return 1.0 - (z / ztop - sin(2.0 * M_PI * z / ztop) / (2.0 * M_PI));
}

}  // namespace

extern "C" int obl_vortex_table(int nz, double dz, double dx, double lat, const float* B, const double* spec,
                                int mid, float* tab, size_t tab_floats, double* info) {
  if (nz < 2 || !(dz > 0) || !(dx > 0) || !B || !spec || !tab || !info) return -1;
  const double vmax = spec[0], rmw = spec[1], r0 = spec[2], ztop = spec[3];
  if (!(rmw > 0) || !(r0 > 0) || !(ztop > 0)) return -1;
  const double f = 2.0 * 7.292e-5 * sin(lat * M_PI / 180.0);
  const double dr = dx / 2 > 500.0 ? dx / 2 : 500.0;
  int nr = (int)ceil(r0 * 2.0 / dr) + 3;
  if (nr > OBL_VORTEX_MAX_RADII) nr = OBL_VORTEX_MAX_RADII;
  if (tab_floats < (size_t)nz * nr * 4) return -2;

  // Radial profile: a smooth Rankine-type wind vmax (1 + a) x / (a + x^(1 + a)), x = r / rmw, a = 0.6, tapered
  // to zero between r0 / 2 and r0, then held inertially stable -- marching outward, the absolute angular momentum
  // M = r v + |f| r^2 / 2 may not fall, so where the taper would lower it v keeps M (or drops to 0). The sign
  // makes the vortex cyclonic in either hemisphere.
  const double fa = fabs(f), a = 0.6, sg = f < 0 ? -1.0 : 1.0;
  std::vector<double> vr(nr);
  double M = 0.0;
  for (int i = 0; i < nr; i++) {
    const double r = i * dr, x = r / rmw;
    double s = vmax * (1.0 + a) * x / (a + pow(x, 1.0 + a)) * (1.0 - smooth01(0.5 * r0, r0, r));
    if (r > 0) {
      const double hold = (M - 0.5 * fa * r * r) / r;
      if (hold > s) s = hold;
      if (s < 0.0) s = 0.0;
    }
    M = r * s + 0.5 * fa * r * r;
    vr[i] = sg * s;
  }

  // Gradient-wind balance of the separable wind v(r, z) = V(r) F(z): d phi / dr = v^2 / r + f v, integrated
  // inward from the table's rim (phi = 0 there) by the trapezoid rule, splits into F^2 P2(r) + F P1(r) with
  // P2' = V^2 / r and P1' = f V, so the radial integrals are done once for every level.
  std::vector<double> P2(nr, 0.0), P1(nr, 0.0);
  for (int i = nr - 2; i >= 0; i--) {
    const double c2a = i == 0 ? 0.0 : vr[i] * vr[i] / (i * dr), c2b = vr[i + 1] * vr[i + 1] / ((i + 1) * dr);
    const double c1a = i == 0 ? 0.0 : f * vr[i], c1b = f * vr[i + 1];
    P2[i] = P2[i + 1] - 0.5 * (c2a + c2b) * dr;
    P1[i] = P1[i + 1] - 0.5 * (c1a + c1b) * dr;
  }
  std::vector<double> F(nz);
  for (int k = 0; k < nz; k++) F[k] = vertical_shape((k + 0.5) * dz, ztop, mid);

  // hydrostatic warm core: theta' = theta0 (d phi / dz) / g, centred differences (one-sided at the ends)
  for (int k = 0; k < nz; k++) {
    const int k0 = k > 0 ? k - 1 : 0, k1 = k < nz - 1 ? k + 1 : nz - 1;
    const double dF2 = F[k1] * F[k1] - F[k0] * F[k0], dF1 = F[k1] - F[k0], h = (k1 - k0) * dz;
    const double th0 = (double)B[(size_t)k * 8];
    float* row = tab + (size_t)k * nr * 4;
    for (int i = 0; i < nr; i++) {
      row[4 * i] = (float)(vr[i] * F[k]);
      row[4 * i + 1] = (float)(th0 * ((dF2 * P2[i] + dF1 * P1[i]) / h) / 9.81);
      row[4 * i + 2] = 0.0f;
      row[4 * i + 3] = 0.0f;
    }
  }

  // diagnostics: the central pressure deficit (hPa), the least d theta / dz in the core (K/km) between 2 km and
  // the vortex top, and the warm core's peak (K)
  double min_stab = 1e9, warm = 0.0;
  const int ncore = (int)ceil(2.0 * rmw / dr) < nr ? (int)ceil(2.0 * rmw / dr) : nr;
  for (int k = 1; k < nz; k++) {
    const double z = k * dz;
    if (z < 2000.0 || z > ztop) continue;
    for (int i = 0; i < ncore; i++) {
      const double lo = (double)B[(size_t)(k - 1) * 8] + tab[((size_t)(k - 1) * nr + i) * 4 + 1];
      const double hi = (double)B[(size_t)k * 8] + tab[((size_t)k * nr + i) * 4 + 1];
      const double s = (hi - lo) / dz * 1000.0;
      if (s < min_stab) min_stab = s;
    }
    const double w = tab[(size_t)k * nr * 4 + 1];
    if (w > warm) warm = w;
  }
  info[0] = dr;
  info[1] = nr;
  info[2] = (F[0] * F[0] * P2[0] + F[0] * P1[0]) * (double)B[2] / 100.0;
  info[3] = min_stab;
  info[4] = warm;
  return 0;
}
