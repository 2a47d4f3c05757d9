// Virtual radar cuts (k_radar.hip): the uniform block and buffers of one cut.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace obl {

// RADAR_FIELDS (radar.js:37-39), 11 vec4f, in the reference's order:
//   site = (model x, model y, antenna height MSL, ke R); geo = (lat0, lon0, site lat, site lon) radians;
//   cut = (elevation rad, az0 deg, d az deg, nrad); gates = (REF gates, Doppler gates, first gate m, gate m);
//   nyq = (Nyquist m/s, time s, noise seed, system PHIDP deg); offA = word offsets (REF, VEL, SW, ZDR);
//   offB = (PHI, RHO, -, moment mask); mdl = (nx, ny, nz, dx); mdl2 = (dz, ground MSL, domain x0, y0);
//   phys = (rho at the ground, echo floor dBZ, 0 Lin / 1 Thompson, tornadoes); phys2 = (maritime graupel, -, -, -)
struct RadarU {
  float4 site, geo, cut, gates, nyq, offA, offB, mdl, mdl2, phys, phys2;
};
constexpr int RADAR_U_FLOATS = 44;
constexpr int RADAR_MAX_TOR = 64;

struct RadarArgs {
  const float4* BS;
  const float4* V;   // the current state's copy
  const float4* Q1;
  const float4* Q2;
  const float2* aux;
  const float4* tor;  // per tornado (x, y, core radius, Vmax), (debris top, debris dBZ, debris 0-1, 0)
  uint32_t* out;      // the volume buffer: 8-bit codes, 4 gates per word
  float* kdp;         // scratch, nrad x Doppler gates
};
void launch_radar_cut(hipStream_t s, const RadarU& R, const RadarArgs& a);
// the geometry of listed gates, as the cut computes it: rg = n (radial, gate) pairs, out = n x 7 floats
// (x, y, ex, ey, zc, hl, hu)
void launch_radar_geom(hipStream_t s, const RadarU& R, int n, const int* rg, float* out);

}  // namespace obl
