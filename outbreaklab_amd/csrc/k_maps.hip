// 2-D maps of the current state for the display, storm finding and the scheduler (gfx950),
// restating the reference's map kernel (model.js:1028-1100) behind requestMaps (model.js:1478-1494):
//   map0 = (composite dBZ, dBZ at ~1 km, theta' at the lowest level against the base state, zeta at ~1 km)
//   map1 = (max w below 12 km, 2-5 km updraft helicity, lowest-level convergence 1e-3/s, lowest-level speed)
//   map2 = (u, v at the lowest level, u, v at ~1 km)
//   map3 = (surface pressure perturbation hPa, 10.7 um brightness temperature K, ~10-m wind, zeta at ~5 km)
// One work-item per column, x fastest (coalesced along the rows of every level it visits). V is read
// clamped to its faces as textureLoad(clamp(c, 0, n)); Q1 / Q2 at cells inside the grid.
#include <hip/hip_runtime.h>

#include "obl_device.h"
#include "obl_kernels.h"
#include "obl_refl.h"

namespace obl {

namespace {
struct MapCtx {
  const float4* V;
  int nx, ny, nz;
  __device__ __forceinline__ float4 ldV(int i, int j, int k) const {
    return ldx(V, off3(nx + 1, ny + 1, clampi(i, 0, nx), clampi(j, 0, ny), clampi(k, 0, nz)));
  }
  // cell-centred velocity from the faces
  __device__ __forceinline__ float3 cvel(int i, int j, int k) const {
    float4 p = ldV(i, j, k);
    return make_float3(0.5f * (p.x + ldV(i + 1, j, k).x), 0.5f * (p.y + ldV(i, j + 1, k).y), 0.5f * (p.z + ldV(i, j, k + 1).z));
  }
  // vertical vorticity from centred differences of the cell velocities
  __device__ __forceinline__ float zeta(int i, int j, int k, float dx) const {
    return (cvel(i + 1, j, k).y - cvel(i - 1, j, k).y - cvel(i, j + 1, k).x + cvel(i, j - 1, k).x) / (2.0f * dx);
  }
};
}  // namespace

__global__ void __launch_bounds__(256) k_maps(MU u, const float4* __restrict__ BS, const float4* __restrict__ V,
                                              const float4* __restrict__ Q1, const float4* __restrict__ Q2,
                                              const float* __restrict__ phi, const float4* __restrict__ sfc,
                                              float4* __restrict__ out) {
  Dims d = dims(u);
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= d.nx || j >= d.ny) return;
  const bool thm = u.v[U_MICRO2].z > 0.5f, mar = maritime_of(u);
  const MapCtx m{V, d.nx, d.ny, d.nz};
  const float dz = d.dz, dx = d.dx;
  auto dbz = [&](int k) {
    const uint32_t c = cidx(d, i, j, k);
    return dbz_of(thm, ldx(Q1, c), ldx(Q2, c), bsA(BS, k, d.nz).z, m.ldV(i, j, k).w * bsB(BS, k, d.nz).x, mar);
  };
  float comp = -30.0f, wmx = 0.0f, uh = 0.0f;
  const int k12 = min((int)(12000.0f / dz), d.nz - 1);
  for (int k = 0; k < k12; k++) {
    const float w = m.cvel(i, j, k).z;
    wmx = fmaxf(wmx, w);
    comp = fmaxf(comp, dbz(k));
    const float z = ((float)k + 0.5f) * dz;
    if (z > 2000.0f && z < 5000.0f && w > 0.0f) uh += w * fmaxf(m.zeta(i, j, k, dx), 0.0f) * dz;
  }
  const int k1 = clampi((int)(1000.0f / dz), 1, d.nz - 2);
  const float3 v0 = m.cvel(i, j, 0);
  const float4 s0 = m.ldV(i, j, 0);
  const float conv = -((m.ldV(i + 1, j, 0).x - s0.x + m.ldV(i, j + 1, 0).y - s0.y) / dx) * 1000.0f;
  const uint32_t n2 = __umul24((uint32_t)d.nx, (uint32_t)d.ny), idx = (uint32_t)i + __umul24((uint32_t)d.nx, (uint32_t)j);
  const float spd0 = sqrtf(v0.x * v0.x + v0.y * v0.y);
  stx(out, idx, make_float4(comp, dbz(k1), s0.w - bsA(BS, 0, d.nz).x, m.zeta(i, j, k1, dx)));
  stx(out, idx + n2, make_float4(wmx, uh, conv, spd0));
  const float3 v1 = m.cvel(i, j, k1);
  stx(out, idx + 2 * n2, make_float4(v0.x, v0.y, v1.x, v1.y));
  // brightness temperature: each level emits at its temperature in proportion to the condensate's
  // infrared absorption, seen through the cloud above it; the rest comes from the sea surface where there
  // is one (hurricane mode), else from the lowest level
  float tr = 1.0f, bt = 0.0f;
  for (int k = d.nz - 1; k >= 0; k--) {
    const uint32_t c = cidx(d, i, j, k);
    const float4 q1 = ldx(Q1, c), q2 = ldx(Q2, c);
    const float dtau = (80.0f * q1.y + 60.0f * q1.z + 25.0f * q2.x + 4.0f * q2.y + 1.0f * q1.w) * bsA(BS, k, d.nz).z * dz;
    const float a = 1.0f - expf(-dtau);
    bt += tr * a * m.ldV(i, j, k).w * bsB(BS, k, d.nz).x;
    tr *= 1.0f - a;
    if (tr < 1e-3f) break;
  }
  const float sst = sfc ? ldx(sfc, idx).y : 0.0f;
  bt += tr * (sst > 200.0f ? sst : s0.w * bsB(BS, 0, d.nz).x);
  const int p = (int)u.v[U_PAD].x, nxp = (int)u.v[U_PAD].y;
  const float ps = bsA(BS, 0, d.nz).z * u.v[U_CN].z * ldx(phi, (uint32_t)(i + p) + __umul24((uint32_t)nxp, (uint32_t)(j + p))) / 100.0f;
  const int k5 = clampi((int)(5000.0f / dz), 1, d.nz - 2);
  stx(out, idx + 3 * n2, make_float4(ps, bt, u.v[U_HUR2].w * spd0, m.zeta(i, j, k5, dx)));
}

void launch_maps(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, const float4* Q1,
                 const float4* Q2, const float* phi, const float4* sfc, float4* out) {
  dim3 blk(64, 4), grid((d.nx + 63) / 64, (d.ny + 3) / 4);
  hipLaunchKernelGGL(k_maps, grid, blk, 0, s, u, BS, V, Q1, Q2, phi, sfc, out);
}

}  // namespace obl
