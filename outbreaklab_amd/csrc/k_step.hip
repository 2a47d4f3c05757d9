// Model-step kernels for gfx950: large-scale forcing, CI map, initial state,
// eddy viscosity, semi-Lagrangian + MacCormack advection, fused forces and
// microphysics, divergence and projection.
//
// Each kernel keeps the arithmetic of the reference kernel it replaces
// (model.js, cited per kernel) and the reference's texel conventions; the
// data movement is this build's own: one work-item per point, x fastest so
// that every wave reads contiguous float4 texels, and per-channel gathers for
// the trilinear samples.
#include "../../include/outbreak_hip.h"
#include "obl_device.h"
#include "obl_kernels.h"
#include "obl_micro.h"
#include "obl_mx.h"
#include "obl_tile.h"

namespace obl {

// k_adv.hip compiles this file a second time with OBL_ADV_ONLY defined, keeping only the joint advection kernel and
// its launcher (built with an instruction-scheduling strategy of its own); this compilation leaves those two out.

// ====================== LSF (model.js:142-244) ======================
struct BndView {
  const float* b;  // 148 floats: p0..p4 then 32 points
};
__device__ static void bpt(const float* b, int i, float t, float& x, float& y) {
  const float* q = b + 20 + 4 * i;
  x = q[0] + q[2] * t;
  y = q[1] + q[3] * t;
}
__device__ static void sdist(const float* b, float t, float px, float py, float& sd, float& over_, float& along_) {
  int np = min((int)b[0], OBL_BOUND_POINTS);  // the record holds 32 points; the host rejects more
  float best = 1e30f, sg = 1.0f, over = 0.0f, along = 0.0f, s0 = 0.0f;
  for (int i = 0; i < np - 1; i++) {
    float ax, ay, ex, ey;
    bpt(b, i, t, ax, ay);
    bpt(b, i + 1, t, ex, ey);
    float abx = ex - ax, aby = ey - ay;
    float la = fmaxf(sqrtf(abx * abx + aby * aby), 1.0f);
    float t0 = ((px - ax) * abx + (py - ay) * aby) / (la * la);
    float tt = t0;
    if (i > 0) tt = fmaxf(tt, 0.0f);
    if (i < np - 2) tt = fminf(tt, 1.0f);
    float qx = px - (ax + tt * abx), qy = py - (ay + tt * aby);
    float d = sqrtf(qx * qx + qy * qy);
    if (d < best) {
      best = d;
      sg = (abx * (py - ay) - aby * (px - ax) > 0.0f) ? 1.0f : -1.0f;
      over = (i == 0 ? fmaxf(-t0, 0.0f) * la : 0.0f) + (i == np - 2 ? fmaxf(t0 - 1.0f, 0.0f) * la : 0.0f);
      along = s0 + tt * la;
    }
    s0 += la;
  }
  sd = best * sg * b[15];
  over_ = over;
  along_ = along;
}
__device__ static float alongVar(const float* b, float time, float s) {
  if (b[16] <= 0.0f) return 0.0f;
  return clampf_(fbm2(s / b[18], b[17] + time / 5400.0f, 2) * 1.6f, -1.0f, 1.0f);
}
__device__ static float hotSpots(const float* b, float time, float s) {
  float L = b[18];
  int sd = (int)(b[17] * 1000.0f);
  float ph = time / 4800.0f;
  int e = (int)floorf(ph);
  float fe = sstep(0.0f, 1.0f, ph - floorf(ph));
  int k0 = (int)floorf(s / L);
  float acc = 0.0f;
  for (int k = k0 - 1; k <= k0 + 1; k++) {
    for (int q = 0; q < 2; q++) {
      uint32_t h = hash3i(k, e + q, sd);
      float r1 = (float)(h & 0xffffu) / 65535.0f;
      float r2 = (float)((h >> 16u) & 0xffffu) / 65535.0f;
      float z = (s - ((float)k + 0.5f + 0.7f * (r1 - 0.5f)) * L) / (0.17f * L);
      acc += (q == 1 ? fe : 1.0f - fe) * (0.35f + 0.65f * r2) * ex(-z * z);
    }
  }
  return clampf_(acc, 0.0f, 1.0f);
}

#ifndef OBL_ADV_ONLY
__global__ void __launch_bounds__(256) k_lsf(MU u, const float4* __restrict__ BS, const float* __restrict__ BD,
                                             float4* __restrict__ lsfA, float4* __restrict__ lsfB,
                                             float* __restrict__ phT) {
  Dims d = dims(u);
  int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= d.nx || j >= d.ny) return;
  float time = u.v[U_STEP].y;
  float px = ((float)i + 0.5f) * d.dx, py = ((float)j + 0.5f) * d.dx;
  float sT = 0, sH = 0, sQ = 0, sHq = 0, sU = 0, sV = 0, sHw = 0, sw = 0, nud = 0;
  int nb = min((int)u.v[U_COR].z, OBL_MAX_BOUNDS);
  for (int bi = 0; bi < nb; bi++) {
    const float* b = BD + OBL_BOUND_FLOATS * bi;
    float dd, over, along;
    sdist(b, time, px, py, dd, over, along);
    float w = sstep(-b[3], b[3], dd) * (1.0f - sstep(0.0f, 60000.0f, over));
    if (w <= 0.0f) continue;
    float vr = b[16];
    float fd = 1.0f - sstep(30000.0f, 90000.0f, dd);
    float av = vr * alongVar(b, time, along) * fd;
    float lift = vr > 0.0f ? mixf(1.0f, mixf(1.0f - 0.85f * vr, 1.0f + 1.3f * vr, hotSpots(b, time, along)), fd) : 1.0f;
    float Hmin = b[7] * clampf_(lift, 0.15f, 3.0f);
    float H = clampf_(b[6] * fmaxf(dd + b[3], 0.0f) * (1.0f + 0.4f * av), Hmin, b[5]);
    sT += w * b[4] * ex(0.5f * av);
    sQ += w * b[8];
    sU += w * b[10] * lift;
    sV += w * b[11] * lift;
    sH += w * H;
    sHq += w * H * b[9];
    sHw += w * H * b[12];
    sw += w;
    nud = fmaxf(nud, b[13] * sstep(b[14], b[14] + 40000.0f, dd) * (1.0f - sstep(0.0f, 60000.0f, over)));
  }
  float iw = 1.0f / fmaxf(sw, 1e-6f);
  float4 A = make_float4(clampf_(sT, -20.0f, 10.0f), fmaxf(sH * iw, 100.0f), clampf_(sQ, 0.0f, 0.92f), fmaxf(sHq * iw, 100.0f));
  float4 B = make_float4(sU, sV, fmaxf(sHw * iw, 100.0f), nud);
  lsfA[(size_t)i + (size_t)d.nx * j] = A;
  lsfB[(size_t)i + (size_t)d.nx * j] = B;
  float ph = 0.0f, bprev = 0.0f;
  for (int k = d.nz - 1; k >= 0; k--) {
    float z = ((float)k + 0.5f) * d.dz;
    float4 a = bsA(BS, k, d.nz), bb = bsB(BS, k, d.nz);
    float dth = A.x / bb.x * tprof(z, A.y);
    float dqv = -a.y * A.z * tprof(z, A.w);
    float bt = 9.81f * (dth / a.x + 0.608f * dqv);
    if (k == d.nz - 1) ph = -bt * 0.5f * d.dz;
    else ph -= 0.5f * (bt + bprev) * d.dz;
    bprev = bt;
    phT[cidx(d, i, j, k)] = ph;
  }
}
#endif

// ====================== CI map (model.js:249-281) ======================
#ifndef OBL_ADV_ONLY
__global__ void __launch_bounds__(256) k_ci(MU u, const float* __restrict__ CI, float4* __restrict__ ciMap) {
  Dims d = dims(u);
  int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= d.nx || j >= d.ny) return;
  float t = u.v[U_STEP].y;
  float px = ((float)i + 0.5f) * d.dx, py = ((float)j + 0.5f) * d.dx;
  float4 best = make_float4(0.0f, 0.0f, 1500.0f, 1500.0f);
  int ne = min((int)u.v[U_CI].x, OBL_MAX_CI);
  for (int e = 0; e < ne; e++) {
    const float* E = CI + OBL_CI_FLOATS * e;
    if (t < E[6] || t > E[7]) continue;
    float dxv = px - (E[0] + E[8] * (t - E[6]));
    float dyv = py - (E[1] + E[9] * (t - E[6]));
    float ang = atan2f(dyv, dxv);
    float wob = 1.0f + 0.25f * gnoise2(cosf(ang) * 1.6f + E[10], sinf(ang) * 1.6f + 3.7f * E[10]);
    float g = sstep(E[6], E[6] + fmaxf(E[11], 60.0f), t);
    float r = sqrtf(dxv * dxv + dyv * dyv) / (E[3] * wob * (0.35f + 0.65f * g));
    if (r >= 1.0f) continue;
    float Lx = (float)d.nx * d.dx, Ly = (float)d.ny * d.dx;
    float edge = fminf(fminf(px, Lx - px), fminf(py, Ly - py));
    float h = sstep(E[6], E[6] + 60.0f, t) * (1.0f - sstep(E[7] - 180.0f, E[7], t)) *
              sstep(u.v[U_SPONGE].x, u.v[U_SPONGE].x + 10000.0f, edge);
    float wt = E[5] * (0.4f + 0.6f * g) * cosf(1.5707963f * r);
    if (h * wt > best.x * best.y) best = make_float4(h, wt, E[2], E[4]);
  }
  ciMap[(size_t)i + (size_t)d.nx * j] = best;
}
#endif

// ---------------- target state helpers (model.js:125-134) ----------------
__device__ __forceinline__ float tgtTh(const float4* BS, const Dims& d, float4 A, int k) {
  float z = ((float)k + 0.5f) * d.dz;
  return bsA(BS, k, d.nz).x + A.x / bsB(BS, k, d.nz).x * tprof(z, A.y);
}
__device__ __forceinline__ float tgtQv(const float4* BS, const Dims& d, float4 A, int k) {
  float z = ((float)k + 0.5f) * d.dz;
  return bsA(BS, k, d.nz).y * (1.0f - A.z * tprof(z, A.w));
}
// 2-D map reads for the target wind: a functor (i, j) -> clamped texel
struct G2d {
  const float4* m;
  Dims d;
  __device__ __forceinline__ float4 operator()(int i, int j) const { return ld2d(m, d, i, j); }
};
template <class B2>
__device__ __forceinline__ float tgtU(const float4* BS, const B2& lb, const Dims& d, int x, int y, int z) {
  float zz = ((float)z + 0.5f) * d.dz;
  float4 a = lb(x - 1, y), b = lb(x, y);
  return bsB(BS, z, d.nz).z + 0.5f * (a.x * tprof(zz, a.z) + b.x * tprof(zz, b.z));
}
template <class B2>
__device__ __forceinline__ float tgtV(const float4* BS, const B2& lb, const Dims& d, int x, int y, int z) {
  float zz = ((float)z + 0.5f) * d.dz;
  float4 a = lb(x, y - 1), b = lb(x, y);
  return bsB(BS, z, d.nz).w + 0.5f * (a.y * tprof(zz, a.z) + b.y * tprof(zz, b.z));
}
__device__ __forceinline__ float tgtU(const float4* BS, const float4* lsfB, const Dims& d, int x, int y, int z) {
  return tgtU(BS, G2d{lsfB, d}, d, x, y, z);
}
__device__ __forceinline__ float tgtV(const float4* BS, const float4* lsfB, const Dims& d, int x, int y, int z) {
  return tgtV(BS, G2d{lsfB, d}, d, x, y, z);
}

// ====================== INIT (model.js:284-339) ======================
// the vortex table's (tangential wind, theta') at radius r, level k (vxAt, model.js:292-297)
__device__ __forceinline__ float2 vxAt(const MU& u, const float4* VX, int nz, float r, int k) {
  int nr = (int)u.v[U_VTX].w;
  float fr = r / u.v[U_VTX].z;
  int i0 = (int)floorf(fr);
  if (i0 >= nr - 1) return make_float2(0.0f, 0.0f);
  int o = clampi(k, 0, nz - 1) * nr + i0;
  float t = fr - (float)i0;
  float4 a = VX[o], b = VX[o + 1];
  return make_float2(a.x * (1.0f - t) + b.x * t, a.y * (1.0f - t) + b.y * t);
}
#ifndef OBL_ADV_ONLY
__global__ void __launch_bounds__(256) k_init(MU u, const float4* __restrict__ BS, const float4* __restrict__ lsfA,
                                              const float4* __restrict__ lsfB, const float4* __restrict__ VX,
                                              float4* __restrict__ V, float4* __restrict__ Q1, float4* __restrict__ Q2,
                                              int seed) {
  Dims d = dims(u);
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x > d.nx || y > d.ny || z > d.nz) return;
  int k = min(z, d.nz - 1);
  float4 A = ld2d(lsfA, d, x, y);
  float zz = ((float)k + 0.5f) * d.dz;
  float nz0 = sstep(800.0f, 1500.0f, zz);
  // hurricane mode: the initial vortex (setVortex, model.js:1318-1365) and its noise reaching up through the core
  const bool vortex = u.v[U_VTX2].z > 0.5f && VX;
  const float xc = u.v[U_VTX].x, yc = u.v[U_VTX].y, dx = d.dx;
  const float pcx = ((float)x + 0.5f) * dx - xc, pcy = ((float)y + 0.5f) * dx - yc;
  if (vortex) nz0 = sqrtf(pcx * pcx + pcy * pcy) < u.v[U_VTX2].x ? sstep(6000.0f, 8000.0f, zz) : 1.0f;
  float noise = u.v[U_MISC].z * (rnd3(x + 17, y + 3 + 65536 * seed, z + 5) - 0.5f) * 2.0f * (1.0f - nz0);
  float th = tgtTh(BS, d, A, k) + noise;
  float uu = (y < d.ny && z < d.nz) ? tgtU(BS, lsfB, d, x, y, k) : 0.0f;
  float vv = (x < d.nx && z < d.nz) ? tgtV(BS, lsfB, d, x, y, k) : 0.0f;
  if (vortex) {
    const float pux = (float)x * dx - xc, puy = ((float)y + 0.5f) * dx - yc, ru = fmaxf(sqrtf(pux * pux + puy * puy), 1.0f);
    const float pvx = ((float)x + 0.5f) * dx - xc, pvy = (float)y * dx - yc, rv = fmaxf(sqrtf(pvx * pvx + pvy * pvy), 1.0f);
    uu -= vxAt(u, VX, d.nz, ru, k).x * puy / ru;
    vv += vxAt(u, VX, d.nz, rv, k).x * pvx / rv;
    th += vxAt(u, VX, d.nz, sqrtf(pcx * pcx + pcy * pcy), k).y;
  }
  float4 B = bsB(BS, k, d.nz);
  float T = th * B.x;
  float qv = fminf(tgtQv(BS, d, A, k), 0.95f * qvsat(T, B.y, fIce(T)));
  if (vortex) {  // a nearly saturated core of the same density
    const float f = (1.0f - sstep(0.6f, 1.0f, sqrtf(pcx * pcx + pcy * pcy) / u.v[U_VTX2].x)) * sstep(500.0f, 1500.0f, zz) *
                    (1.0f - sstep(7000.0f, 10000.0f, zz));
    const float q2 = fmaxf(qv, f * u.v[U_VTX2].y * qvsat(T, B.y, fIce(T)));
    const float dqv = q2 - qv;
    qv = q2;
    th -= 0.608f * dqv * bsA(BS, k, d.nz).x;
  }
  V[vidx(d, x, y, z)] = make_float4(uu, vv, 0.0f, th);
  if (x < d.nx && y < d.ny && z < d.nz) {
    Q1[cidx(d, x, y, z)] = make_float4(qv, 0.0f, 0.0f, 0.0f);
    Q2[cidx(d, x, y, z)] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}
#endif

// Build-defined warm bubble (no reference counterpart): theta += dth cos^2(pi r / 2)
// at each texel's cell centre, applied to both state copies.
#ifndef OBL_ADV_ONLY
__global__ void __launch_bounds__(256) k_bubble(MU u, float4* __restrict__ V0, float4* __restrict__ V1, double x0,
                                                double y0, double z0, double rh, double rz, double dth) {
  Dims d = dims(u);
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x > d.nx || y > d.ny || z > d.nz) return;
  double px = (x + 0.5) * (double)d.dx, py = (y + 0.5) * (double)d.dx, pz = (z + 0.5) * (double)d.dz;
  double a = (px - x0) / rh, b = (py - y0) / rh, c = (pz - z0) / rz;
  double r = sqrt(a * a + b * b + c * c);
  if (r < 1.0) {
    double cc = cos(1.5707963267948966 * r);
    float add = (float)(dth * cc * cc);
    size_t i = vidx(d, x, y, z);
    V0[i].w += add;
    V1[i].w += add;
  }
}
#endif

// ====================== NUT (model.js:342-389), outbreak mode ======================
template <class A>
__device__ __forceinline__ float nut_uc(const A& V, const float4* BS, int nz, int x, int y, int z) {
  return 0.5f * (V.ldc(x, y, z, 0) + V.ldc(x + 1, y, z, 0)) - bsB(BS, z, nz).z;
}
template <class A>
__device__ __forceinline__ float nut_vc(const A& V, const float4* BS, int nz, int x, int y, int z) {
  return 0.5f * (V.ldc(x, y, z, 1) + V.ldc(x, y + 1, z, 1)) - bsB(BS, z, nz).w;
}
template <class A>
__device__ __forceinline__ float nut_wc(const A& V, int x, int y, int z) {
  return 0.5f * (V.ldc(x, y, z, 2) + V.ldc(x, y, z + 1, 2));
}
template <class A>
__device__ __forceinline__ float2 nut_point(const MU& u, const Dims& d, const float4* BS, const A& V, float4 q1,
                                            int x, int y, int z) {
  float dx = d.dx, dz = d.dz;
  int nz = d.nz;
  float4 v0 = V.ld4(x, y, z);
  // (the divisions by dx, dz, 2 dx and 2 dz as products with their reciprocals, bit for bit)
  float dudx = rmul(V.ldc(x + 1, y, z, 0) - v0.x, d.rdx);
  float dvdy = rmul(V.ldc(x, y + 1, z, 1) - v0.y, d.rdx);
  float dwdz = rmul(V.ldc(x, y, z + 1, 2) - v0.z, d.rdz);
  float dudy = rmul(nut_uc(V, BS, nz, x, y + 1, z) - nut_uc(V, BS, nz, x, y - 1, z), d.r2dx);
  float dudz = rmul(nut_uc(V, BS, nz, x, y, z + 1) - nut_uc(V, BS, nz, x, y, z - 1), d.r2dz);
  float dvdx = rmul(nut_vc(V, BS, nz, x + 1, y, z) - nut_vc(V, BS, nz, x - 1, y, z), d.r2dx);
  float dvdz = rmul(nut_vc(V, BS, nz, x, y, z + 1) - nut_vc(V, BS, nz, x, y, z - 1), d.r2dz);
  float dwdx = rmul(nut_wc(V, x + 1, y, z) - nut_wc(V, x - 1, y, z), d.r2dx);
  float dwdy = rmul(nut_wc(V, x, y + 1, z) - nut_wc(V, x, y - 1, z), d.r2dx);
  float s12 = 0.5f * (dudy + dvdx), s13 = 0.5f * (dudz + dwdx), s23 = 0.5f * (dvdz + dwdy);
  float S2 = 2.0f * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + 4.0f * (s12 * s12 + s13 * s13 + s23 * s23);
  int kz0 = max(z - 1, 0), kz1 = min(z + 1, nz - 1);
  float N2 = rmul(9.81f / bsA(BS, z, nz).x * (V.ldc(x, y, kz1, 3) - V.ldc(x, y, kz0, 3)),
                       kz1 - kz0 == 2 ? d.r2dz : (kz1 > kz0 ? d.rdz : __builtin_inff()));
  if (u.v[U_COR].w > 0.5f && q1.y + q1.z > 1e-5f) N2 = fminf(N2, 0.0f);
  float Ri = N2 / fmaxf(S2, 1e-10f);
  float fst = sqrtf(fmaxf(0.0f, 1.0f - Ri / 0.33f));
  float sS = sqrtf(fmaxf(S2, 0.0f));
  float lh = u.v[U_TURB].x * dx, lv = 0.18f * dz;
  float nuh = fminf(lh * lh * sS * fst, 0.08f * dx * dx / fmaxf(d.dt, 0.1f));
  float nuv = fminf(lv * lv * sS * fst, 0.15f * dz * dz / fmaxf(d.dt, 0.1f));
  return make_float2(nuh, nuv);
}

// ====================== advection of V (model.js:392-481) ======================
struct Loc {
  float3 vu, vv, vw, vc;
};
template <class A>
__device__ __forceinline__ Loc localVel(const A& V, int x, int y, int z) {
  float4 p = V.ld4(x, y, z);
  float4 xm = V.ld4(x - 1, y, z), xp = V.ld4(x + 1, y, z);
  float4 ym = V.ld4(x, y - 1, z), yp = V.ld4(x, y + 1, z);
  float4 zm = V.ld4(x, y, z - 1), zp = V.ld4(x, y, z + 1);
  float xmyp = V.ldc(x - 1, y + 1, z, 1), xmzp = V.ldc(x - 1, y, z + 1, 2);
  float xpym = V.ldc(x + 1, y - 1, z, 0), ymzp = V.ldc(x, y - 1, z + 1, 2);
  float xpzm = V.ldc(x + 1, y, z - 1, 0), ypzm = V.ldc(x, y + 1, z - 1, 1);
  Loc L;
  L.vu = make_float3(p.x, 0.25f * (xm.y + p.y + xmyp + yp.y), 0.25f * (xm.z + p.z + xmzp + zp.z));
  L.vv = make_float3(0.25f * (ym.x + p.x + xpym + xp.x), p.y, 0.25f * (ym.z + p.z + ymzp + zp.z));
  L.vw = make_float3(0.25f * (zm.x + p.x + xpzm + xp.x), 0.25f * (zm.y + p.y + ypzm + yp.y), p.z);
  L.vc = make_float3(0.5f * (p.x + xp.x), 0.5f * (p.y + yp.y), 0.5f * (p.z + zp.z));
  return L;
}
// first-order departure (s = 1) / forward (s = -1) point, clamped to [0, N]
__device__ __forceinline__ float3 traceV(const Dims& d, float3 k, float3 p, float3 v0, float s) {
  return make_float3(clampt_(p.x - s * k.x * v0.x, 0.0f, (float)d.nx), clampt_(p.y - s * k.y * v0.y, 0.0f, (float)d.ny),
                     clampt_(p.z - s * k.z * v0.z, 0.0f, (float)d.nz));
}
// texU / texVv / texW / texC (model.js:411-414) in texel coordinates
template <class A>
__device__ __forceinline__ float sampU(const A& T, float3 p) { return samp1(T, p.x, p.y - 0.5f, p.z - 0.5f, 0); }
template <class A>
__device__ __forceinline__ float sampV(const A& T, float3 p) { return samp1(T, p.x - 0.5f, p.y, p.z - 0.5f, 1); }
template <class A>
__device__ __forceinline__ float sampW(const A& T, float3 p) { return samp1(T, p.x - 0.5f, p.y - 0.5f, p.z, 2); }
template <class A>
__device__ __forceinline__ float sampC(const A& T, float3 p) { return samp1(T, p.x - 0.5f, p.y - 0.5f, p.z - 0.5f, 3); }

// The 2 x 2 x 2 texel box of every departure / forward point of a face stays within one point of
// it while each Courant number of the local velocities is below one (traceV moves a point by
// k * v; clamping moves it inward). Then every sample falls in an LDS ring's window.
__device__ __forceinline__ bool near_traces(const Dims& d, const Loc& L) {
  const float kx = d.kx, kz = d.kz;
  float mx = fmaxf(fmaxf(fabsf(L.vu.x), fabsf(L.vv.x)), fmaxf(fabsf(L.vw.x), fabsf(L.vc.x)));
  float my = fmaxf(fmaxf(fabsf(L.vu.y), fabsf(L.vv.y)), fmaxf(fabsf(L.vw.y), fabsf(L.vc.y)));
  float mz = fmaxf(fmaxf(fabsf(L.vu.z), fabsf(L.vv.z)), fmaxf(fabsf(L.vw.z), fabsf(L.vc.z)));
  return mx * kx < 0.999f && my * kx < 0.999f && mz * kz < 0.999f;
}

template <class A>
__device__ __forceinline__ float2 mmV(const A& T, float3 p, float ox, float oy, float oz, int ch) {
  return minmax1(T, floori_(p.x - ox), floori_(p.y - oy), floori_(p.z - oz), ch);
}
// ====================== advection of moisture (model.js:482-576) ======================
template <class A>
__device__ __forceinline__ float3 cellVel(const A& V, int x, int y, int z) {
  float4 p = V.ld4(x, y, z);
  return make_float3(0.5f * (p.x + V.ldc(x + 1, y, z, 0)), 0.5f * (p.y + V.ldc(x, y + 1, z, 1)),
                     0.5f * (p.z + V.ldc(x, y, z + 1, 2)));
}
// the same in two parts: the face texels (issued as loads), then the averages
struct CellVelTexels {
  float4 p;
  float xp, yp, zp;
};
__device__ __forceinline__ CellVelTexels cellVelTexels(const GAcc& V, int x, int y, int z) {
  return CellVelTexels{V.ld4(x, y, z), V.ldc(x + 1, y, z, 0), V.ldc(x, y + 1, z, 1), V.ldc(x, y, z + 1, 2)};
}
__device__ __forceinline__ float3 cellVelOf(const CellVelTexels& t) {
  return make_float3(0.5f * (t.p.x + t.xp), 0.5f * (t.p.y + t.yp), 0.5f * (t.p.z + t.zp));
}
__device__ __forceinline__ float3 traceQ(const Dims& d, float3 k, float3 p, float3 v0, float fall, float s) {
  return make_float3(clampt_(p.x - s * k.x * v0.x, 0.5f, (float)d.nx - 0.5f),
                     clampt_(p.y - s * k.y * v0.y, 0.5f, (float)d.ny - 0.5f),
                     clampt_(p.z - s * k.z * (v0.z - fall), 0.5f, (float)d.nz - 0.5f));
}
// texS(x) = x / N: texel coordinate x - 1/2
template <class A>
__device__ __forceinline__ float4 sampQ(const A& T, float3 p) { return samp4(T, p.x - 0.5f, p.y - 0.5f, p.z - 0.5f); }
template <class A>
__device__ __forceinline__ float sampQ1(const A& T, float3 p, int ch) { return samp1(T, p.x - 0.5f, p.y - 0.5f, p.z - 0.5f, ch); }
template <class A>
__device__ __forceinline__ float2 mmQ1(const A& T, float3 p, int ch) {
  return minmax1(T, floori_(p.x - 0.5f), floori_(p.y - 0.5f), floori_(p.z - 0.5f), ch);
}

// Fall speeds of one cell (model.js:495-500): rain mass- and number-weighted, snow, graupel.
struct Fall {
  float r, n, s, g;  // n: number-weighted rain (Thompson), else 0
};
__device__ __forceinline__ Fall fall_speeds(bool thm, float4 q1, float4 q2, float rho, float rsfc) {
  float2 fv = vtRain(thm, q1.w, q2.z, rho, rsfc);
  return Fall{fv.x, thm ? fv.y : 0.0f, vtS(q2.x, rho, rsfc), vtGrau(thm, q2.y, rho, rsfc)};
}
// A cell whose hydrometeors all lie below the clamps of their fall-speed laws falls as a clear cell: the same
// function of the level's density alone. vtS, vtG and vtGT clamp 1000 rho q or rho q from below, lamR clamps q and
// Nr, vtR is 0 for q <= 0, and vtG's 1 + (q / 0.009)^2 rounds to exactly 1 there. The tests below are the laws'
// own clamped expressions, so such a cell's speeds equal the clear cell's bit for bit.
__device__ __forceinline__ bool fall_clear(bool thm, float4 q1, float4 q2, float rho) {
  const bool s = 1000.0f * rho * q2.x <= 1e-6f;
  if (thm) return s && q1.w <= 1e-12f && q2.z <= 1e-3f && rho * q2.y <= 1e-12f;
  return s && q1.w <= 0.0f && 1000.0f * rho * q2.y <= 1e-6f;
}
constexpr int FALL_TAB = 128;  // levels of a block's table of clear-cell fall speeds
// as near_traces, for cell-centred traces with per-species fall speeds: |k_z (w - f)| is largest
// at the extreme f, so the bound over f in [min(0, f...), max(0, f...)] covers every species
__device__ __forceinline__ bool near_traces_q(const Dims& d, float3 v, const Fall& f) {
  const float kx = d.kx, kz = d.kz;
  float fmin = fminf(fminf(0.0f, fminf(f.r, f.n)), fminf(f.s, f.g));
  float fmax = fmaxf(fmaxf(0.0f, fmaxf(f.r, f.n)), fmaxf(f.s, f.g));
  return fabsf(v.x) * kx < 0.999f && fabsf(v.y) * kx < 0.999f && fabsf(v.z - fmin) * kz < 0.999f &&
         fabsf(v.z - fmax) * kz < 0.999f;
}

// ====================== forces + microphysics (model.js:579-907) ======================
template <class A>
__device__ __forceinline__ float lapH(const A& V, int x, int y, int z, int ch, float me) {
  return V.ldc(x + 1, y, z, ch) + V.ldc(x - 1, y, z, ch) + V.ldc(x, y + 1, z, ch) + V.ldc(x, y - 1, z, ch) - 4.0f * me;
}
__device__ __forceinline__ float bsV(const float4* BS, const Dims& d, int k, int ch) {
  int kk = clampi(k, 0, d.nz - 1);
  if (ch == 0) return bsB(BS, kk, d.nz).z;
  if (ch == 1) return bsB(BS, kk, d.nz).w;
  if (ch == 3) return bsA(BS, kk, d.nz).x;
  return 0.0f;
}
template <class A>
__device__ __forceinline__ float lapZp(const A& V, const float4* BS, const Dims& d, int x, int y, int z, int ch, float me,
                                       bool lo, bool hi) {
  float m0 = me - bsV(BS, d, z, ch);
  float a = lo ? m0 : V.ldc(x, y, z - 1, ch) - bsV(BS, d, z - 1, ch);
  float b = hi ? m0 : V.ldc(x, y, z + 1, ch) - bsV(BS, d, z + 1, ch);
  return a + b - 2.0f * m0;
}
__device__ __forceinline__ float spongeLat(const MU& u, const Dims& d, float xm, float ym) {
  float Lx = (float)d.nx * d.dx, Ly = (float)d.ny * d.dx;
  float e = fminf(fminf(xm, Lx - xm), fminf(ym, Ly - ym));
  float dd = clampf_(1.0f - e / u.v[U_SPONGE].x, 0.0f, 1.0f);
  return dd * (0.3f + 0.7f * dd) / u.v[U_SPONGE].y;
}
__device__ __forceinline__ float spongeTop(const MU& u, const Dims& d, float z) {
  float ztop = (float)d.nz * d.dz;
  if (z < u.v[U_SPONGE].z) return 0.0f;
  float s = sinf(1.5707963f * (z - u.v[U_SPONGE].z) / (ztop - u.v[U_SPONGE].z));
  return s * s / u.v[U_SPONGE].w;
}
__device__ __forceinline__ float2 ldA(const float2* aux, const Dims& d, int x, int y, int z) {
  return ldx(aux, cidx(d, clampi(x, 0, d.nx - 1), clampi(y, 0, d.ny - 1), clampi(z, 0, d.nz - 1)));
}
__device__ __forceinline__ float ldP(const float* phT, const Dims& d, int x, int y, int z) {
  return ldx(phT, cidx(d, clampi(x, 0, d.nx - 1), clampi(y, 0, d.ny - 1), clampi(z, 0, d.nz - 1)));
}

// Accessors of phys_point's cell fields and 2-D map: global (clamped) reads, or reads of a block's
// LDS plane tile for the points of its own plane (the rest from global memory). The tile
// loader applies the same clamping, so both give the same values.
struct GAux {
  const float2* p;
  Dims d;
  __device__ __forceinline__ float2 operator()(int x, int y, int z) const { return ldA(p, d, x, y, z); }
};
struct GPh {
  const float* p;
  Dims d;
  __device__ __forceinline__ float operator()(int x, int y, int z) const { return ldP(p, d, x, y, z); }
};
struct TAux {
  const lds_float* s;  // [2][TPS]
  const float2* g;
  Dims d;
  int x0, y0, zc;
  __device__ __forceinline__ float2 operator()(int x, int y, int z) const {
    if (z != zc) return ldA(g, d, x, y, z);
    int i = (y - y0 + 1) * TPW + (x - x0 + 1);
    return make_float2(s[i], s[TPS + i]);
  }
};
struct TPh {
  const lds_float* s;  // [TPS]
  const float* g;
  Dims d;
  int x0, y0, zc;
  __device__ __forceinline__ float operator()(int x, int y, int z) const {
    if (z != zc) return ldP(g, d, x, y, z);
    return s[(y - y0 + 1) * TPW + (x - x0 + 1)];
  }
};
struct T2d {
  const lds_float* s;  // [4][TPS]
  int x0, y0;
  __device__ __forceinline__ float4 operator()(int i, int j) const {
    int b = (j - y0 + 1) * TPW + (i - x0 + 1);
    return make_float4(s[b], s[TPS + b], s[2 * TPS + b], s[3 * TPS + b]);
  }
};
// the three texels of one column at z - 1, z, z + 1 (for the vertical stencils)
struct ZCol {
  float4 m, c, p;
  int z;
  __device__ __forceinline__ float ldc(int, int, int k, int ch) const {
    float4 v = k < z ? m : (k > z ? p : c);
    return ch == 0 ? v.x : (ch == 1 ? v.y : (ch == 2 ? v.z : v.w));
  }
};

// The values of a point's column that the vertical stencils and the w-face buoyancy need from
// z - 1 / z + 1 (clamped texels), its Q2 texel and its 2-D map texels.
struct PhysCol {
  float4 Vm, Vp, Q1m, Q1p, Q2m, Q2c, LA, ci;
  float2 am;
  float pm;
};
// buoyancy of a cell (model.js:624-627) from its texels
__device__ __forceinline__ float buoyv(float4 A, float th, float4 q1, float4 q2) {
  return 9.81f * ((th - A.x) / A.x + 0.608f * (q1.x - A.y) - (q1.y + q1.z + q1.w + q2.x + q2.y));
}

enum { PHYS_PLAIN = 0, PHYS_DUMP = 1, PHYS_SURR = 2 };

// Hand a point to the surrogate: one atomic per wave reserves consecutive record slots for the
// wave's candidates (lane order). False when the list is full: the point then runs the plain block.
__device__ __forceinline__ bool surr_slot(bool cand, unsigned* count, unsigned cap, unsigned& slot) {
  unsigned long long m = __ballot(cand);
  if (!cand) return false;
  int leader = __ffsll((long long)m) - 1;
  unsigned below = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
  unsigned base = 0u;
  if (below == 0u) base = atomicAdd(count, (unsigned)__popcll(m));
  base = __shfl(base, leader);
  slot = base + below;
  return slot < cap;
}

// hurricane mode: the sea surface map's column (clamped), and the lowest-level wind speed at a cell
__device__ __forceinline__ float4 sfcAt(const HurArgs& hu, const Dims& d, int i, int j) {
  return ldx(hu.sfc, (uint32_t)clampi(i, 0, d.nx - 1) + __umul24((uint32_t)d.nx, (uint32_t)clampi(j, 0, d.ny - 1)));
}
__device__ __forceinline__ float u1At(const HurArgs& hu, const Dims& d, int x, int y) {
  const float4 s0 = ldx(hu.V, vidx(d, x, y, 0));
  const float ua = 0.5f * (s0.x + ldxc(hu.V, vidx(d, x + 1, y, 0), 0)), va = 0.5f * (s0.y + ldxc(hu.V, vidx(d, x, y + 1, 0), 1));
  return sqrtf(ua * ua + va * va);
}

template <int MODE, bool HUR, class AV, class AQ, class AX, class AP, class AB>
__device__ __forceinline__ void phys_point(const MU& u, const Dims& d, const float4* __restrict__ BS, const AV& Vin,
                                           const AQ& Q1in, const PhysCol& col, const AX& aux, const AB& lsfB,
                                           const AP& phT,
                                           float4* __restrict__ Vout, float4* __restrict__ Q1o, float4* __restrict__ Q2o,
                                           const MicroDump& dump, const SurrPhys& surr, const HurArgs& hu, int x, int y,
                                           int z) {
  int nx = d.nx, ny = d.ny, nz = d.nz;
  float dx = d.dx, dz = d.dz, dt = d.dt;
  float ix2 = 1.0f / (dx * dx), iz2 = 1.0f / (dz * dz);
  float fc = u.v[U_COR].x, nudR = 1.0f / u.v[U_COR].y;
  float nu0 = u.v[U_TURB].y;
  bool thm = u.v[U_MICRO2].z > 0.5f;
  bool drag = u.v[U_TURB].w > 0.5f;
  const uint32_t vi = vidx(d, x, y, z);
  float4 V = Vin.ld4(x, y, z), o = V;
  // the column's z neighbours (clamped texels)
  const ZCol zcol{col.Vm, V, col.Vp, z};
  // ---- u ----
  if (y < ny && z < nz) {
    float xmx = (float)x * dx, xmy = ((float)y + 0.5f) * dx, zz = ((float)z + 0.5f) * dz;
    float2 a0 = aux(x - 1, y, z), a1 = aux(x, y, z);
    float nux = 0.5f * (a0.x + a1.x) + nu0, nuy = 0.5f * (a0.y + a1.y) + nu0;
    float du = nux * lapH(Vin, x, y, z, 0, V.x) * ix2 + nuy * lapZp(zcol, BS, d, x, y, z, 0, V.x, z == 0, z == nz - 1) * iz2;
    float ut = tgtU(BS, lsfB, d, x, y, z);
    float vbar = 0.25f * (Vin.ldc(x - 1, y, z, 1) + V.y + Vin.ldc(x - 1, y + 1, z, 1) + Vin.ldc(x, y + 1, z, 1));
    float vtb = 0.5f * (tgtV(BS, lsfB, d, x - 1, y, z) + tgtV(BS, lsfB, d, x, y + 1, z));
    float nw = 0.5f * (lsfB(x - 1, y).w + lsfB(x, y).w);
    du += (spongeLat(u, d, xmx, xmy) + spongeTop(u, d, zz) + nw * nudR) * (ut - V.x);
    du += fc * (vbar - vtb);
    if (x > 0 && x < nx) du -= (phT(x, y, z) - phT(x - 1, y, z)) / dx;
    if (z == 0 && drag) {
      float lg = sqrtf(V.x * V.x + vbar * vbar), lg0 = sqrtf(ut * ut + vtb * vtb);
      float cd = u.v[U_MICRO].w;
      if (HUR) cd = cdSfc(u, lg, 0.5f * (sfcAt(hu, d, max(x - 1, 0), y).x + sfcAt(hu, d, x, y).x));
      du += cd * (-lg * V.x + lg0 * ut) / dz;
    }
    o.x = V.x + dt * du;
  } else {
    o.x = 0.0f;
  }
  // ---- v ----
  if (x < nx && z < nz) {
    float xmx = ((float)x + 0.5f) * dx, xmy = (float)y * dx, zz = ((float)z + 0.5f) * dz;
    float2 a0 = aux(x, y - 1, z), a1 = aux(x, y, z);
    float nux = 0.5f * (a0.x + a1.x) + nu0, nuy = 0.5f * (a0.y + a1.y) + nu0;
    float dv = nux * lapH(Vin, x, y, z, 1, V.y) * ix2 + nuy * lapZp(zcol, BS, d, x, y, z, 1, V.y, z == 0, z == nz - 1) * iz2;
    float vt = tgtV(BS, lsfB, d, x, y, z);
    float ubar = 0.25f * (Vin.ldc(x, y - 1, z, 0) + V.x + Vin.ldc(x + 1, y - 1, z, 0) + Vin.ldc(x + 1, y, z, 0));
    float utb = 0.5f * (tgtU(BS, lsfB, d, x, y - 1, z) + tgtU(BS, lsfB, d, x + 1, y, z));
    float nw = 0.5f * (lsfB(x, y - 1).w + lsfB(x, y).w);
    dv += (spongeLat(u, d, xmx, xmy) + spongeTop(u, d, zz) + nw * nudR) * (vt - V.y);
    dv -= fc * (ubar - utb);
    if (y > 0 && y < ny) dv -= (phT(x, y, z) - phT(x, y - 1, z)) / dx;
    if (z == 0 && drag) {
      float lg = sqrtf(ubar * ubar + V.y * V.y), lg0 = sqrtf(utb * utb + vt * vt);
      float cd = u.v[U_MICRO].w;
      if (HUR) cd = cdSfc(u, lg, 0.5f * (sfcAt(hu, d, x, max(y - 1, 0)).x + sfcAt(hu, d, x, y).x));
      dv += cd * (-lg * V.y + lg0 * vt) / dz;
    }
    o.y = V.y + dt * dv;
  } else {
    o.y = 0.0f;
  }
  // ---- w ----
  if (x < nx && y < ny && z > 0 && z < nz) {
    float xmx = ((float)x + 0.5f) * dx, xmy = ((float)y + 0.5f) * dx, zz = (float)z * dz;
    float2 a0 = col.am, a1 = aux(x, y, z);
    float nux = 0.5f * (a0.x + a1.x) + nu0, nuy = 0.5f * (a0.y + a1.y) + nu0;
    float lz = zcol.m.z + zcol.p.z - 2.0f * V.z;
    float dw = nux * lapH(Vin, x, y, z, 2, V.z) * ix2 + nuy * lz * iz2;
    dw += 0.5f * (buoyv(bsA(BS, z - 1, nz), col.Vm.w, col.Q1m, col.Q2m) +
                  buoyv(bsA(BS, z, nz), V.w, Q1in.ld4(x, y, z), col.Q2c)) -
          (phT(x, y, z) - col.pm) / dz;
    dw += (spongeLat(u, d, xmx, xmy) + spongeTop(u, d, zz)) * (0.0f - V.z);
    float4 ci = col.ci;
    if (ci.x > 0.0f && !HUR) {  // hurricane mode starts its storms as warm, moist bubbles instead (theta below)
      float zr = fabsf(zz - ci.z) / ci.w;
      if (zr < 1.0f) {
        float cv = cosf(1.5707963f * zr);
        dw += u.v[U_MISC].w * ci.x * cv * fmaxf(ci.y * cv - V.z, 0.0f);
      }
    }
    o.z = V.z + dt * dw;
  } else {
    o.z = 0.0f;
  }
  // ---- theta + microphysics ----
  if (x < nx && y < ny && z < nz) {
    float xmx = ((float)x + 0.5f) * dx, xmy = ((float)y + 0.5f) * dx, zz = ((float)z + 0.5f) * dz;
    float4 A = bsA(BS, z, nz), B = bsB(BS, z, nz);
    float4 LA = col.LA;
    float nw = lsfB(x, y).w;
    float tth = tgtTh(BS, d, LA, z), tqv = tgtQv(BS, d, LA, z);
    float sl = spongeLat(u, d, xmx, xmy);
    float relax = sl + nw * nudR;
    float2 a1 = aux(x, y, z);
    float nux = a1.x * 3.0f + nu0, nuy = a1.y * 3.0f + nu0;
    float th = V.w + dt * (nux * lapH(Vin, x, y, z, 3, V.w) * ix2 +
                           nuy * lapZp(zcol, BS, d, x, y, z, 3, V.w, z == 0, z == nz - 1) * iz2 +
                           (relax + spongeTop(u, d, zz)) * (tth - V.w));
    const uint32_t ci_ = cidx(d, x, y, z);
    float4 q1 = Q1in.ld4(x, y, z), q2 = col.Q2c;
    float4 e1 = Q1in.ld4(x + 1, y, z), w1 = Q1in.ld4(x - 1, y, z);
    float4 n1 = Q1in.ld4(x, y + 1, z), s1 = Q1in.ld4(x, y - 1, z);
    float4 dn = z == 0 ? q1 : col.Q1m;
    float4 up = z == nz - 1 ? q1 : col.Q1p;
    float l1x = e1.x + w1.x + n1.x + s1.x - 4.0f * q1.x, l1y = e1.y + w1.y + n1.y + s1.y - 4.0f * q1.y;
    float l1zz = e1.z + w1.z + n1.z + s1.z - 4.0f * q1.z;
    float lzx = dn.x + up.x - 2.0f * q1.x, lzy = dn.y + up.y - 2.0f * q1.y, lzz = dn.z + up.z - 2.0f * q1.z;
    float qvb = (z == 0 ? A.y : bsA(BS, z - 1, nz).y) + (z == nz - 1 ? A.y : bsA(BS, z + 1, nz).y) - 2.0f * A.y;
    lzx -= qvb;
    q1.x = q1.x + dt * (nux * l1x * ix2 + nuy * lzx * iz2);
    q1.y = q1.y + dt * (nux * l1y * ix2 + nuy * lzy * iz2);
    q1.z = q1.z + dt * (nux * l1zz * ix2 + nuy * lzz * iz2);
    q1.x += dt * relax * (tqv - q1.x);
    float envDecay = 0.0f;
    if (HUR) {
      // hurricane mode (model.js:713-774): radiative relaxation, storm-start bubbles, the far field held to the
      // sounding, boundary-layer thermals under the storm, the sea's heat and moisture
      th += dt * clampf_((tth - V.w) / u.v[U_HUR2].x, -u.v[U_HUR2].y, u.v[U_HUR2].y);
      const float4 cs = col.ci;
      if (cs.x > 0.0f) {
        const float zr = fabsf(zz - cs.z) / cs.w;
        if (zr < 1.0f) {
          const float f = cs.x * cosf(1.5707963f * zr) * cs.y / 8.0f;
          th += dt * 1.2e-3f * f;
          q1.x += dt * f / 600.0f * fmaxf(0.97f * qvsat(V.w * B.x, B.y, 0.0f) - q1.x, 0.0f);
        }
      }
      const float ex = xmx - u.v[U_HUR4].x, ey = xmy - u.v[U_HUR4].y;
      const float rS = sqrtf(ex * ex + ey * ey);
      const float wEnv = sstep(u.v[U_HUR4].w, 1.5f * u.v[U_HUR4].w, rS) / 3600.0f;
      th += dt * wEnv * (tth - V.w);
      q1.x += dt * wEnv * (tqv - q1.x);
      if (zz < u.v[U_HUR2].z) {
        const float u1 = u1At(hu, d, x, y);
        const float e = u.v[U_STEP].y / 900.0f, e0 = floorf(e), qx = xmx / 6000.0f, qy = xmy / 6000.0f;
        const float ga = gnoise2(qx + 37.1f * e0, qy + 11.3f * e0);
        const float gb = gnoise2(qx + 37.1f * (e0 + 1.0f), qy + 11.3f * (e0 + 1.0f));
        const float t = sstep(0.0f, 1.0f, e - floorf(e));
        const float rr = 0.5f * (ga * (1.0f - t) + gb * t);
        const float fu = fminf(u1 * u1 / 225.0f, 1.0f) * (1.0f - sstep(20.0f, 32.0f, u1)) *
                         (1.0f - sstep(u.v[U_HUR4].z, 1.5f * u.v[U_HUR4].z, rS));
        th += dt * u.v[U_HUR3].z * 2.0f * rr * fu * (1.0f - zz / u.v[U_HUR2].z);
      }
      if (z == 0) {
        const float4 sf = sfcAt(hu, d, x, y);
        const float land = clampf_(sf.x, 0.0f, 1.0f);
        const float u1 = u1At(hu, d, x, y);
        const float ch = u.v[U_HUR].y * u.v[U_HUR2].w * fmaxf(u1, u.v[U_HUR3].y) / dz;
        const int pd = (int)u.v[U_PAD].x;
        const float phc = ldx(hu.phi, (uint32_t)(x + pd) + __umul24((uint32_t)u.v[U_PAD].y, (uint32_t)(y + pd)));
        const float ps = fmaxf(u.v[U_HUR].w + A.z * u.v[U_CN].y * phc, 50000.0f);
        const float ths = sf.y * powf(1.0e5f / ps, 0.2854f);
        th += dt * ch * (1.0f - land) * (ths - V.w);
        if (u.v[U_HUR5].y > 0.5f) th += dt * cdSfc(u, u1, land) * u1 * u1 * u1 / (OBL_CPD * dz) / B.x;
        const float qsS = qvsat(sf.y, ps, 0.0f), qsL = qvsat(V.w * B.x, ps, 0.0f);
        q1.x += dt * ch * ((1.0f - land) * (qsS - q1.x) + land * u.v[U_HUR].z * fmaxf(qsL - q1.x, 0.0f));
      }
      envDecay = sstep(u.v[U_HUR4].w, 1.5f * u.v[U_HUR4].w, rS) / 1800.0f;
    }
    q1.x = fmaxf(q1.x, 0.0f); q1.y = fmaxf(q1.y, 0.0f); q1.z = fmaxf(q1.z, 0.0f); q1.w = fmaxf(q1.w, 0.0f);
    q2.x = fmaxf(q2.x, 0.0f); q2.y = fmaxf(q2.y, 0.0f); q2.z = fmaxf(q2.z, 0.0f); q2.w = fmaxf(q2.w, 0.0f);
    float T = th * B.x;
    float qv = q1.x, qc = q1.y, qi = q1.z, qr = q1.w, qs = q2.x, qg = q2.y, nr = fmaxf(q2.z, 0.0f);
    const float T0 = T, qv0 = qv, qc0 = qc, qi0 = qi, qr0 = qr, qs0 = qs, qg0 = qg, nr0 = nr;
    bool handed = false;
    if (MODE == PHYS_SURR) {
      // points with at least qmin of condensate, inside the training envelope, go to the surrogate (k_surr.hip)
      float wc = 0.5f * (V.z + zcol.p.z);
      float in[14] = {T, B.y, A.z, qv, qc, qr, qi, qs, qg, nr, 0.0f, wc, dz, dt};
      bool cand = micro_active(qc, qi, qr, qs, qg) && qc + qi + qr + qs + qg >= surr.qmin;
#pragma unroll
      for (int t = 0; t < 14; t++) cand = cand && in[t] >= surr.lo[t] && in[t] <= surr.hi[t];
      unsigned slot;
      if (surr_slot(cand, surr.count, surr.cap, slot)) {
        float keep = 1.0f / (1.0f + 2.0f * dt * sl + dt * 0.0f);
        float4* R = reinterpret_cast<float4*>(surr.rec + (size_t)slot * SURR_REC);
        R[0] = make_float4(__uint_as_float((unsigned)cidx(d, x, y, z)), T, B.y, A.z);
        R[1] = make_float4(qv, qc, qi, qr);
        R[2] = make_float4(qs, qg, nr, wc);
        R[3] = make_float4(keep, B.x, q2.w, 0.0f);
        handed = true;
      }
    }
    float Tp = 0.0f, qvp = 0.0f, qcp = 0.0f, qip = 0.0f;  // after the process part (teacher targets)
    if (!handed) {
      // Condensate-free air below saturation: every process rate of the block is zero (each is a product with a
      // zero mixing ratio, or sits behind a qr / qs / qg threshold), and each adjustment iteration's step
      // (qv - qs) / (...) <= 0 is clamped to -qcl = 0, so the block returns T, qv and the zero condensate
      // unchanged and only sets Thompson's Nr to 0 (model.js:856-866). The 0.1 % margin keeps the test clear of
      // the adjustment's own saturation value. Whole waves of clear air skip the block's arithmetic.
      const bool clear = qc == 0.0f && qi == 0.0f && qr == 0.0f && qs == 0.0f && qg == 0.0f &&
                         qv < 0.999f * qvsat(T, B.y, fIce(T));
      if (clear) {
        if (thm) nr = 0.0f;
        Tp = T; qvp = qv; qcp = qc; qip = qi;
      } else {
        micro_processes(u, dt, B.y, A.z, thm, T, qv, qc, qi, qr, qs, qg, nr);
        Tp = T; qvp = qv; qcp = qc; qip = qi;
        micro_satadj(B.y, T, qv, qc, qi);
      }
    }
    if (MODE == PHYS_DUMP) {
      // teacher pairs of the microphysics block: stratified sample, points where the block is
      // active (the surrogate's domain) with probability p_active, the rest with p_clear
      bool active = micro_active(qc0, qi0, qr0, qs0, qg0);
      float r = (float)(hash3i(x, y, z ^ (int)dump.salt) >> 8) * (1.0f / 16777216.0f);
      if (r < (active ? dump.p_active : dump.p_clear)) {
        unsigned slot = atomicAdd(dump.count, 1u);
        if (slot < dump.cap) {
          float wc = 0.5f * (V.z + zcol.p.z);
          float* row = dump.buf + (size_t)slot * OBL_DUMP_STRIDE;
          float in[14] = {T0, B.y, A.z, qv0, qc0, qr0, qi0, qs0, qg0, nr0, 0.0f, wc, dz, dt};
          // targets: the process part's tendencies (the saturation adjustment is re-applied exactly)
          float out[10] = {Tp - T0, qvp - qv0, qcp - qc0, qip - qi0, qr - qr0, qs - qs0, qg - qg0, nr - nr0, 0.0f,
                           (float)z};
#pragma unroll
          for (int t = 0; t < 14; t++) row[t] = in[t];
#pragma unroll
          for (int t = 0; t < 10; t++) row[14 + t] = out[t];
        }
      }
    }
    if (!handed) {
      th = T / B.x;
      float keep = 1.0f / (1.0f + 2.0f * dt * sl + dt * envDecay);
      qc *= keep; qi *= keep; qr *= keep; qs *= keep; qg *= keep; nr *= keep;
      stnt(Q1o, ci_, make_float4(fmaxf(qv, 0.0f), fmaxf(qc, 0.0f), fmaxf(qi, 0.0f), fmaxf(qr, 0.0f)));
      stnt(Q2o, ci_, make_float4(fmaxf(qs, 0.0f), fmaxf(qg, 0.0f), thm ? fmaxf(nr, 0.0f) : q2.z, q2.w));
    }
    o.w = th;  // a handed point's theta is written by the surrogate
  }
  float sp = sqrtf(o.x * o.x + o.y * o.y + o.z * o.z);
  float vcap = u.v[U_MISC].x;
  if (sp > vcap) {
    float f = vcap / sp;
    o.x *= f; o.y *= f; o.z *= f;
  }
  if (o.x != o.x || o.y != o.y || o.z != o.z || o.w != o.w) {
    float4 Bz = bsB(BS, z, nz);
    o = make_float4(Bz.z, Bz.w, 0.0f, bsA(BS, z, nz).x);
  }
  stnt(Vout, vi, o);
}


// ====================== z-marching drivers ======================
// One block per TX x TY column of points, walking z upward; every field a kernel gathers
// from is held as a three-plane LDS ring (obl_tile.h), the next plane prefetched into
// registers while the current one is computed.
#define TILE_SETUP_AT(BX, BY)                               \
  Dims d = dims(u);                                         \
  const int tid = threadIdx.x;                              \
  const int x0 = (BX) * TX, y0 = (BY) * TY;                 \
  const int x = x0 + tid % TX, y = y0 + tid / TX;
#define TILE_SETUP TILE_SETUP_AT(blockIdx.x, blockIdx.y)

template <bool HUR>
__global__ void __launch_bounds__(TILE_THREADS) k_nut_t(MU u, const float4* __restrict__ BS, const float4* __restrict__ V,
                                                        const float4* __restrict__ Q1, const float4* __restrict__ sfc,
                                                        float2* __restrict__ aux) {
  __shared__ float lds[RING_FLOATS];
  TILE_SETUP
  Ring rv{lds, V, d.nx + 1, d.ny + 1, d.nz + 1, x0, y0, 0};
  rv.load(-1, tid); rv.load(0, tid); rv.load(1, tid);
  __syncthreads();
  const bool act = x < d.nx && y < d.ny;
  // hurricane mode: the K-profile boundary layer's friction velocity, from the column's lowest-level wind
  // (model.js:377-386)
  float ust = 0.0f;
  if (HUR && act) {
    const float4 s0 = ldx(V, vidx(d, x, y, 0));
    const float ua = 0.5f * (s0.x + ldxc(V, vidx(d, x + 1, y, 0), 0)), va = 0.5f * (s0.y + ldxc(V, vidx(d, x, y + 1, 0), 1));
    const float u1 = sqrtf(ua * ua + va * va);
    ust = sqrtf(cdSfc(u, u1, ldx(sfc, (uint32_t)x + __umul24((uint32_t)d.nx, (uint32_t)y)).x)) * fmaxf(u1, u.v[U_HUR3].y);
  }
  // Reads a level needs besides its rings are loaded one level ahead, like the ring planes: the
  // loads a level issues are then only needed by later levels, and its compute starts without
  // waiting for memory.
  float4 q1n = ldQ(Q1, d, x, y, 0);
  for (int z = 0; z < d.nz; z++) {
    Ring::Pre pv;
    const float4 q1 = q1n;
    q1n = ldQ(Q1, d, x, y, z + 1);
    rv.fetch(z + 2, tid, pv);
    rv.center(z);
    if (act) {
      float2 nu = nut_point(u, d, BS, rv.w(), q1, x, y, z);
      if (HUR) {
        const float zk = ((float)z + 0.5f) * d.dz, hb = u.v[U_HUR2].z;
        if (zk < hb) {
          const float q = 1.0f - zk / hb;
          nu.y = fmaxf(nu.y, fminf(0.4f * ust * zk * q * q, 0.15f * d.dz * d.dz / fmaxf(d.dt, 0.1f)));
        }
      }
      stnt(aux, cidx(d, x, y, z), nu);
    }
    __syncthreads();
    rv.store(z + 2, tid, pv);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------------------
// Moisture advection in one pass: predictor (WGSL_M_ADVQ1, model.js:502-530) and corrector with its limiter
// (WGSL_M_ADVQ2, model.js:532-576) in one z-marching kernel, so the predicted fields never go through HBM.
//
// A block owns a 32 x 8 column tile. At level z it first predicts plane z + 1 for the tile and a one-point halo
// (the points whose predicted values the corrector's forward traces can reach) into an LDS ring of three
// predicted planes, from LDS rings of Q1 and Q2 with a two-point halo; then it corrects level z from the ring.
// The limiter's min / max over the 8 texels around each departure point are the predictor's own trilinear
// texels (same departure point, same fall speed), so they are taken while predicting and carried in registers
// to the next level, with the point's own texels, fall speeds and cell velocity.
//
// Waves whose traces all stay within one point (|Courant| < 1 on each axis for every species) read LDS only.
// Any other wave takes the same arithmetic on global memory: the predictor samples Q1 / Q2 there, and the
// corrector recomputes the predicted value of each texel it samples from Q1 / Q2 / V in global memory.
// ---------------------------------------------------------------------------------------------------
constexpr int QW = TX + 4, QH = TY + 4, QS = QW * QH;  // two-point halo
constexpr int QRING_FLOATS = 3 * 4 * QS;
constexpr int NHALO = TPS - TILE_THREADS;  // the one-point halo's slots (84)

// three planes of a (TX+4) x (TY+4) tile, channel-major, clamped texels (as Ring, two-point halo)
struct Ring2 {
  float* s;
  const float4* g;
  int W, H, D, x0, y0;
  __device__ __forceinline__ static int slot(int k) { return (k + 3 * 4096) % 3; }
  __device__ __forceinline__ uint32_t tofs(int t) const {
    return (uint32_t)clampi(x0 - 2 + t % QW, 0, W - 1) + __umul24((uint32_t)W, (uint32_t)clampi(y0 - 2 + t / QW, 0, H - 1));
  }
  __device__ __forceinline__ uint32_t pofs(int k) const { return __umul24((uint32_t)clampi(k, 0, D - 1), (uint32_t)(W * H)); }
  struct Pre {
    float4 v[(QS + TILE_THREADS - 1) / TILE_THREADS];
  };
  __device__ __forceinline__ void fetch(int k, int tid, Pre& r) const {
    const uint32_t po = pofs(k);
#pragma unroll
    for (int q = 0; q < (QS + TILE_THREADS - 1) / TILE_THREADS; q++) r.v[q] = ldx(g, tofs(min(tid + q * TILE_THREADS, QS - 1)) + po);
  }
  __device__ __forceinline__ void store(int k, int tid, const Pre& r) {
    float* p0 = s + slot(k) * 4 * QS;
#pragma unroll
    for (int q = 0; q < (QS + TILE_THREADS - 1) / TILE_THREADS; q++) {
      int t = tid + q * TILE_THREADS;
      if (t < QS) {
        p0[t] = r.v[q].x;
        p0[QS + t] = r.v[q].y;
        p0[2 * QS + t] = r.v[q].z;
        p0[3 * QS + t] = r.v[q].w;
      }
    }
  }
  __device__ __forceinline__ void load(int k, int tid) {
    Pre r;
    fetch(k, tid, r);
    store(k, tid, r);
  }
};
// window view of a Ring2 centred at plane kc (reads within one point of the halo-1 region and of plane kc)
struct Win2 {
  const lds_float* s;
  int x0, y0, kc, om, oc, op;
  __device__ __forceinline__ int base(int i0, int j0, int k0, int ch, int& dz) const {
    bool lo = k0 < kc;
    dz = lo ? oc - om : op - oc;
    return (lo ? om : oc) + ch * QS + __mul24(j0 - y0 + 2, QW) + (i0 - x0 + 2);
  }
  __device__ __forceinline__ float4 ld4(int i, int j) const {
    int b = oc + __mul24(j - y0 + 2, QW) + (i - x0 + 2);
    return make_float4(s[b], s[b + QS], s[b + 2 * QS], s[b + 3 * QS]);
  }
};
__device__ __forceinline__ Win2 win2(const Ring2& r, int kc) {
  return Win2{(const lds_float*)r.s, r.x0, r.y0, kc, Ring2::slot(kc - 1) * 4 * QS, Ring2::slot(kc) * 4 * QS,
              Ring2::slot(kc + 1) * 4 * QS};
}
// trilinear sample of one channel at texel coordinates (tx, ty, tz) and the min / max of its 8 texels
__device__ __forceinline__ float samp_mm(const Win2& a, float tx, float ty, float tz, int ch, float2& mm) {
  float ax = __builtin_amdgcn_fractf(tx), ay = __builtin_amdgcn_fractf(ty), az = __builtin_amdgcn_fractf(tz);
  int dz;
  int b = a.base(floori_(tx), floori_(ty), floori_(tz), ch, dz);
  const Box8 c = box8(a.s, b, dz, QW);
  mm = make_float2(vmin3_(vmin3_(c.c000, c.c100, c.c010), vmin3_(c.c110, c.c001, c.c101), vmin_(c.c011, c.c111)),
                   vmax3_(vmax3_(c.c000, c.c100, c.c010), vmax3_(c.c110, c.c001, c.c101), vmax_(c.c011, c.c111)));
  float c00 = lerpf(c.c000, c.c100, ax), c10 = lerpf(c.c010, c.c110, ax);
  float c01 = lerpf(c.c001, c.c101, ax), c11 = lerpf(c.c011, c.c111, ax);
  return lerpf(lerpf(c00, c10, ay), lerpf(c01, c11, ay), az);
}
// the same from global memory (clamped texels)
__device__ __forceinline__ float samp_mm(const GAcc& a, float tx, float ty, float tz, int ch, float2& mm) {
  int i0 = floori_(tx), j0 = floori_(ty), k0 = floori_(tz);
  mm = minmax1_global(a.g, a.W, a.H, a.D, i0, j0, k0, ch);
  return samp1_global(a.g, a.W, a.H, a.D, i0, j0, k0, __builtin_amdgcn_fractf(tx), __builtin_amdgcn_fractf(ty),
                      __builtin_amdgcn_fractf(tz), ch);
}

// The predictor of one cell (model.js:502-530): Q1t = (qv, qc, qi at the departure point, qr at the rain's),
// Q2t = (qs at the snow's, qg at the graupel's, Nr at the number-weighted rain's (Thompson) or the departure
// point, spare at the departure point); mm: the min / max of each channel's 8 texels (the corrector's limiter)
struct Pred {
  float4 h1, h2;
  float2 mm[8];
};
// Channels 0-3 are Q1's, 4-7 Q2's. `clear` (wave-uniform) says that channels 1-7 are +0 at every texel the wave's
// samples can reach: a trilinear sample of +0 texels is +0 and their min / max (+0, +0), so those channels are set to
// exactly that without their traces or their reads, and only vapour is sampled.
template <class A>
__device__ __forceinline__ void predict(const Dims& d, bool thm, const A& Q1, const A& Q2, float3 v, const Fall& f,
                                        int x, int y, int z, bool clear, Pred& o) {
  float3 k = make_float3(d.kx, d.kx, d.kz);
  float3 p = make_float3((float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f);
  float3 p0 = traceQ(d, k, p, v, 0.0f, 1.0f);
  float tx = p0.x - 0.5f, ty = p0.y - 0.5f, tz = p0.z - 0.5f;
  o.h1.x = samp_mm(Q1, tx, ty, tz, 0, o.mm[0]);
  if (clear) {
    o.h1 = make_float4(o.h1.x, 0.0f, 0.0f, 0.0f);
    o.h2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int ch = 1; ch < 8; ch++) o.mm[ch] = make_float2(0.0f, 0.0f);
    return;
  }
  float3 pr = traceQ(d, k, p, v, f.r, 1.0f);
  float3 ps = traceQ(d, k, p, v, f.s, 1.0f), pg = traceQ(d, k, p, v, f.g, 1.0f), pn = traceQ(d, k, p, v, f.n, 1.0f);
  o.h1.y = samp_mm(Q1, tx, ty, tz, 1, o.mm[1]);
  o.h1.z = samp_mm(Q1, tx, ty, tz, 2, o.mm[2]);
  o.h1.w = samp_mm(Q1, pr.x - 0.5f, pr.y - 0.5f, pr.z - 0.5f, 3, o.mm[3]);
  o.h2.x = samp_mm(Q2, ps.x - 0.5f, ps.y - 0.5f, ps.z - 0.5f, 0, o.mm[4]);
  o.h2.y = samp_mm(Q2, pg.x - 0.5f, pg.y - 0.5f, pg.z - 0.5f, 1, o.mm[5]);
  // f.n is 0 in the Lin scheme: Nr then comes from the departure point, as model.js:527
  o.h2.z = samp_mm(Q2, pn.x - 0.5f, pn.y - 0.5f, pn.z - 0.5f, 2, o.mm[6]);
  o.h2.w = samp_mm(Q2, tx, ty, tz, 3, o.mm[7]);
  (void)thm;
}
// channel bits (as `live`) of the texels of a float4 that are not +0
__device__ __forceinline__ uint32_t nz_bits(float4 v) {
  return (__float_as_uint(v.x) != 0u ? 1u : 0u) | (__float_as_uint(v.y) != 0u ? 2u : 0u) |
         (__float_as_uint(v.z) != 0u ? 4u : 0u) | (__float_as_uint(v.w) != 0u ? 8u : 0u);
}
// OR of an 8-bit mask over the wave (every lane active), wave-uniform
__device__ __forceinline__ uint32_t wave_or8(uint32_t m) {
  uint32_t r = 0;
  if (__ballot(m != 0u)) {
#pragma unroll
    for (int c = 0; c < 8; c++)
      if (__ballot((m >> c) & 1u)) r |= 1u << c;
  }
  return r;
}

// the predicted value of channel ch (0-3 of Q1t, 4-7 of Q2t) at texel (i, j, k) (clamped), from global memory
__device__ __noinline__ static float qhat_at(const MU& u, const Dims& d, const float4* BS, const float4* V,
                                             const float4* Q1, const float4* Q2, int i, int j, int k, int ch) {
  const bool thm = u.v[U_MICRO2].z > 0.5f;
  i = clampi(i, 0, d.nx - 1); j = clampi(j, 0, d.ny - 1); k = clampi(k, 0, d.nz - 1);
  GAcc gv{V, d.nx + 1, d.ny + 1, d.nz + 1, nullptr};
  GAcc g1{Q1, d.nx, d.ny, d.nz, nullptr}, g2{Q2, d.nx, d.ny, d.nz, nullptr};
  float3 v = cellVel(gv, i, j, k);
  float4 q1 = ldx(Q1, cidx(d, i, j, k)), q2 = ldx(Q2, cidx(d, i, j, k));
  Fall f = fall_speeds(thm, q1, q2, bsA(BS, k, d.nz).z, u.v[U_STEP].z);
  const float fall = ch == 3 ? f.r : ch == 4 ? f.s : ch == 5 ? f.g : ch == 6 ? f.n : 0.0f;
  float3 kk = make_float3(d.kx, d.kx, d.kz);
  float3 p = traceQ(d, kk, make_float3((float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f), v, fall, 1.0f);
  return ch < 4 ? samp1(g1, p.x - 0.5f, p.y - 0.5f, p.z - 0.5f, ch) : samp1(g2, p.x - 0.5f, p.y - 0.5f, p.z - 0.5f, ch - 4);
}
// trilinear sample of the predicted channel ch at texel coordinates (tx, ty, tz), every texel recomputed
__device__ __noinline__ static float qhat_sample_global(const MU& u, const Dims& d, const float4* BS, const float4* V,
                                                        const float4* Q1, const float4* Q2, float tx, float ty, float tz,
                                                        int ch) {
  int i0 = floori_(tx), j0 = floori_(ty), k0 = floori_(tz);
  float ax = __builtin_amdgcn_fractf(tx), ay = __builtin_amdgcn_fractf(ty), az = __builtin_amdgcn_fractf(tz);
  float c[8];
#pragma unroll
  for (int q = 0; q < 8; q++) c[q] = qhat_at(u, d, BS, V, Q1, Q2, i0 + (q & 1), j0 + ((q >> 1) & 1), k0 + (q >> 2), ch);
  float c00 = lerpf(c[0], c[1], ax), c10 = lerpf(c[2], c[3], ax), c01 = lerpf(c[4], c[5], ax), c11 = lerpf(c[6], c[7], ax);
  return lerpf(lerpf(c00, c10, ay), lerpf(c01, c11, ay), az);
}

// what the corrector of a cell needs from its predictor, carried to the next level in registers
struct Carry {
  float4 q1, q2;  // the cell's own texels
  float3 v;       // its cell velocity
  Fall f;
  bool near;
  uint32_t liveq;  // the channels not +0 somewhere in the Q window of the cell's plane (wave-uniform)
  float2 mm[8];
  // the face texels of the cell velocities of a later plane (own cell and halo slot), loaded a level ahead
  CellVelTexels pv1, pv2;
};

// the corrector of one cell (model.js:532-576) from the predicted ring (halo 1) or, for far traces, recomputed.
// On the near path `clear` (wave-uniform) says that channels 1-7 are +0 in the Q window of the plane and in the
// three predicted planes around it: the cell's texel, its predicted value, the forward sample and the limiter's
// bounds are then all +0, and so is the corrected value, which is set without computing it.
template <bool NEAR>
__device__ __forceinline__ void correct(const MU& u, const Dims& d, const RingW& w1t, const RingW& w2t, const Carry& c,
                                        const float4* BS, const float4* V, const float4* Q1, const float4* Q2, int x,
                                        int y, int z, bool clear, float4& o1, float4& o2) {
  float3 k = make_float3(d.kx, d.kx, d.kz);
  float3 p = make_float3((float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f);
  const float3 v = c.v;
  float cx = v.x * k.x, cy = v.y * k.y, cz = v.z * k.z;
  float cour = sqrtf(cx * cx + cy * cy + cz * cz);
  float sm = 0.5f * u.v[U_MISC].y * clampf_(cour / 0.35f, u.v[U_DZV].w, 1.0f);
  float3 f0 = traceQ(d, k, p, v, 0.0f, -1.0f);
  auto samp = [&](float3 at, int ch) {
    float tx = at.x - 0.5f, ty = at.y - 0.5f, tz = at.z - 0.5f;
    if (NEAR) return ch < 4 ? samp1(w1t, tx, ty, tz, ch) : samp1(w2t, tx, ty, tz, ch - 4);
    return qhat_sample_global(u, d, BS, V, Q1, Q2, tx, ty, tz, ch);
  };
  const float4 h1 = w1t.ld4(x, y, z), q1 = c.q1;
  if (NEAR && clear) {
    o1 = make_float4(vmax_(vclamp_(h1.x + sm * (q1.x - samp(f0, 0)), c.mm[0]), 0.0f), 0.0f, 0.0f, 0.0f);
    o2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  float3 fr = traceQ(d, k, p, v, c.f.r, -1.0f);
  float3 fs = traceQ(d, k, p, v, c.f.s, -1.0f), fg = traceQ(d, k, p, v, c.f.g, -1.0f), fn = traceQ(d, k, p, v, c.f.n, -1.0f);
  float b[8];
  const float3 at[8] = {f0, f0, f0, fr, fs, fg, fn, f0};
#pragma unroll
  for (int ch = 0; ch < 8; ch++) b[ch] = samp(at[ch], ch);
  const float4 h2 = w2t.ld4(x, y, z), q2 = c.q2;
  float r[8] = {h1.x + sm * (q1.x - b[0]), h1.y + sm * (q1.y - b[1]), h1.z + sm * (q1.z - b[2]), h1.w + sm * (q1.w - b[3]),
                h2.x + sm * (q2.x - b[4]), h2.y + sm * (q2.y - b[5]), h2.z + sm * (q2.z - b[6]), h2.w + sm * (q2.w - b[7])};
#pragma unroll
  for (int ch = 0; ch < 8; ch++) r[ch] = vmax_(vclamp_(r[ch], c.mm[ch]), 0.0f);
  o1 = make_float4(r[0], r[1], r[2], r[3]);
  o2 = make_float4(r[4], r[5], r[6], r[7]);
}

// the slot of the one-point halo ring that thread tid predicts in round 1 (its own cell) or round 2 (the halo)
__device__ __forceinline__ void halo_slot(int t, int& si, int& sj) {
  // t < 34: the row below the tile, t < 68: the row above, then the left and right columns
  if (t < TPW) { si = t; sj = 0; }
  else if (t < 2 * TPW) { si = t - TPW; sj = TPH - 1; }
  else if (t < 2 * TPW + TY) { si = 0; sj = 1 + t - 2 * TPW; }
  else { si = TPW - 1; sj = 1 + t - 2 * TPW - TY; }
}

// the moisture advection's LDS
struct AdvqSmem {
  float lds[2 * QRING_FLOATS + 2 * RING_FLOATS];
  float4 ftab[FALL_TAB];  // per level: a clear cell's fall speeds (r, n, s, g)
  // per ring slot and wave: the channels (Q1 bits 0-3, Q2 bits 4-7) not +0 somewhere in the texels the wave stored
  // into the Q rings' plane (tile and two-point halo) / the predicted rings' plane (tile and one-point halo)
  uint32_t nzq[3][4] __attribute__((aligned(16)));
  uint32_t nzh[3][4] __attribute__((aligned(16)));
};
// the moisture advection of column tile (bx, by)
__device__ __forceinline__ void advq_tile(const MU& u, const float4* __restrict__ BS, const float4* __restrict__ V,
                                          const float4* __restrict__ Q1, const float4* __restrict__ Q2,
                                          float4* __restrict__ Q1o, float4* __restrict__ Q2o, int bx, int by,
                                          AdvqSmem& sm) {
  float* const lds = sm.lds;
  float4* const ftab = sm.ftab;
  uint32_t(*const nzq)[4] = sm.nzq;
  uint32_t(*const nzh)[4] = sm.nzh;
  TILE_SETUP_AT(bx, by)
  const bool thm = u.v[U_MICRO2].z > 0.5f;
  const float rsfc = u.v[U_STEP].z;
  const int wave = tid >> 6;
  const bool lane0 = (tid & 63) == 0;
  {
    // computed at run time from zeros the compiler cannot see, so the table holds what the hardware's exp2 / log2
    // give a clear cell (read after the prologue's barrier)
    float zq = 0.0f;
    asm("" : "+v"(zq));
    const float4 z4 = make_float4(zq, zq, zq, zq);
    for (int k = tid; k < min(d.nz, FALL_TAB); k += TILE_THREADS) {
      const Fall f = fall_speeds(thm, z4, z4, bsA(BS, k, d.nz).z, rsfc);
      ftab[k] = make_float4(f.r, f.n, f.s, f.g);
    }
  }
  // a cell's fall speeds: the level's table entry when the whole wave is clear, else the laws
  auto fall_at = [&](float4 q1, float4 q2, float rho, int k) {
    if (k < FALL_TAB && wave_all(fall_clear(thm, q1, q2, rho))) {
      const float4 t = ftab[k];
      return Fall{t.x, t.y, t.z, t.w};
    }
    return fall_speeds(thm, q1, q2, rho, rsfc);
  };
  Ring2 q1r{lds, Q1, d.nx, d.ny, d.nz, x0, y0};
  Ring2 q2r{lds + QRING_FLOATS, Q2, d.nx, d.ny, d.nz, x0, y0};
  // the predicted planes (one-point halo): written by the predictor, read through RingW views
  Ring h1r{lds + 2 * QRING_FLOATS, nullptr, d.nx, d.ny, d.nz, x0, y0, 0};
  Ring h2r{lds + 2 * QRING_FLOATS + RING_FLOATS, nullptr, d.nx, d.ny, d.nz, x0, y0, 0};
  GAcc gv{V, d.nx + 1, d.ny + 1, d.nz + 1, nullptr};
  GAcc g1{Q1, d.nx, d.ny, d.nz, nullptr}, g2{Q2, d.nx, d.ny, d.nz, nullptr};
  const bool act = x < d.nx && y < d.ny;
  // round 1: this thread's own cell (clamped into the field); round 2 (threads below NHALO): a halo slot
  const int cx1 = clampi(x, 0, d.nx - 1), cy1 = clampi(y, 0, d.ny - 1);
  int hsi = 0, hsj = 0;
  halo_slot(min(tid, NHALO - 1), hsi, hsj);
  const int cx2 = clampi(x0 - 1 + hsi, 0, d.nx - 1), cy2 = clampi(y0 - 1 + hsj, 0, d.ny - 1);
  const int own = __mul24(tid / TX + 1, TPW) + tid % TX + 1, hal = __mul24(hsj, TPW) + hsi;

  // the Q rings' planes: stored with their channel masks
  auto store_q = [&](int k, const Ring2::Pre& a, const Ring2::Pre& b) {
    q1r.store(k, tid, a);
    q2r.store(k, tid, b);
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < (QS + TILE_THREADS - 1) / TILE_THREADS; q++) m |= nz_bits(a.v[q]) | (nz_bits(b.v[q]) << 4);
    m = wave_or8(m);
    if (lane0) nzq[Ring2::slot(k)][wave] = m;
  };
  // the channels not +0 somewhere in plane k of a ring (its slot's four wave words)
  auto nzw = [&](const uint32_t (*nz)[4], int k) {
    const uint4 w = *reinterpret_cast<const uint4*>(nz[Ring::slot(k)]);
    return w.x | w.y | w.z | w.w;
  };

  // predicts plane kp (0..nz-1) into the predicted ring, from the Q rings centred at kp and the cell-velocity
  // texels already in cr; returns this thread's own cell's carry
  auto predict_plane = [&](int kp, Carry& cr) {
    const Win2 w1 = win2(q1r, kp), w2 = win2(q2r, kp);
    const float rho = bsA(BS, kp, d.nz).z;
    float* H1 = h1r.s + Ring::slot(kp) * 4 * TPS;
    float* H2 = h2r.s + Ring::slot(kp) * 4 * TPS;
    // the channels not +0 somewhere in the Q window of plane kp (planes kp - 1 .. kp + 1, two-point halo)
    const uint32_t liveq = __builtin_amdgcn_readfirstlane(nzw(nzq, kp - 1) | nzw(nzq, kp) | nzw(nzq, kp + 1));
    // clear air: vapour the only channel (qv) not +0 in the window
    const bool clear = (liveq & ~0x01u) == 0u;
    Pred pr;
    uint32_t mh;
    {
      cr.q1 = w1.ld4(cx1, cy1);
      cr.q2 = w2.ld4(cx1, cy1);
      cr.v = cellVelOf(cr.pv1);
      cr.f = fall_at(cr.q1, cr.q2, rho, kp);
      cr.near = near_traces_q(d, cr.v, cr.f);
      cr.liveq = liveq;
      if (wave_all(cr.near)) predict(d, thm, w1, w2, cr.v, cr.f, cx1, cy1, kp, clear, pr);
      else predict(d, thm, g1, g2, cr.v, cr.f, cx1, cy1, kp, false, pr);
      H1[own] = pr.h1.x; H1[TPS + own] = pr.h1.y; H1[2 * TPS + own] = pr.h1.z; H1[3 * TPS + own] = pr.h1.w;
      H2[own] = pr.h2.x; H2[TPS + own] = pr.h2.y; H2[2 * TPS + own] = pr.h2.z; H2[3 * TPS + own] = pr.h2.w;
      mh = nz_bits(pr.h1) | (nz_bits(pr.h2) << 4);
#pragma unroll
      for (int q = 0; q < 8; q++) cr.mm[q] = pr.mm[q];
    }
    if (tid < NHALO) {
      const float4 q1 = w1.ld4(cx2, cy2), q2 = w2.ld4(cx2, cy2);
      const float3 v = cellVelOf(cr.pv2);
      const Fall f = fall_at(q1, q2, rho, kp);
      if (wave_all(near_traces_q(d, v, f))) predict(d, thm, w1, w2, v, f, cx2, cy2, kp, clear, pr);
      else predict(d, thm, g1, g2, v, f, cx2, cy2, kp, false, pr);
      H1[hal] = pr.h1.x; H1[TPS + hal] = pr.h1.y; H1[2 * TPS + hal] = pr.h1.z; H1[3 * TPS + hal] = pr.h1.w;
      H2[hal] = pr.h2.x; H2[TPS + hal] = pr.h2.y; H2[2 * TPS + hal] = pr.h2.z; H2[3 * TPS + hal] = pr.h2.w;
      mh |= nz_bits(pr.h1) | (nz_bits(pr.h2) << 4);
    }
    mh = wave_or8(mh);
    if (lane0) nzh[Ring::slot(kp)][wave] = mh;
  };

  // prologue: Q planes -1, 0, 1 (clamped), then plane 0's prediction
  {
    Ring2::Pre a, b;
#pragma unroll
    for (int k = -1; k <= 1; k++) {
      q1r.fetch(k, tid, a); q2r.fetch(k, tid, b);
      store_q(k, a, b);
    }
  }
  Carry cur, alt;
  auto load_vel = [&](int k, Carry& cr) {
    cr.pv1 = cellVelTexels(gv, cx1, cy1, k);
    if (tid < NHALO) cr.pv2 = cellVelTexels(gv, cx2, cy2, k);
  };
  load_vel(0, cur);
  load_vel(1, alt);
  __syncthreads();
  predict_plane(0, cur);
  __syncthreads();
  {
    // plane 2 replaces plane -1 (slot(2) == slot(-1)): the prediction of plane 1 needs planes 0..2
    Ring2::Pre a, b;
    q1r.fetch(2, tid, a); q2r.fetch(2, tid, b);
    store_q(2, a, b);
  }
  __syncthreads();
  // one level: corrects plane z from cur and predicts plane z + 1 into nxt
  auto level = [&](int z, Carry& cur, Carry& nxt) {
    Ring2::Pre p1, p2;
    const bool more = z + 1 < d.nz;
    // plane z + 3 for the prediction of plane z + 2 at the next level, and that prediction's cell velocities
    // (into cur, which the next level refills)
    q1r.fetch(z + 3, tid, p1); q2r.fetch(z + 3, tid, p2);
    load_vel(z + 2, cur);
    if (more) predict_plane(z + 1, nxt);
    __syncthreads();
    store_q(z + 3, p1, p2);
    // the predicted ring centred at z: below the ground and above the lid the clamped planes 0 and nz - 1
    h1r.center(z); h2r.center(z);
    if (z == 0) { h1r.om = h1r.oc; h2r.om = h2r.oc; }
    if (!more) { h1r.op = h1r.oc; h2r.op = h2r.oc; }
    // the channels not +0 somewhere in the cell's Q window or in the predicted planes its corrector reads
    const uint32_t live = __builtin_amdgcn_readfirstlane(
        cur.liveq | nzw(nzh, max(z - 1, 0)) | nzw(nzh, z) | nzw(nzh, min(z + 1, d.nz - 1)));
    if (act) {
      float4 o1, o2;
      if (wave_all(cur.near))
        correct<true>(u, d, h1r.w(), h2r.w(), cur, BS, V, Q1, Q2, x, y, z, (live & ~0x01u) == 0u, o1, o2);
      else correct<false>(u, d, h1r.w(), h2r.w(), cur, BS, V, Q1, Q2, x, y, z, false, o1, o2);
      const uint32_t i = cidx(d, x, y, z);
      stx(Q1o, i, o1);
      stx(Q2o, i, o2);
    }
    __syncthreads();
  };
  // two levels per trip, the carries swapping roles (no copy of a carry between levels)
  for (int z = 0; z < d.nz; z += 2) {
    level(z, cur, alt);
    if (z + 1 < d.nz) level(z + 1, alt, cur);
  }
}

// ---------------------------------------------------------------------------------------------------
// Wind advection in one pass, with the turbulent viscosity (outbreak mode): predictor (WGSL_M_ADVV1,
// model.js:432-449), corrector with its limiter (WGSL_M_ADVV2, model.js:451-481) and WGSL_M_NUT
// (model.js:342-389) in one z-marching kernel over the faces grid, as the moisture advection below: plane z + 1
// is predicted for the tile and a one-point halo into an LDS ring, from an LDS ring of V with a two-point halo
// that also holds every texel nut's stencil reads; the limiter's min / max come from the predictor's texels.
// ---------------------------------------------------------------------------------------------------
// a 3-plane view of a Ring2 for the accessor-templated point code (nut_point, localVel): reads within two
// points of the tile and one plane of kc
struct Win2A {
  const lds_float* s;
  int x0, y0, kc, om, oc, op;
  __device__ __forceinline__ int pl(int k) const { return k < kc ? om : (k == kc ? oc : op); }
  __device__ __forceinline__ int sidx(int i, int j, int k, int ch) const {
    return pl(k) + ch * QS + __mul24(j - y0 + 2, QW) + (i - x0 + 2);
  }
  __device__ __forceinline__ float ldc(int i, int j, int k, int ch) const { return s[sidx(i, j, k, ch)]; }
  __device__ __forceinline__ float4 ld4(int i, int j, int k) const {
    int b = sidx(i, j, k, 0);
    return make_float4(s[b], s[b + QS], s[b + 2 * QS], s[b + 3 * QS]);
  }
};
__device__ __forceinline__ Win2A win2a(const Ring2& r, int kc) {
  return Win2A{(const lds_float*)r.s, r.x0, r.y0, kc, Ring2::slot(kc - 1) * 4 * QS, Ring2::slot(kc) * 4 * QS,
               Ring2::slot(kc + 1) * 4 * QS};
}

// the predictor of one face (model.js:432-449): each channel sampled at its own departure point, with the
// min / max of the 8 texels there (the corrector's limiter)
struct PredV {
  float4 h;
  float2 mm[4];
};
template <class A>
__device__ __forceinline__ void predict_v(const Dims& d, const A& V, const Loc& L, int x, int y, int z, PredV& o) {
  float3 k = make_float3(d.kx, d.kx, d.kz);
  float fx = (float)x, fy = (float)y, fz = (float)z;
  float3 pu = traceV(d, k, make_float3(fx, fy + 0.5f, fz + 0.5f), L.vu, 1.0f);
  float3 pv = traceV(d, k, make_float3(fx + 0.5f, fy, fz + 0.5f), L.vv, 1.0f);
  float3 pw = traceV(d, k, make_float3(fx + 0.5f, fy + 0.5f, fz), L.vw, 1.0f);
  float3 pc = traceV(d, k, make_float3(fx + 0.5f, fy + 0.5f, fz + 0.5f), L.vc, 1.0f);
  o.h.x = samp_mm(V, pu.x, pu.y - 0.5f, pu.z - 0.5f, 0, o.mm[0]);
  o.h.y = samp_mm(V, pv.x - 0.5f, pv.y, pv.z - 0.5f, 1, o.mm[1]);
  o.h.z = samp_mm(V, pw.x - 0.5f, pw.y - 0.5f, pw.z, 2, o.mm[2]);
  o.h.w = samp_mm(V, pc.x - 0.5f, pc.y - 0.5f, pc.z - 0.5f, 3, o.mm[3]);
}
// samp_mm on a 3-plane view (the predictor's texel boxes lie within one plane of kc)
__device__ __forceinline__ float samp_mm(const Win2A& a, float tx, float ty, float tz, int ch, float2& mm) {
  const Win2 w{a.s, a.x0, a.y0, a.kc, a.om, a.oc, a.op};
  return samp_mm(w, tx, ty, tz, ch, mm);
}
// the predicted channel ch of face texel (i, j, k) (clamped), from global memory
__device__ __noinline__ static float vhat_at(const Dims& d, const float4* V, int i, int j, int k, int ch) {
  i = clampi(i, 0, d.nx); j = clampi(j, 0, d.ny); k = clampi(k, 0, d.nz);
  GAcc gv{V, d.nx + 1, d.ny + 1, d.nz + 1, nullptr};
  const Loc L = localVel(gv, i, j, k);
  float3 kk = make_float3(d.kx, d.kx, d.kz);
  float fx = (float)i, fy = (float)j, fz = (float)k;
  if (ch == 0) return sampU(gv, traceV(d, kk, make_float3(fx, fy + 0.5f, fz + 0.5f), L.vu, 1.0f));
  if (ch == 1) return sampV(gv, traceV(d, kk, make_float3(fx + 0.5f, fy, fz + 0.5f), L.vv, 1.0f));
  if (ch == 2) return sampW(gv, traceV(d, kk, make_float3(fx + 0.5f, fy + 0.5f, fz), L.vw, 1.0f));
  return sampC(gv, traceV(d, kk, make_float3(fx + 0.5f, fy + 0.5f, fz + 0.5f), L.vc, 1.0f));
}
__device__ __noinline__ static float vhat_sample_global(const Dims& d, const float4* V, float tx, float ty, float tz,
                                                        int ch) {
  int i0 = floori_(tx), j0 = floori_(ty), k0 = floori_(tz);
  float ax = __builtin_amdgcn_fractf(tx), ay = __builtin_amdgcn_fractf(ty), az = __builtin_amdgcn_fractf(tz);
  float c[8];
#pragma unroll
  for (int q = 0; q < 8; q++) c[q] = vhat_at(d, V, i0 + (q & 1), j0 + ((q >> 1) & 1), k0 + (q >> 2), ch);
  float c00 = lerpf(c[0], c[1], ax), c10 = lerpf(c[2], c[3], ax), c01 = lerpf(c[4], c[5], ax), c11 = lerpf(c[6], c[7], ax);
  return lerpf(lerpf(c00, c10, ay), lerpf(c01, c11, ay), az);
}
struct CarryV {
  Loc L;
  float4 old;
  bool near;
  float2 mm[4];
};
// the corrector of one face (model.js:451-481)
template <bool NEAR>
__device__ __forceinline__ float4 correct_v(const MU& u, const Dims& d, const RingW& wt, const CarryV& c,
                                            const float4* V, int x, int y, int z) {
  float3 k = make_float3(d.kx, d.kx, d.kz);
  float fx = (float)x, fy = (float)y, fz = (float)z;
  const Loc& L = c.L;
  float3 pu = traceV(d, k, make_float3(fx, fy + 0.5f, fz + 0.5f), L.vu, -1.0f);
  float3 pv = traceV(d, k, make_float3(fx + 0.5f, fy, fz + 0.5f), L.vv, -1.0f);
  float3 pw = traceV(d, k, make_float3(fx + 0.5f, fy + 0.5f, fz), L.vw, -1.0f);
  float3 pc = traceV(d, k, make_float3(fx + 0.5f, fy + 0.5f, fz + 0.5f), L.vc, -1.0f);
  float4 b;
  if (NEAR) {
    b = make_float4(samp1(wt, pu.x, pu.y - 0.5f, pu.z - 0.5f, 0), samp1(wt, pv.x - 0.5f, pv.y, pv.z - 0.5f, 1),
                    samp1(wt, pw.x - 0.5f, pw.y - 0.5f, pw.z, 2), samp1(wt, pc.x - 0.5f, pc.y - 0.5f, pc.z - 0.5f, 3));
  } else {
    b = make_float4(vhat_sample_global(d, V, pu.x, pu.y - 0.5f, pu.z - 0.5f, 0),
                    vhat_sample_global(d, V, pv.x - 0.5f, pv.y, pv.z - 0.5f, 1),
                    vhat_sample_global(d, V, pw.x - 0.5f, pw.y - 0.5f, pw.z, 2),
                    vhat_sample_global(d, V, pc.x - 0.5f, pc.y - 0.5f, pc.z - 0.5f, 3));
  }
  const float4 hat = wt.ld4(x, y, z), old = c.old;
  float cx = L.vc.x * k.x, cy = L.vc.y * k.y, cz = L.vc.z * k.z;
  float cour = sqrtf(cx * cx + cy * cy + cz * cz);
  float sm = 0.5f * u.v[U_MISC].y * clampf_(cour / 0.35f, u.v[U_DZV].w, 1.0f);
  return make_float4(vclamp_(hat.x + sm * (old.x - b.x), c.mm[0]), vclamp_(hat.y + sm * (old.y - b.y), c.mm[1]),
                     vclamp_(hat.z + sm * (old.z - b.z), c.mm[2]), vclamp_(hat.w + sm * (old.w - b.w), c.mm[3]));
}

// (the wind advection's point code needs ~210 VGPRs: two blocks per CU; ADVV_MIN_BLOCKS = 3 makes the compiler fit
// three, with spills)
#ifndef ADVV_MIN_BLOCKS
#define ADVV_MIN_BLOCKS 2
#endif
// the wind advection's LDS
struct AdvvSmem {
  float lds[QRING_FLOATS + RING_FLOATS];
};
// nut and the wind advection of face tile (bx, by)
__device__ __forceinline__ void advv_tile(const MU& u, const float4* __restrict__ BS, const float4* __restrict__ V,
                                          const float4* __restrict__ Q1, float4* __restrict__ Vout,
                                          float2* __restrict__ aux, int bx, int by, AdvvSmem& sm) {
  float* const lds = sm.lds;
  TILE_SETUP_AT(bx, by)
  const int W = d.nx + 1, H = d.ny + 1, D = d.nz + 1;
  Ring2 vr{lds, V, W, H, D, x0, y0};
  Ring hr{lds + QRING_FLOATS, nullptr, W, H, D, x0, y0, 0};
  GAcc gv{V, W, H, D, nullptr};
  const bool act = x <= d.nx && y <= d.ny, cell = x < d.nx && y < d.ny;
  const int cx1 = clampi(x, 0, d.nx), cy1 = clampi(y, 0, d.ny);
  int hsi = 0, hsj = 0;
  halo_slot(min(tid, NHALO - 1), hsi, hsj);
  const int cx2 = clampi(x0 - 1 + hsi, 0, d.nx), cy2 = clampi(y0 - 1 + hsj, 0, d.ny);
  const int own = __mul24(tid / TX + 1, TPW) + tid % TX + 1, hal = __mul24(hsj, TPW) + hsi;

  // predicts face plane kp into the predicted ring, and the viscosity of the cells of level kp (kp < nz)
  auto predict_plane = [&](int kp, float4 q1, CarryV& cr) {
    const Win2A w = win2a(vr, kp);
    float* Hp = hr.s + Ring::slot(kp) * 4 * TPS;
    PredV pr;
    cr.L = localVel(w, cx1, cy1, kp);
    cr.old = w.ld4(cx1, cy1, kp);
    cr.near = near_traces(d, cr.L);
    if (wave_all(cr.near)) predict_v(d, w, cr.L, cx1, cy1, kp, pr);
    else predict_v(d, gv, cr.L, cx1, cy1, kp, pr);
    Hp[own] = pr.h.x; Hp[TPS + own] = pr.h.y; Hp[2 * TPS + own] = pr.h.z; Hp[3 * TPS + own] = pr.h.w;
#pragma unroll
    for (int q = 0; q < 4; q++) cr.mm[q] = pr.mm[q];
    if (cell && kp < d.nz) stx(aux, cidx(d, x, y, kp), nut_point(u, d, BS, w, q1, x, y, kp));
    if (tid < NHALO) {
      const Loc L = localVel(w, cx2, cy2, kp);
      if (wave_all(near_traces(d, L))) predict_v(d, w, L, cx2, cy2, kp, pr);
      else predict_v(d, gv, L, cx2, cy2, kp, pr);
      Hp[hal] = pr.h.x; Hp[TPS + hal] = pr.h.y; Hp[2 * TPS + hal] = pr.h.z; Hp[3 * TPS + hal] = pr.h.w;
    }
  };

  vr.load(-1, tid); vr.load(0, tid); vr.load(1, tid);
  // the cell's Q1 texel for nut, loaded one level ahead (ldQ clamps into the field; lanes past the grid's cells
  // load a clamped texel they do not use: unconditional loads, no select of addresses)
  float4 q1n = ldQ(Q1, d, x, y, 0);
  __syncthreads();
  CarryV cur;
  {
    const float4 q1 = q1n;
    q1n = ldQ(Q1, d, x, y, 1);
    predict_plane(0, q1, cur);
  }
  __syncthreads();
  {
    Ring2::Pre a;
    vr.fetch(2, tid, a);
    vr.store(2, tid, a);
  }
  __syncthreads();
  // one face level: corrects plane z from cur and predicts plane z + 1 into nxt, with the cells' Q1 of level
  // z + 1 (q1); returns level z + 2's, loaded meanwhile
  auto level = [&](int z, const CarryV& cur, CarryV& nxt, float4 q1) {
    Ring2::Pre pv;
    const bool more = z < d.nz;
    vr.fetch(z + 3, tid, pv);
    const float4 q1next = ldQ(Q1, d, x, y, z + 2);
    if (more) predict_plane(z + 1, q1, nxt);
    __syncthreads();
    vr.store(z + 3, tid, pv);
    hr.center(z);
    if (z == 0) hr.om = hr.oc;
    if (!more) hr.op = hr.oc;
    if (act) {
      float4 o;
      if (wave_all(cur.near)) o = correct_v<true>(u, d, hr.w(), cur, V, x, y, z);
      else o = correct_v<false>(u, d, hr.w(), cur, V, x, y, z);
      stx(Vout, vidx(d, x, y, z), o);
    }
    __syncthreads();
    return q1next;
  };
  // two levels per trip, the carries swapping roles (no carry copies between levels)
  CarryV alt;
  for (int z = 0; z <= d.nz; z += 2) {
    const float4 q1b = level(z, cur, alt, q1n);
    if (z + 1 <= d.nz) q1n = level(z + 1, alt, cur, q1b);
  }
}

#ifdef OBL_ADV_ONLY
// Both advections in one launch (they read the same state and write different fields), so that a CU runs a wind
// tile and a moisture tile side by side and the moisture advection's waits at its barriers and on memory overlap
// the wind advection's instruction issue (the two kernels back to back leave each bottleneck alone). Block b runs
// on XCD b mod 8 and an XCD's 32 CUs take its blocks in turn, so roles alternate in runs of 256 blocks (one block
// per CU of every XCD); measured against runs of 1 (roles split between XCDs) and of 8. One LDS union: a block
// holds one role's rings.
union AdvSmem {
  AdvqSmem q;
  AdvvSmem v;
};
__global__ void __launch_bounds__(TILE_THREADS, 2) k_adv_t(MU u, const float4* __restrict__ BS, const float4* __restrict__ V,
                                                           const float4* __restrict__ Q1, const float4* __restrict__ Q2,
                                                           float4* __restrict__ Vout, float2* __restrict__ aux,
                                                           float4* __restrict__ Q1o, float4* __restrict__ Q2o, int gvx,
                                                           int nv, int gqx, int nq) {
  __shared__ AdvSmem sm;
  const int b = blockIdx.x, m = min(nv, nq);
  bool wind;
  int i;
  constexpr int S = 8;
  const int M = (2 * m) & ~((2 << S) - 1);
  if (b < M) {
    wind = ((b >> S) & 1) == 0;
    i = ((b >> (S + 1)) << S) | (b & ((1 << S) - 1));
  } else if (b < 2 * m) {
    wind = (b & 1) == 0;
    i = M / 2 + ((b - M) >> 1);
  } else {
    wind = nv > nq;
    i = m + (b - 2 * m);
  }
  if (wind) advv_tile(u, BS, V, Q1, Vout, aux, i % gvx, i / gvx, sm.v);
  else advq_tile(u, BS, V, Q1, Q2, Q1o, Q2o, i % gqx, i / gqx, sm.q);
}
#endif  // OBL_ADV_ONLY

// launch grids of the tile kernels
static dim3 tface_grid(const Dims& d) { return dim3((d.nx + 1 + TX - 1) / TX, (d.ny + 1 + TY - 1) / TY, 1); }
static dim3 tcell_grid(const Dims& d) { return dim3((d.nx + TX - 1) / TX, (d.ny + TY - 1) / TY, 1); }

#ifndef OBL_ADV_ONLY
// One block per 32 x 8 column tile of the faces grid, walking z upward like the advection
// kernels: V and Q1 are three-plane LDS rings (the next plane prefetched into registers during a
// level's compute), aux and phT the current plane, lsfB the 2-D tile; the vertical stencils read
// z - 1 / z + 1 from the rings, and Q2 / aux / phT at z - 1 are carried from the level below. Every
// field crosses HBM about once per step instead of once per neighbouring level.
template <int MODE, bool HUR>
__global__ void __launch_bounds__(TILE_THREADS) k_phys(MU u, const float4* __restrict__ BS, const float4* __restrict__ Vin,
                                                       const float4* __restrict__ Q1in, const float4* __restrict__ Q2in,
                                                       const float2* __restrict__ aux, const float4* __restrict__ lsfA,
                                                       const float4* __restrict__ lsfB, const float* __restrict__ phT,
                                                       const float4* __restrict__ ciMap, float4* __restrict__ Vout,
                                                       float4* __restrict__ Q1o, float4* __restrict__ Q2o, MicroDump dump,
                                                       SurrPhys sp, HurArgs hu) {
  __shared__ float lds[2 * RING_FLOATS];
  __shared__ float sA[2 * TPS], sP[TPS], sB[4 * TPS];
  TILE_SETUP
  const bool act = x <= d.nx && y <= d.ny;
  Ring rv{lds, Vin, d.nx + 1, d.ny + 1, d.nz + 1, x0, y0, 0};
  Ring rq{lds + RING_FLOATS, Q1in, d.nx, d.ny, d.nz, x0, y0, 0};
  for (int k = -1; k <= 1; k++) { rv.load(k, tid); rq.load(k, tid); }
  // aux / phT tile texels of this thread (clamped like ldA / ldP), and the 2-D lsfB tile
  constexpr int NQ = (TPS + TILE_THREADS - 1) / TILE_THREADS;
  uint32_t cof[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    int t = tid + q * TILE_THREADS;
    int i = clampi(x0 - 1 + t % TPW, 0, d.nx - 1), j = clampi(y0 - 1 + t / TPW, 0, d.ny - 1);
    cof[q] = (uint32_t)i + __umul24((uint32_t)d.nx, (uint32_t)j);
    if (t < TPS) {
      float4 b = ldx(lsfB, cof[q]);
      sB[t] = b.x; sB[TPS + t] = b.y; sB[2 * TPS + t] = b.z; sB[3 * TPS + t] = b.w;
      float2 a = ldx(aux, cof[q]);
      sA[t] = a.x; sA[TPS + t] = a.y;
      sP[t] = ldx(phT, cof[q]);
    }
  }
  const uint32_t plane = __umul24((uint32_t)d.nx, (uint32_t)d.ny);
  PhysCol col;
  col.LA = ld2d(lsfA, d, x, y);
  col.ci = ldx(ciMap, (uint32_t)clampi(x, 0, d.nx - 1) + __umul24((uint32_t)d.nx, (uint32_t)clampi(y, 0, d.ny - 1)));
  col.Q2m = ldQ(Q2in, d, x, y, 0);
  col.am = ldA(aux, d, x, y, 0);
  col.pm = ldP(phT, d, x, y, 0);
  // the cell's Q2 texel, loaded one level ahead
  float4 q2n = ldQ(Q2in, d, x, y, 0);
  __syncthreads();
  const T2d tb{(const lds_float*)sB, x0, y0};
  for (int z = 0; z <= d.nz; z++) {
    Ring::Pre pv, pq;
    float2 na[NQ];
    float np[NQ];
    col.Q2c = q2n;
    q2n = ldQ(Q2in, d, x, y, z + 1);
    {
      rv.fetch(z + 2, tid, pv);
      rq.fetch(z + 2, tid, pq);
      const uint32_t zo = __umul24(plane, (uint32_t)clampi(z + 1, 0, d.nz - 1));
      // unconditional (cof is a clamped, valid offset for every lane)
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        na[q] = ldx(aux, cof[q] + zo);
        np[q] = ldx(phT, cof[q] + zo);
      }
    }
    rv.center(z);
    rq.center(z);
    const RingW wv = rv.w(), wq = rq.w();
    if (act) {
      col.Vm = wv.ld4(x, y, z - 1);
      col.Vp = wv.ld4(x, y, z + 1);
      col.Q1m = wq.ld4(x, y, z - 1);
      col.Q1p = wq.ld4(x, y, z + 1);
      const TAux ta{(const lds_float*)sA, aux, d, x0, y0, z};
      const TPh tp{(const lds_float*)sP, phT, d, x0, y0, z};
      phys_point<MODE, HUR>(u, d, BS, wv, wq, col, ta, tb, tp, Vout, Q1o, Q2o, dump, sp, hu, x, y, z);
      col.Q2m = col.Q2c;
      col.am = ta(x, y, z);
      col.pm = tp(x, y, z);
    }
    __syncthreads();
    {
      rv.store(z + 2, tid, pv);
      rq.store(z + 2, tid, pq);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        int t = tid + q * TILE_THREADS;
        if (t < TPS) {
          sA[t] = na[q].x; sA[TPS + t] = na[q].y;
          sP[t] = np[q];
        }
      }
    }
    __syncthreads();
  }
}

// ====================== DIV (model.js:910-925) ======================
// Each block takes ZL = 2 consecutive levels of a 64 x 4 column tile: level z's top faces (w at z + 1) are the next
// level's own texels, read once for both (one block per level read each plane twice; 4 or 8 levels per block lose the
// row reuse that neighbouring blocks of one plane share in L2: 0.43 -> 0.38 ms with 2, 0.39 with 4, 0.43 with 8).
constexpr int ZL = 2;
// The divergence right-hand side of cell (x, y, z) from its face texel p, the texel above (its top w), u of the
// face at x + 1 and v of the face at y + 1 (model.js:910-925)
__device__ __forceinline__ float div_cell(const Dims& d, const float4* __restrict__ BS, float4 p, float4 pup, float ux,
                                          float vy, int z) {
  const float rc = bsA(BS, z, d.nz).z;
  const float wt = z == d.nz - 1 ? 0.0f : pup.z;
  const float wb = z == 0 ? 0.0f : p.z;
  const float dd = rmul(rc * (ux - p.x + vy - p.y), d.rdx) + rmul(rhoW(BS, z + 1, d.nz) * wt - rhoW(BS, z, d.nz) * wb, d.rdz);
  return rmul(dd, d.rdt);
}
__global__ void __launch_bounds__(256) k_div(MU u, const float4* __restrict__ BS, const float4* __restrict__ V,
                                             float* __restrict__ rhs) {
  Dims d = dims(u);
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= d.nx || y >= d.ny) return;
  const int z0 = blockIdx.z * ZL;
  int pd = (int)u.v[U_PAD].x, nxp = (int)u.v[U_PAD].y, nyp = (int)u.v[U_PAD].z;
  float4 pv[ZL + 1];
  float ux[ZL], vy[ZL];
#pragma unroll
  for (int j = 0; j <= ZL; j++) pv[j] = ldx(V, vidx(d, x, y, min(z0 + j, d.nz)));
#pragma unroll
  for (int j = 0; j < ZL; j++) {
    ux[j] = ldVc(V, d, x + 1, y, z0 + j, 0);
    vy[j] = ldVc(V, d, x, y + 1, z0 + j, 1);
  }
#pragma unroll
  for (int j = 0; j < ZL; j++) {
    const int z = z0 + j;
    if (z < d.nz) stx(rhs, off3(nxp, nyp, x + pd, y + pd, z), div_cell(d, BS, pv[j], pv[j + 1], ux[j], vy[j], z));
  }
}

// DIV and the fold-and-split of the multi-piece BF16 solve's x forward pass in one kernel. A block takes two levels
// of four rows y = 4 g .. 4 g + 3 of the solve's padded grid (Nx x Ny, the interior at pd) over a 64-wide span of
// the fold index x' (64 i .. 64 i + 63, x' < Nx / 2): threads 0-255 the cells x = x', threads 256-511 their mirrors
// x = Nx - 1 - x' (32- and 128-wide spans measured slower). Each computes its cell's divergence as k_div does
// (stored to rhs; zero outside the interior) into LDS, and two waves then write the x forward pass's tiled operand
// for the span as the fold kernel does from rhs:
// parity 0 holds S = F(x') + F(Nx - 1 - x'), parity 1 D = F(x') - F(Nx - 1 - x'), in P BF16 pieces. Row y of level
// z is operand row t 32 MF + 16 MF h + (q mod 16 MF) of pair q = z Ny / 2 + y' (y' = y in the lower half, h = 0;
// Ny - 1 - y in the upper, h = 1; t = q / 16 MF), so with Ny a multiple of 8 the four rows are four consecutive
// operand rows, and 8 lanes (4 rows x 2 k-halves) fill one 128-byte line. The divergence skips its round trip
// through rhs for the fold; the loads are k_div's (two levels per block, the level above read once for both).
template <int P>
__global__ void __launch_bounds__(512) k_div_fold(MU u, const float4* __restrict__ BS, const float4* __restrict__ V,
                                                  float* __restrict__ rhs, uint16_t* __restrict__ out, int RB, int KB) {
  __shared__ __align__(16) float fs[2][2][4][64];  // [cell / mirror][level][row][x' - 64 i]
  const Dims d = dims(u);
  const int pd = (int)u.v[U_PAD].x, Nx = (int)u.v[U_PAD].y, Ny = (int)u.v[U_PAD].z;
  const int Hx = Nx >> 1, Hy = Ny >> 1;
  const int tid = threadIdx.x, half = tid >> 8, r = (tid >> 6) & 3, xi = tid & 63;
  const int z0 = 2 * blockIdx.z, xq = 64 * blockIdx.x + xi;
  const int xp = half ? Nx - 1 - xq : xq, x = xp - pd, yp = 4 * blockIdx.y + r, y = yp - pd;
  float f0 = 0.0f, f1 = 0.0f;
  // (x' >= Nx / 2: the cell is another span's mirror)
  if (xq < Hx && x >= 0 && x < d.nx && y >= 0 && y < d.ny) {
    float4 pv[3];
    float ux[2], vy[2];
#pragma unroll
    for (int j = 0; j <= 2; j++) pv[j] = ldx(V, vidx(d, x, y, min(z0 + j, d.nz)));
#pragma unroll
    for (int j = 0; j < 2; j++) {
      ux[j] = ldVc(V, d, x + 1, y, z0 + j, 0);
      vy[j] = ldVc(V, d, x, y + 1, z0 + j, 1);
    }
    f0 = div_cell(d, BS, pv[0], pv[1], ux[0], vy[0], z0);
    stnt(rhs, off3(Nx, Ny, xp, yp, z0), f0);
    if (z0 + 1 < d.nz) {
      f1 = div_cell(d, BS, pv[1], pv[2], ux[1], vy[1], z0 + 1);
      stnt(rhs, off3(Nx, Ny, xp, yp, z0 + 1), f1);
    }
  }
  fs[half][0][r][xi] = f0;
  fs[half][1][r][xi] = f1;
  __syncthreads();
  if (tid >= 128) return;
  // lane: row (bits 0-1), k-half (bit 2), k-block of the span (bits 3-4), level (bit 5), parity (bit 6)
  const int lr = tid & 3, kc = (tid >> 2) & 1, kbl = (tid >> 3) & 3, j = (tid >> 5) & 1, par = tid >> 6;
  const int z = z0 + j, kb = 4 * blockIdx.x + kbl;
  if (z >= d.nz || kb >= KB) return;
  const int xo = 16 * kbl + 8 * kc;
  const float4 a0 = *reinterpret_cast<const float4*>(&fs[0][j][lr][xo]);
  const float4 a1 = *reinterpret_cast<const float4*>(&fs[0][j][lr][xo + 4]);
  const float4 m0 = *reinterpret_cast<const float4*>(&fs[1][j][lr][xo]);
  const float4 m1 = *reinterpret_cast<const float4*>(&fs[1][j][lr][xo + 4]);
  uint2 lo[P], hi[P];
  if (par == 0) {
    mx_split4<P>(a0.x + m0.x, a0.y + m0.y, a0.z + m0.z, a0.w + m0.w, lo);
    mx_split4<P>(a1.x + m1.x, a1.y + m1.y, a1.z + m1.z, a1.w + m1.w, hi);
  } else {
    mx_split4<P>(a0.x - m0.x, a0.y - m0.y, a0.z - m0.z, a0.w - m0.w, lo);
    mx_split4<P>(a1.x - m1.x, a1.y - m1.y, a1.z - m1.z, a1.w - m1.w, hi);
  }
  const int yq = 4 * blockIdx.y + lr, h = yq >= Hy, q = z * Hy + (h ? Ny - 1 - yq : yq);
  const int ro = q / (16 * MX_MF) * (32 * MX_MF) + 16 * MX_MF * h + q % (16 * MX_MF);
  uint16_t* const o = out + (size_t)(par * P) * ((size_t)RB * (KB + 1) * 512) + mx_off(ro, 16 * kb + 8 * kc, KB);
#pragma unroll
  for (int w = 0; w < P; w++)
    *reinterpret_cast<uint4*>(o + (size_t)w * ((size_t)RB * (KB + 1) * 512)) = make_uint4(lo[w].x, lo[w].y, hi[w].x, hi[w].y);
}

// ====================== PROJ (model.js:928-946) ======================
__device__ __forceinline__ float Pphi(const float* phi, const MU& u, const Dims& d, int x, int y, int z) {
  int p = (int)u.v[U_PAD].x, nxp = (int)u.v[U_PAD].y, nyp = (int)u.v[U_PAD].z;
  x = clampi(x, -p, d.nx + p - 1);
  y = clampi(y, -p, d.ny + p - 1);
  z = clampi(z, 0, d.nz - 1);
  return ldx(phi, off3(nxp, nyp, x + p, y + p, z));
}
__global__ void __launch_bounds__(256) k_proj(MU u, const float4* __restrict__ V, const float* __restrict__ phi,
                                              float4* __restrict__ Vout) {
  Dims d = dims(u);
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x > d.nx || y > d.ny || z > d.nz) return;
  const float kx = d.kx, kz = d.kz;
  const uint32_t i = vidx(d, x, y, z);
  float4 Vv = ldx(V, i), o = Vv;
  float pc = Pphi(phi, u, d, x, y, z);
  if (y < d.ny && z < d.nz) o.x = Vv.x - kx * (pc - Pphi(phi, u, d, x - 1, y, z));
  if (x < d.nx && z < d.nz) o.y = Vv.y - kx * (pc - Pphi(phi, u, d, x, y - 1, z));
  if (x < d.nx && y < d.ny && z > 0 && z < d.nz) o.z = Vv.z - kz * (pc - Pphi(phi, u, d, x, y, z - 1));
  else o.z = 0.0f;
  stx(Vout, i, o);
}

// ====================== hurricane mode: advection, PREP, FILT ======================
// Hurricane runs trace with the wind at the trajectory's midpoint (cn.w, model.js:420-430, 494-499) and carry the
// values of Vp (V less part of the last pressure force, WGSL_M_PREP) in the wind advection; ice settles (vtI) and
// Lin graupel is maritime. The advections walk z over LDS rings like the outbreak ones; PREP and FILT are face
// kernels over global memory (GIn, clamped texel reads).
struct GIn {
  const float4* g;
  int W, H, D;
  __device__ __forceinline__ uint32_t gidx(int i, int j, int k) const {
    return off3(W, H, clampi(i, 0, W - 1), clampi(j, 0, H - 1), clampi(k, 0, D - 1));
  }
  __device__ __forceinline__ float ldc(int i, int j, int k, int ch) const { return ldxc(g, gidx(i, j, k), ch); }
  __device__ __forceinline__ float4 ld4(int i, int j, int k) const { return ldx(g, gidx(i, j, k)); }
};

template <class A>
__device__ __forceinline__ float3 velAt(const A& V, float3 p) { return make_float3(sampU(V, p), sampV(V, p), sampW(V, p)); }
__device__ __forceinline__ float3 clampX3(const Dims& d, float3 p) {
  return make_float3(clampt_(p.x, 0.0f, (float)d.nx), clampt_(p.y, 0.0f, (float)d.ny), clampt_(p.z, 0.0f, (float)d.nz));
}
template <class A>
__device__ __forceinline__ float3 traceM(const Dims& d, const A& V, float3 k, float3 p, float3 v0, float s, bool mid) {
  if (!mid) return traceV(d, k, p, v0, s);
  float3 m = make_float3(p.x - 0.5f * s * k.x * v0.x, p.y - 0.5f * s * k.y * v0.y, p.z - 0.5f * s * k.z * v0.z);
  return traceV(d, k, p, velAt(V, clampX3(d, m)), s);
}
template <class A>
__device__ __forceinline__ float3 traceQM(const Dims& d, const A& V, float3 k, float3 p, float3 v0, float fall, float s,
                                          bool mid) {
  if (!mid) return traceQ(d, k, p, v0, fall, s);
  float3 m = make_float3(p.x - 0.5f * s * k.x * v0.x, p.y - 0.5f * s * k.y * v0.y, p.z - 0.5f * s * k.z * (v0.z - fall));
  return traceQ(d, k, p, velAt(V, clampX3(d, m)), fall, s);
}
#define HUR_FACE                                                                                     \
  Dims d = dims(u);                                                                                  \
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
// The hurricane-mode advections (midpoint trajectories, the prepared winds Vval, settling ice and
// maritime graupel) over LDS rings. Each sample reads through the window-guaranteed view (RingW:
// no window tests, plane offsets selected) when its whole wave's boxes lie within one point of
// their cells, and through the checked ring (global fallback) otherwise; both read the same texels
// with the same arithmetic. Two wave votes decide it: one per point on the Courant numbers, which
// keeps every midpoint sample's box in the window (|m - p| < 1/2), and one per trace on its final
// point (|t - p| < 1 on each axis).
__device__ __forceinline__ bool near_box(float3 t, float3 p) {
  return fabsf(t.x - p.x) < 0.999f && fabsf(t.y - p.y) < 0.999f && fabsf(t.z - p.z) < 0.999f;
}
__device__ __forceinline__ float sampU_r(const Ring& r, float3 t, float3 p) {
  return wave_all(near_box(t, p)) ? sampU(r.w(), t) : sampU(r, t);
}
__device__ __forceinline__ float sampV_r(const Ring& r, float3 t, float3 p) {
  return wave_all(near_box(t, p)) ? sampV(r.w(), t) : sampV(r, t);
}
__device__ __forceinline__ float sampW_r(const Ring& r, float3 t, float3 p) {
  return wave_all(near_box(t, p)) ? sampW(r.w(), t) : sampW(r, t);
}
__device__ __forceinline__ float sampC_r(const Ring& r, float3 t, float3 p) {
  return wave_all(near_box(t, p)) ? sampC(r.w(), t) : sampC(r, t);
}
__device__ __forceinline__ float2 mmV_r(const Ring& r, float3 t, float3 p, float ox, float oy, float oz, int ch) {
  return wave_all(near_box(t, p)) ? mmV(r.w(), t, ox, oy, oz, ch) : mmV(r, t, ox, oy, oz, ch);
}
__device__ __forceinline__ float4 sampQ_r(const Ring& r, float3 t, float3 p) {
  return wave_all(near_box(t, p)) ? sampQ(r.w(), t) : sampQ(r, t);
}
__device__ __forceinline__ float sampQ1_r(const Ring& r, float3 t, float3 p, int ch) {
  return wave_all(near_box(t, p)) ? sampQ1(r.w(), t, ch) : sampQ1(r, t, ch);
}
__device__ __forceinline__ float2 mmQ1_r(const Ring& r, float3 t, float3 p, int ch) {
  return wave_all(near_box(t, p)) ? mmQ1(r.w(), t, ch) : mmQ1(r, t, ch);
}
// the trace of one point: the midpoint's wind from the window view when the point's wave is near
__device__ __forceinline__ float3 traceM_r(const Dims& d, const Ring& V, bool nearA, float3 k, float3 p, float3 v0,
                                           float s, bool mid) {
  return nearA ? traceM(d, V.w(), k, p, v0, s, mid) : traceM(d, V, k, p, v0, s, mid);
}
__device__ __forceinline__ float3 traceQM_r(const Dims& d, const Ring& V, bool nearA, float3 k, float3 p, float3 v0,
                                            float fall, float s, bool mid) {
  return nearA ? traceQM(d, V.w(), k, p, v0, fall, s, mid) : traceQM(d, V, k, p, v0, fall, s, mid);
}
__device__ __forceinline__ float4 advv1_h_point(const MU& u, const Dims& d, const Ring& gv, const Ring& gq, int x, int y,
                                                int z) {
  const bool mid = u.v[U_CN].w > 0.5f;
  const Loc L = localVel(gv.w(), x, y, z);
  const bool nA = wave_all(near_traces(d, L));
  const float3 k = make_float3(d.kx, d.kx, d.kz);
  const float fx = (float)x, fy = (float)y, fz = (float)z;
  const float3 pu = make_float3(fx, fy + 0.5f, fz + 0.5f), pv = make_float3(fx + 0.5f, fy, fz + 0.5f);
  const float3 pw = make_float3(fx + 0.5f, fy + 0.5f, fz), pc = make_float3(fx + 0.5f, fy + 0.5f, fz + 0.5f);
  float4 o;
  o.x = sampU_r(gq, traceM_r(d, gv, nA, k, pu, L.vu, 1.0f, mid), pu);
  o.y = sampV_r(gq, traceM_r(d, gv, nA, k, pv, L.vv, 1.0f, mid), pv);
  o.z = sampW_r(gq, traceM_r(d, gv, nA, k, pw, L.vw, 1.0f, mid), pw);
  o.w = sampC_r(gq, traceM_r(d, gv, nA, k, pc, L.vc, 1.0f, mid), pc);
  return o;
}
__device__ __forceinline__ float4 advv2_h_point(const MU& u, const Dims& d, const Ring& gv, const Ring& gq, const Ring& gt,
                                                int x, int y, int z) {
  const bool mid = u.v[U_CN].w > 0.5f;
  const Loc L = localVel(gv.w(), x, y, z);
  const bool nA = wave_all(near_traces(d, L));
  const float3 k = make_float3(d.kx, d.kx, d.kz);
  const float fx = (float)x, fy = (float)y, fz = (float)z;
  const float3 pu = make_float3(fx, fy + 0.5f, fz + 0.5f), pv = make_float3(fx + 0.5f, fy, fz + 0.5f);
  const float3 pw = make_float3(fx + 0.5f, fy + 0.5f, fz), pc = make_float3(fx + 0.5f, fy + 0.5f, fz + 0.5f);
  const float4 hat = gt.w().ld4(x, y, z), old = gq.w().ld4(x, y, z);
  float4 b;
  b.x = sampU_r(gt, traceM_r(d, gv, nA, k, pu, L.vu, -1.0f, mid), pu);
  b.y = sampV_r(gt, traceM_r(d, gv, nA, k, pv, L.vv, -1.0f, mid), pv);
  b.z = sampW_r(gt, traceM_r(d, gv, nA, k, pw, L.vw, -1.0f, mid), pw);
  b.w = sampC_r(gt, traceM_r(d, gv, nA, k, pc, L.vc, -1.0f, mid), pc);
  const float cx = L.vc.x * k.x, cy = L.vc.y * k.y, cz = L.vc.z * k.z;
  const float cour = sqrtf(cx * cx + cy * cy + cz * cz);
  const float sm = 0.5f * u.v[U_MISC].y * clampf_(cour / 0.35f, u.v[U_DZV].w, 1.0f);
  float4 r = make_float4(hat.x + sm * (old.x - b.x), hat.y + sm * (old.y - b.y), hat.z + sm * (old.z - b.z),
                         hat.w + sm * (old.w - b.w));
  float2 m;
  m = mmV_r(gq, traceM_r(d, gv, nA, k, pu, L.vu, 1.0f, mid), pu, 0.0f, 0.5f, 0.5f, 0); r.x = vclamp_(r.x, m);
  m = mmV_r(gq, traceM_r(d, gv, nA, k, pv, L.vv, 1.0f, mid), pv, 0.5f, 0.0f, 0.5f, 1); r.y = vclamp_(r.y, m);
  m = mmV_r(gq, traceM_r(d, gv, nA, k, pw, L.vw, 1.0f, mid), pw, 0.5f, 0.5f, 0.0f, 2); r.z = vclamp_(r.z, m);
  m = mmV_r(gq, traceM_r(d, gv, nA, k, pc, L.vc, 1.0f, mid), pc, 0.5f, 0.5f, 0.5f, 3); r.w = vclamp_(r.w, m);
  return r;
}
struct FallH {
  float r, n, s, g, i;
};
__device__ __forceinline__ FallH fall_h(const MU& u, float4 q1, float4 q2, float rho) {
  const bool thm = u.v[U_MICRO2].z > 0.5f;
  const float rs = u.v[U_STEP].z;
  const float2 fv = vtRain(thm, q1.w, q2.z, rho, rs);
  return FallH{fv.x, thm ? fv.y : 0.0f, vtS(q2.x, rho, rs), vtGrauM(thm, maritime_of(u), q2.y, rho, rs), vtI(q1.z, rho)};
}
// The midpoints of a cell's traces in one direction (s) differ only in z (each species' fall speed), so
// the wind's x-y bilinears on the window's three planes are shared: computed once, each midpoint wind
// is then one z lerp per channel between the planes around it. The same texels and arithmetic as
// velAt over the window view (x, then y, then z lerps), valid when the point's vote keeps every
// midpoint within half a cell.
struct MidWind {
  float u[3], v[3], w[3];  // bilinears of planes zc - 1, zc, zc + 1
  int zc;
  __device__ __forceinline__ static float bil(const RingW& R, float tx, float ty, int r, int ch) {
    const float ax = __builtin_amdgcn_fractf(tx), ay = __builtin_amdgcn_fractf(ty);
    const int b = (r == 0 ? R.om : (r == 1 ? R.oc : R.op)) + ch * TPS + __mul24(floori_(ty) - R.y0 + 1, TPW) +
                  (floori_(tx) - R.x0 + 1);
    return lerpf(lerpf(R.s[b], R.s[b + 1], ax), lerpf(R.s[b + TPW], R.s[b + TPW + 1], ax), ay);
  }
  __device__ __forceinline__ MidWind(const RingW& R, float mx, float my) : zc(R.zc) {
#pragma unroll
    for (int r = 0; r < 3; r++) {
      u[r] = bil(R, mx, my - 0.5f, r, 0);
      v[r] = bil(R, mx - 0.5f, my, r, 1);
      w[r] = bil(R, mx - 0.5f, my - 0.5f, r, 2);
    }
  }
  __device__ __forceinline__ static float zl(const float* c, float tz, int zc) {
    const float az = __builtin_amdgcn_fractf(tz);
    const bool lo = floori_(tz) < zc;  // the box's lower plane is zc - 1, else zc
    return lerpf(lo ? c[0] : c[1], lo ? c[1] : c[2], az);
  }
  __device__ __forceinline__ float3 at(float mz) const {
    return make_float3(zl(u, mz - 0.5f, zc), zl(v, mz - 0.5f, zc), zl(w, mz, zc));
  }
};
// traceQM's midpoint trace with the midpoint wind from a MidWind of the same direction
__device__ __forceinline__ float3 traceQ_mw(const Dims& d, const MidWind& M, float3 k, float3 p, float3 v0, float fall,
                                            float s) {
  const float mz = clampt_(p.z - 0.5f * s * k.z * (v0.z - fall), 0.0f, (float)d.nz);
  return traceQ(d, k, p, M.at(mz), fall, s);
}
// the midpoint's x and y (shared by every species' trace in direction s)
__device__ __forceinline__ float2 mid_xy(const Dims& d, float3 k, float3 p, float3 v0, float s) {
  return make_float2(clampt_(p.x - 0.5f * s * k.x * v0.x, 0.0f, (float)d.nx),
                     clampt_(p.y - 0.5f * s * k.y * v0.y, 0.0f, (float)d.ny));
}

// as near_traces_q, with the settling ice's fall speed
__device__ __forceinline__ bool near_traces_qh(const Dims& d, float3 v, const FallH& f) {
  return near_traces_q(d, v, Fall{f.r, f.n, f.s, f.g}) && near_traces_q(d, v, Fall{f.i, 0.0f, 0.0f, 0.0f});
}
template <class TR>
__device__ __forceinline__ void advq1_h_body(bool thm, const Ring& g1, const Ring& g2, const FallH& f, float3 p,
                                             const TR& tr, float4& o1, float4& o2) {
  const float3 t0 = tr(0.0f, 1.0f);
  float4 a = sampQ_r(g1, t0, p);
  if (f.i > 0.0f) a.z = sampQ1_r(g1, tr(f.i, 1.0f), p, 2);
  const float r = sampQ1_r(g1, tr(f.r, 1.0f), p, 3);
  const float4 b = sampQ_r(g2, t0, p);
  const float sv = sampQ1_r(g2, tr(f.s, 1.0f), p, 0);
  const float g = sampQ1_r(g2, tr(f.g, 1.0f), p, 1);
  const float nr = thm ? sampQ1_r(g2, tr(f.n, 1.0f), p, 2) : b.z;
  o1 = make_float4(a.x, a.y, a.z, r);
  o2 = make_float4(sv, g, nr, b.w);
}
__device__ __forceinline__ void advq1_h_point(const MU& u, const Dims& d, const float4* BS, const Ring& gv, const Ring& g1,
                                              const Ring& g2, int x, int y, int z, float4& o1, float4& o2) {
  const bool mid = u.v[U_CN].w > 0.5f, thm = u.v[U_MICRO2].z > 0.5f;
  const float3 v = cellVel(gv.w(), x, y, z), k = make_float3(d.kx, d.kx, d.kz);
  const float3 p = make_float3((float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f);
  const float4 q1 = g1.w().ld4(x, y, z), q2 = g2.w().ld4(x, y, z);
  const FallH f = fall_h(u, q1, q2, bsA(BS, z, d.nz).z);
  const bool nA = wave_all(near_traces_qh(d, v, f));
  if (mid && nA) {
    const float2 md = mid_xy(d, k, p, v, 1.0f);
    const MidWind M(gv.w(), md.x, md.y);
    advq1_h_body(thm, g1, g2, f, p, [&](float fall, float s) { return traceQ_mw(d, M, k, p, v, fall, s); }, o1, o2);
  } else {
    advq1_h_body(thm, g1, g2, f, p, [&](float fall, float s) { return traceQM_r(d, gv, nA, k, p, v, fall, s, mid); },
                 o1, o2);
  }
}
template <class TR>
__device__ __forceinline__ void advq2_h_body(const MU& u, const Ring& g1, const Ring& g2, const Ring& g1t, const Ring& g2t,
                                             const FallH& f, float3 p, float3 v, float3 k, float4 q1, float4 q2, float4 h1,
                                             float4 h2, const TR& tr, float4& o1, float4& o2) {
  const float cx = v.x * k.x, cy = v.y * k.y, cz = v.z * k.z;
  const float cour = sqrtf(cx * cx + cy * cy + cz * cz);
  const float sm = 0.5f * u.v[U_MISC].y * clampf_(cour / 0.35f, u.v[U_DZV].w, 1.0f);
  const float3 f0 = tr(0.0f, -1.0f);
  const float4 b1 = sampQ_r(g1t, f0, p);
  const float br = sampQ1_r(g1t, tr(f.r, -1.0f), p, 3);
  const float4 b2 = sampQ_r(g2t, f0, p);
  const float bs = sampQ1_r(g2t, tr(f.s, -1.0f), p, 0);
  const float bg = sampQ1_r(g2t, tr(f.g, -1.0f), p, 1);
  const float bn = sampQ1_r(g2t, tr(f.n, -1.0f), p, 2);
  const float bi = f.i > 0.0f ? sampQ1_r(g1t, tr(f.i, -1.0f), p, 2) : b1.z;
  float4 r1 = make_float4(h1.x + sm * (q1.x - b1.x), h1.y + sm * (q1.y - b1.y), h1.z + sm * (q1.z - bi),
                          h1.w + sm * (q1.w - br));
  float4 r2 = make_float4(h2.x + sm * (q2.x - bs), h2.y + sm * (q2.y - bg), h2.z + sm * (q2.z - bn),
                          h2.w + sm * (q2.w - b2.w));
  const float3 d0 = tr(0.0f, 1.0f);
  float2 m;
  m = mmQ1_r(g1, d0, p, 0); r1.x = vclamp_(r1.x, m);
  m = mmQ1_r(g1, d0, p, 1); r1.y = vclamp_(r1.y, m);
  m = mmQ1_r(g1, tr(f.i, 1.0f), p, 2); r1.z = vclamp_(r1.z, m);
  m = mmQ1_r(g1, tr(f.r, 1.0f), p, 3); r1.w = vclamp_(r1.w, m);
  m = mmQ1_r(g2, tr(f.s, 1.0f), p, 0); r2.x = vclamp_(r2.x, m);
  m = mmQ1_r(g2, tr(f.g, 1.0f), p, 1); r2.y = vclamp_(r2.y, m);
  m = mmQ1_r(g2, tr(f.n, 1.0f), p, 2); r2.z = vclamp_(r2.z, m);
  m = mmQ1_r(g2, d0, p, 3); r2.w = vclamp_(r2.w, m);
  o1 = make_float4(vmax_(r1.x, 0.0f), vmax_(r1.y, 0.0f), vmax_(r1.z, 0.0f), vmax_(r1.w, 0.0f));
  o2 = make_float4(vmax_(r2.x, 0.0f), vmax_(r2.y, 0.0f), vmax_(r2.z, 0.0f), vmax_(r2.w, 0.0f));
}
__device__ __forceinline__ void advq2_h_point(const MU& u, const Dims& d, const float4* BS, const Ring& gv, const Ring& g1,
                                              const Ring& g2, const Ring& g1t, const Ring& g2t, int x, int y, int z,
                                              float4& o1, float4& o2) {
  const bool mid = u.v[U_CN].w > 0.5f;
  const float3 v = cellVel(gv.w(), x, y, z), k = make_float3(d.kx, d.kx, d.kz);
  const float3 p = make_float3((float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f);
  const float4 q1 = g1.w().ld4(x, y, z), q2 = g2.w().ld4(x, y, z), h1 = g1t.w().ld4(x, y, z), h2 = g2t.w().ld4(x, y, z);
  const FallH f = fall_h(u, q1, q2, bsA(BS, z, d.nz).z);
  const bool nA = wave_all(near_traces_qh(d, v, f));
  if (mid && nA) {
    const float2 mf = mid_xy(d, k, p, v, -1.0f), md = mid_xy(d, k, p, v, 1.0f);
    const MidWind Mf(gv.w(), mf.x, mf.y), Md(gv.w(), md.x, md.y);
    advq2_h_body(u, g1, g2, g1t, g2t, f, p, v, k, q1, q2, h1, h2,
                 [&](float fall, float s) { return traceQ_mw(d, s > 0.0f ? Md : Mf, k, p, v, fall, s); }, o1, o2);
  } else {
    advq2_h_body(u, g1, g2, g1t, g2t, f, p, v, k, q1, q2, h1, h2,
                 [&](float fall, float s) { return traceQM_r(d, gv, nA, k, p, v, fall, s, mid); }, o1, o2);
  }
}

// z-marching drivers (as the outbreak advections). PREP: the carried winds Vval are a field of
// their own (the prep field), with a ring of their own; otherwise they are V.
template <bool PREP>
__global__ void __launch_bounds__(TILE_THREADS) k_advv1_hr(MU u, const float4* __restrict__ V,
                                                           const float4* __restrict__ Vval, float4* __restrict__ Vt) {
  __shared__ float lds[(PREP ? 2 : 1) * RING_FLOATS];
  TILE_SETUP
  Ring rv{lds, V, d.nx + 1, d.ny + 1, d.nz + 1, x0, y0, 0};
  Ring rq{lds + (PREP ? RING_FLOATS : 0), Vval, d.nx + 1, d.ny + 1, d.nz + 1, x0, y0, 0};
  for (int k = -1; k <= 1; k++) {
    rv.load(k, tid);
    if (PREP) rq.load(k, tid);
  }
  __syncthreads();
  const bool act = x <= d.nx && y <= d.ny;
  for (int z = 0; z <= d.nz; z++) {
    Ring::Pre pv, pq;
    const bool more = z < d.nz;
    if (more) {
      rv.fetch(z + 2, tid, pv);
      if (PREP) rq.fetch(z + 2, tid, pq);
    }
    rv.center(z);
    if (PREP) rq.center(z);
    if (act) stx(Vt, vidx(d, x, y, z), advv1_h_point(u, d, rv, PREP ? rq : rv, x, y, z));
    __syncthreads();
    if (more) {
      rv.store(z + 2, tid, pv);
      if (PREP) rq.store(z + 2, tid, pq);
    }
    __syncthreads();
  }
}
template <bool PREP>
__global__ void __launch_bounds__(TILE_THREADS) k_advv2_hr(MU u, const float4* __restrict__ V,
                                                           const float4* __restrict__ Vval, const float4* __restrict__ Vt,
                                                           float4* __restrict__ Vout) {
  __shared__ float lds[(PREP ? 3 : 2) * RING_FLOATS];
  TILE_SETUP
  Ring rv{lds, V, d.nx + 1, d.ny + 1, d.nz + 1, x0, y0, 0};
  Ring rt{lds + RING_FLOATS, Vt, d.nx + 1, d.ny + 1, d.nz + 1, x0, y0, 0};
  Ring rq{lds + (PREP ? 2 * RING_FLOATS : 0), Vval, d.nx + 1, d.ny + 1, d.nz + 1, x0, y0, 0};
  for (int k = -1; k <= 1; k++) {
    rv.load(k, tid);
    rt.load(k, tid);
    if (PREP) rq.load(k, tid);
  }
  __syncthreads();
  const bool act = x <= d.nx && y <= d.ny;
  for (int z = 0; z <= d.nz; z++) {
    Ring::Pre pv, pt, pq;
    const bool more = z < d.nz;
    if (more) {
      rv.fetch(z + 2, tid, pv);
      rt.fetch(z + 2, tid, pt);
      if (PREP) rq.fetch(z + 2, tid, pq);
    }
    rv.center(z);
    rt.center(z);
    if (PREP) rq.center(z);
    if (act) stx(Vout, vidx(d, x, y, z), advv2_h_point(u, d, rv, PREP ? rq : rv, rt, x, y, z));
    __syncthreads();
    if (more) {
      rv.store(z + 2, tid, pv);
      rt.store(z + 2, tid, pt);
      if (PREP) rq.store(z + 2, tid, pq);
    }
    __syncthreads();
  }
}
__global__ void __launch_bounds__(TILE_THREADS) k_advq1_hr(MU u, const float4* __restrict__ BS, const float4* __restrict__ V,
                                                           const float4* __restrict__ Q1, const float4* __restrict__ Q2,
                                                           float4* __restrict__ Q1t, float4* __restrict__ Q2t) {
  __shared__ float lds[3 * RING_FLOATS];
  TILE_SETUP
  Ring rv{lds, V, d.nx + 1, d.ny + 1, d.nz + 1, x0, y0, 0};
  Ring r1{lds + RING_FLOATS, Q1, d.nx, d.ny, d.nz, x0, y0, 0};
  Ring r2{lds + 2 * RING_FLOATS, Q2, d.nx, d.ny, d.nz, x0, y0, 0};
  for (int k = -1; k <= 1; k++) { rv.load(k, tid); r1.load(k, tid); r2.load(k, tid); }
  __syncthreads();
  const bool act = x < d.nx && y < d.ny;
  for (int z = 0; z < d.nz; z++) {
    Ring::Pre pv, p1, p2;
    rv.fetch(z + 2, tid, pv); r1.fetch(z + 2, tid, p1); r2.fetch(z + 2, tid, p2);
    rv.center(z); r1.center(z); r2.center(z);
    if (act) {
      float4 o1, o2;
      advq1_h_point(u, d, BS, rv, r1, r2, x, y, z, o1, o2);
      const uint32_t i = cidx(d, x, y, z);
      stx(Q1t, i, o1);
      stx(Q2t, i, o2);
    }
    __syncthreads();
    rv.store(z + 2, tid, pv); r1.store(z + 2, tid, p1); r2.store(z + 2, tid, p2);
    __syncthreads();
  }
}
__global__ void __launch_bounds__(TILE_THREADS) k_advq2_hr(MU u, const float4* __restrict__ BS, const float4* __restrict__ V,
                                                           const float4* __restrict__ Q1, const float4* __restrict__ Q2,
                                                           const float4* __restrict__ Q1t, const float4* __restrict__ Q2t,
                                                           float4* __restrict__ Q1o, float4* __restrict__ Q2o) {
  __shared__ float lds[5 * RING_FLOATS];
  TILE_SETUP
  Ring rv{lds, V, d.nx + 1, d.ny + 1, d.nz + 1, x0, y0, 0};
  Ring r1{lds + RING_FLOATS, Q1, d.nx, d.ny, d.nz, x0, y0, 0};
  Ring r2{lds + 2 * RING_FLOATS, Q2, d.nx, d.ny, d.nz, x0, y0, 0};
  Ring r3{lds + 3 * RING_FLOATS, Q1t, d.nx, d.ny, d.nz, x0, y0, 0};
  Ring r4{lds + 4 * RING_FLOATS, Q2t, d.nx, d.ny, d.nz, x0, y0, 0};
  for (int k = -1; k <= 1; k++) { rv.load(k, tid); r1.load(k, tid); r2.load(k, tid); r3.load(k, tid); r4.load(k, tid); }
  __syncthreads();
  const bool act = x < d.nx && y < d.ny;
  for (int z = 0; z < d.nz; z++) {
    Ring::Pre pv, p1, p2, p3, p4;
    rv.fetch(z + 2, tid, pv); r1.fetch(z + 2, tid, p1); r2.fetch(z + 2, tid, p2);
    r3.fetch(z + 2, tid, p3); r4.fetch(z + 2, tid, p4);
    rv.center(z); r1.center(z); r2.center(z); r3.center(z); r4.center(z);
    if (act) {
      float4 o1, o2;
      advq2_h_point(u, d, BS, rv, r1, r2, r3, r4, x, y, z, o1, o2);
      const uint32_t i = cidx(d, x, y, z);
      stx(Q1o, i, o1);
      stx(Q2o, i, o2);
    }
    __syncthreads();
    rv.store(z + 2, tid, pv); r1.store(z + 2, tid, p1); r2.store(z + 2, tid, p2);
    r3.store(z + 2, tid, p3); r4.store(z + 2, tid, p4);
    __syncthreads();
  }
}
// PREP (model.js:996-1026): the winds the advection carries, V less cn.x of the last pressure force
__global__ void __launch_bounds__(256) k_prep(MU u, const float4* __restrict__ V, const float* __restrict__ phi,
                                              float4* __restrict__ Vp) {
  HUR_FACE
  if (x > d.nx || y > d.ny || z > d.nz) return;
  const uint32_t vi = vidx(d, x, y, z);
  const float4 Vv = ldx(V, vi);
  float4 o = Vv;
  const float f = u.v[U_CN].x * u.v[U_CN].y;
  const float kx = f * d.dt / d.dx, kz = f * d.dt / d.dz;
  const float pc = Pphi(phi, u, d, x, y, z);
  if (y < d.ny && z < d.nz) o.x = Vv.x - kx * (pc - Pphi(phi, u, d, x - 1, y, z));
  if (x < d.nx && z < d.nz) o.y = Vv.y - kx * (pc - Pphi(phi, u, d, x, y - 1, z));
  if (x < d.nx && y < d.ny && z > 0 && z < d.nz) o.z = Vv.z - kz * (pc - Pphi(phi, u, d, x, y, z - 1));
  stx(Vp, vi, o);
}
// FILT (model.js:948-994): a weak 4th-order horizontal hyperdiffusion, every 4th step at 4x strength
__device__ __forceinline__ float4 bih4(const GIn& g, int x, int y, int z) {
  const float4 c0 = g.ld4(x, y, z);
  const float4 xp = g.ld4(x + 1, y, z), xm = g.ld4(x - 1, y, z), yp = g.ld4(x, y + 1, z), ym = g.ld4(x, y - 1, z);
  const float4 pp = g.ld4(x + 1, y + 1, z), pm = g.ld4(x + 1, y - 1, z), mp = g.ld4(x - 1, y + 1, z), mm = g.ld4(x - 1, y - 1, z);
  const float4 x2p = g.ld4(x + 2, y, z), x2m = g.ld4(x - 2, y, z), y2p = g.ld4(x, y + 2, z), y2m = g.ld4(x, y - 2, z);
#define OBL_BIH(ch)                                                                                                   \
  (20.0f * c0.ch - 8.0f * (xp.ch + xm.ch + yp.ch + ym.ch) + 2.0f * (pp.ch + pm.ch + mp.ch + mm.ch) + x2p.ch + x2m.ch + \
   y2p.ch + y2m.ch)
  const float4 r = make_float4(OBL_BIH(x), OBL_BIH(y), OBL_BIH(z), OBL_BIH(w));
#undef OBL_BIH
  return r;
}
__global__ void __launch_bounds__(256) k_filt(MU u, const float4* __restrict__ V, const float4* __restrict__ Q1,
                                              const float4* __restrict__ Q2, float4* __restrict__ Vo,
                                              float4* __restrict__ Q1o, float4* __restrict__ Q2o) {
  HUR_FACE
  if (x > d.nx || y > d.ny || z > d.nz) return;
  const float a = u.v[U_HUR5].z;
  const GIn gv{V, d.nx + 1, d.ny + 1, d.nz + 1};
  const float4 v = gv.ld4(x, y, z), bv = bih4(gv, x, y, z);
  float4 o = make_float4(v.x - a * bv.x, v.y - a * bv.y, v.z - a * bv.z, v.w - a * bv.w);
  if (z == 0 || z >= d.nz) o.z = 0.0f;  // the lid / ground faces of w stay zero
  stnt(Vo, vidx(d, x, y, z), o);
  if (x < d.nx && y < d.ny && z < d.nz) {
    const GIn g1{Q1, d.nx, d.ny, d.nz}, g2{Q2, d.nx, d.ny, d.nz};
    const uint32_t i = cidx(d, x, y, z);
    const float4 q = ldx(Q1, i), bq = bih4(g1, x, y, z), r = ldx(Q2, i), br = bih4(g2, x, y, z);
    stnt(Q1o, i, make_float4(fmaxf(q.x - a * bq.x, 0.0f), fmaxf(q.y - a * bq.y, 0.0f), fmaxf(q.z - a * bq.z, 0.0f),
                            fmaxf(q.w - a * bq.w, 0.0f)));
    stnt(Q2o, i, make_float4(fmaxf(r.x - a * br.x, 0.0f), fmaxf(r.y - a * br.y, 0.0f), fmaxf(r.z - a * br.z, 0.0f),
                            fmaxf(r.w - a * br.w, 0.0f)));
  }
}
#undef HUR_FACE

// ====================== launchers ======================
static dim3 faces_grid(const Dims& d) { return dim3((d.nx + 1 + 63) / 64, (d.ny + 1 + 3) / 4, d.nz + 1); }
static dim3 cells_grid(const Dims& d) { return dim3((d.nx + 63) / 64, (d.ny + 3) / 4, d.nz); }
static dim3 cols_grid(const Dims& d) { return dim3((d.nx + 15) / 16, (d.ny + 15) / 16, 1); }

void launch_lsf(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float* BD, float4* lsfA,
                float4* lsfB, float* phT) {
  hipLaunchKernelGGL(k_lsf, cols_grid(d), dim3(16, 16, 1), 0, s, u, BS, BD, lsfA, lsfB, phT);
}
void launch_ci(hipStream_t s, const MU& u, const Dims& d, const float* CI, float4* ciMap) {
  hipLaunchKernelGGL(k_ci, cols_grid(d), dim3(16, 16, 1), 0, s, u, CI, ciMap);
}
void launch_init(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* lsfA, const float4* lsfB,
                 const float4* VX, float4* V, float4* Q1, float4* Q2, int seed) {
  hipLaunchKernelGGL(k_init, faces_grid(d), dim3(64, 4, 1), 0, s, u, BS, lsfA, lsfB, VX, V, Q1, Q2, seed);
}
void launch_bubble(hipStream_t s, const MU& u, const Dims& d, float4* V0, float4* V1, double x, double y, double z,
                   double rh, double rz, double dth) {
  hipLaunchKernelGGL(k_bubble, faces_grid(d), dim3(64, 4, 1), 0, s, u, V0, V1, x, y, z, rh, rz, dth);
}
void launch_nut(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, const float4* Q1,
                const float4* sfc, float2* aux) {
  if (sfc && hur_on(u))
    hipLaunchKernelGGL(k_nut_t<true>, tcell_grid(d), dim3(TILE_THREADS), 0, s, u, BS, V, Q1, sfc, aux);
  else
    hipLaunchKernelGGL(k_nut_t<false>, tcell_grid(d), dim3(TILE_THREADS), 0, s, u, BS, V, Q1, sfc, aux);
}
#endif  // OBL_ADV_ONLY
#ifdef OBL_ADV_ONLY
void launch_adv(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, const float4* Q1,
                const float4* Q2, float4* Vout, float2* aux, float4* Q1o, float4* Q2o) {
  const dim3 gv = tface_grid(d), gq = tcell_grid(d);
  const int nv = (int)(gv.x * gv.y), nq = (int)(gq.x * gq.y);
  hipLaunchKernelGGL(k_adv_t, dim3(nv + nq), dim3(TILE_THREADS), 0, s, u, BS, V, Q1, Q2, Vout, aux, Q1o, Q2o, (int)gv.x, nv,
                     (int)gq.x, nq);
}
#else
void launch_phys(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* Vin, const float4* Q1in,
                 const float4* Q2in, const float2* aux, const float4* lsfA, const float4* lsfB, const float* phT,
                 const float4* ciMap, float4* Vout, float4* Q1o, float4* Q2o) {
  MicroDump none = {nullptr, nullptr, 0, 0u, 0.0f, 0.0f};
  SurrPhys nos = {};
  HurArgs noh = {nullptr, nullptr, nullptr};
  hipLaunchKernelGGL((k_phys<PHYS_PLAIN, false>), tface_grid(d), dim3(TILE_THREADS), 0, s, u, BS, Vin, Q1in, Q2in, aux,
                     lsfA, lsfB, phT, ciMap, Vout, Q1o, Q2o, none, nos, noh);
}
void launch_phys_hur(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* Vin, const float4* Q1in,
                     const float4* Q2in, const float2* aux, const float4* lsfA, const float4* lsfB, const float* phT,
                     const float4* ciMap, float4* Vout, float4* Q1o, float4* Q2o, const HurArgs& hu) {
  MicroDump none = {nullptr, nullptr, 0, 0u, 0.0f, 0.0f};
  SurrPhys nos = {};
  hipLaunchKernelGGL((k_phys<PHYS_PLAIN, true>), tface_grid(d), dim3(TILE_THREADS), 0, s, u, BS, Vin, Q1in, Q2in, aux,
                     lsfA, lsfB, phT, ciMap, Vout, Q1o, Q2o, none, nos, hu);
}
void launch_phys_surr(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* Vin, const float4* Q1in,
                      const float4* Q2in, const float2* aux, const float4* lsfA, const float4* lsfB, const float* phT,
                      const float4* ciMap, float4* Vout, float4* Q1o, float4* Q2o, const SurrPhys& sp) {
  MicroDump none = {nullptr, nullptr, 0, 0u, 0.0f, 0.0f};
  HurArgs noh = {nullptr, nullptr, nullptr};
  hipLaunchKernelGGL((k_phys<PHYS_SURR, false>), tface_grid(d), dim3(TILE_THREADS), 0, s, u, BS, Vin, Q1in, Q2in, aux,
                     lsfA, lsfB, phT, ciMap, Vout, Q1o, Q2o, none, sp, noh);
}
void launch_phys_dump(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* Vin, const float4* Q1in,
                      const float4* Q2in, const float2* aux, const float4* lsfA, const float4* lsfB, const float* phT,
                      const float4* ciMap, float4* Vout, float4* Q1o, float4* Q2o, const MicroDump& dump) {
  SurrPhys nos = {};
  HurArgs noh = {nullptr, nullptr, nullptr};
  hipLaunchKernelGGL((k_phys<PHYS_DUMP, false>), tface_grid(d), dim3(TILE_THREADS), 0, s, u, BS, Vin, Q1in, Q2in, aux,
                     lsfA, lsfB, phT, ciMap, Vout, Q1o, Q2o, dump, nos, noh);
}
void launch_div(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, float* rhs) {
  hipLaunchKernelGGL(k_div, dim3((d.nx + 63) / 64, (d.ny + 3) / 4, (d.nz + ZL - 1) / ZL), dim3(64, 4, 1), 0, s, u, BS, V, rhs);
}
bool div_fold_fits(int Nx, int Ny) { return Nx % 8 == 0 && Ny % 8 == 0; }
void launch_div_fold(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, float* rhs, int P,
                     uint16_t* out, int RB, int KB) {
  const int Nx = (int)u.v[U_PAD].y, Ny = (int)u.v[U_PAD].z;
  const dim3 grid((Nx / 2 + 63) / 64, Ny / 4, (d.nz + 1) / 2);
  if (P == 2) hipLaunchKernelGGL(k_div_fold<2>, grid, dim3(512), 0, s, u, BS, V, rhs, out, RB, KB);
  else hipLaunchKernelGGL(k_div_fold<3>, grid, dim3(512), 0, s, u, BS, V, rhs, out, RB, KB);
}
void launch_proj(hipStream_t s, const MU& u, const Dims& d, const float4* V, const float* phi, float4* Vout) {
  hipLaunchKernelGGL(k_proj, faces_grid(d), dim3(64, 4, 1), 0, s, u, V, phi, Vout);
}

void launch_advv1_hur(hipStream_t s, const MU& u, const Dims& d, const float4* V, const float4* Vval, float4* Vt) {
  if (Vval != V) hipLaunchKernelGGL(k_advv1_hr<true>, tface_grid(d), dim3(TILE_THREADS), 0, s, u, V, Vval, Vt);
  else hipLaunchKernelGGL(k_advv1_hr<false>, tface_grid(d), dim3(TILE_THREADS), 0, s, u, V, Vval, Vt);
}
void launch_advv2_hur(hipStream_t s, const MU& u, const Dims& d, const float4* V, const float4* Vval, const float4* Vt,
                      float4* Vout) {
  if (Vval != V) hipLaunchKernelGGL(k_advv2_hr<true>, tface_grid(d), dim3(TILE_THREADS), 0, s, u, V, Vval, Vt, Vout);
  else hipLaunchKernelGGL(k_advv2_hr<false>, tface_grid(d), dim3(TILE_THREADS), 0, s, u, V, Vval, Vt, Vout);
}
void launch_advq1_hur(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, const float4* Q1,
                      const float4* Q2, float4* Q1t, float4* Q2t) {
  hipLaunchKernelGGL(k_advq1_hr, tcell_grid(d), dim3(TILE_THREADS), 0, s, u, BS, V, Q1, Q2, Q1t, Q2t);
}
void launch_advq2_hur(hipStream_t s, const MU& u, const Dims& d, const float4* BS, const float4* V, const float4* Q1,
                      const float4* Q2, const float4* Q1t, const float4* Q2t, float4* Q1o, float4* Q2o) {
  hipLaunchKernelGGL(k_advq2_hr, tcell_grid(d), dim3(TILE_THREADS), 0, s, u, BS, V, Q1, Q2, Q1t, Q2t, Q1o, Q2o);
}
void launch_prep(hipStream_t s, const MU& u, const Dims& d, const float4* V, const float* phi, float4* Vp) {
  hipLaunchKernelGGL(k_prep, faces_grid(d), dim3(64, 4, 1), 0, s, u, V, phi, Vp);
}
void launch_filt(hipStream_t s, const MU& u, const Dims& d, const float4* V, const float4* Q1, const float4* Q2,
                 float4* Vo, float4* Q1o, float4* Q2o) {
  hipLaunchKernelGGL(k_filt, faces_grid(d), dim3(64, 4, 1), 0, s, u, V, Q1, Q2, Vo, Q1o, Q2o);
}

#endif  // OBL_ADV_ONLY

}  // namespace obl
