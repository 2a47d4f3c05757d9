// Virtual WSR-88D cuts of the model state (gfx950), restating the reference's radar kernels
// (radar.js:41-268): pass 1 samples the state along 4/3-earth beams at real great-circle azimuths
// and quantises REF / VEL / SW / ZDR / RHO to Archive II codes (4 gates per 32-bit word), writing
// KDP per Doppler gate to a scratch radial; pass 2 accumulates PHIDP (two-way) along each radial.
// One work-item per gate in pass 1 (four lanes pack a word) and one wave per radial in pass 2.
// The state is sampled trilinearly with clamp-to-edge at texel coordinate u*S - 1/2 (the
// reference's linear sampler), with exact FP32 weights.
#include <hip/hip_runtime.h>

#include "obl_device.h"
#include "obl_radar.h"

namespace obl {

namespace {
constexpr float RE_ = 6371000.0f;

struct Tex {  // a W x H x D texture of float4 (or of float2 when two = true) in HBM
  const float* p;
  int W, H, D;
  int stride;  // floats per texel
  __device__ __forceinline__ float at(int x, int y, int z, int ch) const {
    return ldx(p, (uint32_t)stride * off3(W, H, x, y, z) + (uint32_t)ch);
  }
  // trilinear, clamp-to-edge, at texel coordinates (tx, ty, tz)
  __device__ __forceinline__ float samp(float tx, float ty, float tz, int ch) const {
    float fx = floorf(tx), fy = floorf(ty), fz = floorf(tz);
    float ax = tx - fx, ay = ty - fy, az = tz - fz;
    int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
    int x1 = clampi(x0 + 1, 0, W - 1), y1 = clampi(y0 + 1, 0, H - 1), z1 = clampi(z0 + 1, 0, D - 1);
    x0 = clampi(x0, 0, W - 1); y0 = clampi(y0, 0, H - 1); z0 = clampi(z0, 0, D - 1);
    float c00 = lerpf(at(x0, y0, z0, ch), at(x1, y0, z0, ch), ax), c10 = lerpf(at(x0, y1, z0, ch), at(x1, y1, z0, ch), ax);
    float c01 = lerpf(at(x0, y0, z1, ch), at(x1, y0, z1, ch), ax), c11 = lerpf(at(x0, y1, z1, ch), at(x1, y1, z1, ch), ax);
    return lerpf(lerpf(c00, c10, ay), lerpf(c01, c11, ay), az);
  }
  // textureSampleLevel at normalised coordinates uc (texel coordinate uc * S - 1/2)
  __device__ __forceinline__ float samp_uv(float3 uc, int ch) const {
    return samp(uc.x * (float)W - 0.5f, uc.y * (float)H - 0.5f, uc.z * (float)D - 0.5f, ch);
  }
};

struct Smp {
  float zr, zs, zg, vt;
  float3 u;
  float T, qr, rho, nu, inside, lr;
};

struct RadarCtx {
  const RadarU& R;
  const float4* BS;
  Tex V, Q1, Q2, aux;
  __device__ __forceinline__ float4 bsA_(int k) const { return bsA(BS, k, (int)R.mdl.z); }
  __device__ __forceinline__ float4 bsB_(int k) const { return bsB(BS, k, (int)R.mdl.z); }

  // great-circle destination (bearing a rad, ground distance s m) -> azimuthal-equidistant model metres
  __device__ __forceinline__ float2 gatePos(float a, float s) const {
    float p1 = R.geo.z, l1 = R.geo.w, d = s / RE_;
    float sp2 = sinf(p1) * cosf(d) + cosf(p1) * sinf(d) * cosf(a);
    float p2 = asinf(clampf_(sp2, -1.0f, 1.0f));
    float l2 = l1 + atan2f(sinf(a) * sinf(d) * cosf(p1), cosf(d) - sinf(p1) * sp2);
    float p0 = R.geo.x, dl = l2 - R.geo.y;
    float cc = clampf_(sinf(p0) * sinf(p2) + cosf(p0) * cosf(p2) * cosf(dl), -1.0f, 1.0f);
    float c = acosf(cc);
    float k = c < 1e-6f ? 1.0f : c / sinf(c);
    float x = RE_ * k * cosf(p2) * sinf(dl);
    float y = RE_ * k * (cosf(p0) * sinf(p2) - sinf(p0) * cosf(p2) * cosf(dl));
    return make_float2(x - R.mdl2.z, y - R.mdl2.w);
  }
  // 4/3-earth beam: height above the antenna and ground distance at slant range r
  __device__ __forceinline__ float2 beam(float r, float el) const {
    float a = R.site.w;
    float q = r * r + 2.0f * r * a * sinf(el);
    float h = q / (sqrtf(q + a * a) + a);
    float s = a * asinf(clampf_(r * cosf(el) / (a + h), -1.0f, 1.0f));
    return make_float2(h, s);
  }

  __device__ Smp sample(float2 xy, float z) const {
    Smp o;
    const float3 n = make_float3(R.mdl.x, R.mdl.y, R.mdl.z);
    const float dx = R.mdl.w, dz = R.mdl2.x;
    const float3 g = make_float3(xy.x / dx, xy.y / dx, z / dz);
    o.inside = (g.x >= 0.0f && g.y >= 0.0f && g.z >= 0.0f && g.x <= n.x && g.y <= n.y && g.z <= n.z && z >= 0.0f) ? 1.0f : 0.0f;
    const float3 gc = make_float3(clampf_(g.x, 0.0f, n.x), clampf_(g.y, 0.0f, n.y), clampf_(g.z, 0.0f, n.z));
    const float3 uc = make_float3(gc.x / n.x, gc.y / n.y, gc.z / n.z);
    const float q1w = Q1.samp_uv(uc, 3);
    const float q2x = Q2.samp_uv(uc, 0), q2y = Q2.samp_uv(uc, 1), q2z = Q2.samp_uv(uc, 2);
    const float3 N1 = make_float3(n.x + 1.0f, n.y + 1.0f, n.z + 1.0f);
    const float u = V.samp_uv(make_float3((gc.x + 0.5f) / N1.x, gc.y / N1.y, gc.z / N1.z), 0);
    const float v = V.samp_uv(make_float3(gc.x / N1.x, (gc.y + 0.5f) / N1.y, gc.z / N1.z), 1);
    const float w = V.samp_uv(make_float3(gc.x / N1.x, gc.y / N1.y, (gc.z + 0.5f) / N1.z), 2);
    const float th = V.samp_uv(make_float3(gc.x / N1.x, gc.y / N1.y, gc.z / N1.z), 3);
    const int k = (int)clampf_(floorf(gc.z), 0.0f, n.z - 1.0f);
    const float4 A = bsA_(k), B = bsB_(k);
    const float rho = A.z, T = th * B.x;
    const float qr = fmaxf(q1w, 0.0f), qs = fmaxf(q2x, 0.0f), qg = fmaxf(q2y, 0.0f);
    o.zr = 3.63e9f * pw(rho * qr, 1.75f);
    o.zs = 9.80e8f * pw(rho * qs, 1.75f) * 0.224f * 4.0f;
    o.zg = (T < 273.15f ? 1.2e10f : 4.33e10f) * pw(rho * qg, 1.75f);
    const float sq = sqrtf(R.phys.x / rho);
    float vr = fminf(36.34f * pw(fmaxf(0.001f * rho * qr, 0.0f), 0.1364f) * sq, 10.0f * sq);
    const float vs = clampf_(1.1f * pw(fmaxf(1000.0f * rho * qs, 1e-6f), 0.11f), 0.3f, 2.2f) * sq;
    float vg = clampf_(9.0f * pw(fmaxf(1000.0f * rho * qg, 1e-6f), 0.42f), 1.0f, 55.0f) * sq;
    const float rr = qg / 0.009f;
    vg = fminf(vg * (1.0f + rr * rr), 65.0f);
    if (R.phys2.x > 0.5f && R.phys.z < 0.5f) {  // hurricane mode's maritime graupel, as in the model
      o.zg = (T < 273.15f ? 1.3e9f : 1.0e10f) * pw(rho * qg, 1.75f);
      vg = clampf_(5.0f * pw(fmaxf(1000.0f * rho * qg, 1e-6f), 0.25f), 1.0f, 9.0f) * sq;
    }
    // Marshall-Palmer slope for the dual-pol drop size; Thompson: the predicted drop sizes and graupel
    o.lr = pw(3.14159f * 1000.0f * 8e6f / fmaxf(rho * qr, 1e-9f), 0.25f);
    if (R.phys.z > 0.5f) {
      const float nr = fmaxf(q2z, 0.0f);
      if (qr > 1e-9f) {
        const float l = clampf_(pw(3.14159f * 1000.0f * fmaxf(nr, 1e-3f) / qr, 0.33333f), 1000.0f, 1.0e5f);
        const float nb = qr * l * l * l / 3141.59f;
        o.zr = 720e18f * rho * nb / pw(l, 6.0f);
        o.lr = l;
        vr = fminf(3949.0f * pw(l, -0.8f), 10.0f) * sq;
      } else {
        o.zr = 0.0f;
      }
      if (qg > 1e-9f) {
        const float n0 = clampf_(200.0f / (rho * qg), 1e4f, 3e6f);
        const float l = pw(3.14159f * 500.0f * n0 / (rho * qg), 0.25f);
        o.zg = 720e18f * n0 / pw(l, 7.0f) * (T < 273.15f ? 0.0666f : 0.24f);
        vg = fminf(2483.0f * pw(l, -0.89f), 30.0f) * sq;
      } else {
        o.zg = 0.0f;
      }
    }
    const float zt = o.zr + o.zs + o.zg;
    o.vt = zt > 1e-6f ? (o.zr * vr + o.zs * vs + o.zg * vg) / zt : 0.0f;
    o.u = make_float3(u, v, w);
    o.T = T;
    o.qr = qr;
    o.rho = rho;
    o.nu = aux.samp_uv(uc, 0);
    return o;
  }
};

__device__ __forceinline__ uint32_t code(float v, float sc, float ofs) {
  return (uint32_t)clampf_(rintf(v * sc + ofs), 2.0f, 255.0f);  // WGSL round: half to even
}
__device__ __forceinline__ float db2lin(float d) { return pw(10.0f, 0.1f * d); }
__device__ __forceinline__ float sq_(float x) { return x * x; }

struct Gate {
  uint32_t zref, vel, sw, zdr, rho;
  float kdp;
};

// a gate's geometry: position, radial direction (unit, along the beam), beam-centre height and the heights
// 0.4 beam widths below and above it, all above the model's ground
struct GateGeom {
  float2 p, e;
  float zc, hl, hu;
};
constexpr float BEAM_W = 0.0165f;  // ~0.95 degree half-power beam width

__device__ __forceinline__ GateGeom gate_geom(const RadarCtx& C, int ri, int gi) {
  const RadarU& R = C.R;
  GateGeom g;
  const float r = R.gates.z + R.gates.w * (float)gi;
  const float el = R.cut.x;
  const float az = (R.cut.y + R.cut.z * (float)ri) * 0.017453293f;
  const float2 hc = C.beam(r, el);
  g.p = C.gatePos(az, hc.y);
  const float2 p2 = C.gatePos(az, hc.y + 200.0f);
  float2 e = make_float2(p2.x - g.p.x, p2.y - g.p.y);
  const float el_ = sqrtf(e.x * e.x + e.y * e.y);
  g.e = make_float2(e.x / el_, e.y / el_);
  const float zAnt = R.site.z - R.mdl2.y;  // antenna above the model's ground
  g.zc = zAnt + hc.x;
  // beam filling: three heights across the beam, Gaussian weights
  g.hl = C.beam(r, el - 0.4f * BEAM_W).x + zAnt;
  g.hu = C.beam(r, el + 0.4f * BEAM_W).x + zAnt;
  return g;
}

__device__ Gate gate(const RadarCtx& C, const float4* tor, int ri, int gi) {
  const RadarU& R = C.R;
  Gate G{0u, 0u, 0u, 0u, 0u, -1.0f};
  const float r = R.gates.z + R.gates.w * (float)gi;
  const float el = R.cut.x;
  const float bw = BEAM_W;
  const GateGeom gg = gate_geom(C, ri, gi);
  const float2 p = gg.p, e = gg.e;
  const float zc = gg.zc;
  if (zc > R.mdl.z * R.mdl2.x - 200.0f) return G;
  const float hl = gg.hl, hu = gg.hu;
  const Smp s0 = C.sample(p, fmaxf(zc, 5.0f)), s1 = C.sample(p, fmaxf(hl, 5.0f)), s2 = C.sample(p, fmaxf(hu, 5.0f));
  if (s0.inside < 0.5f) return G;
  const float zr = 0.5f * s0.zr + 0.25f * (s1.zr + s2.zr), zs = 0.5f * s0.zs + 0.25f * (s1.zs + s2.zs);
  const float zg = 0.5f * s0.zg + 0.25f * (s1.zg + s2.zg);
  const float zmet = zr + zs + zg;
  // sub-grid tornadoes: Rankine vortex smeared over the beam (circulation kept), lofted debris
  float zdeb = 0.0f, swt = 0.0f;
  float2 dvt = make_float2(0.0f, 0.0f);
  const float bwm = 0.5f * r * bw;
  for (int q = 0; q < (int)R.phys.w; q++) {
    const float4 A = tor[2 * q], B = tor[2 * q + 1];
    const float2 d = make_float2(p.x - A.x, p.y - A.y);
    const float rt = sqrtf(d.x * d.x + d.y * d.y);
    const float Re = sqrtf(A.z * A.z + bwm * bwm + 125.0f * 125.0f);
    if (rt > 8.0f * Re) continue;
    const float vmx = A.w * A.z / Re;
    const float vt = vmx * (rt < Re ? rt / Re : Re / fmaxf(rt, 1.0f));
    const float2 et = make_float2(-d.y / fmaxf(rt, 1.0f), d.x / fmaxf(rt, 1.0f));
    const float2 er = make_float2(d.x / fmaxf(rt, 1.0f), d.y / fmaxf(rt, 1.0f));
    const float fz = 1.0f - sstep(1200.0f, 3500.0f, zc);
    dvt.x += (vt * et.x - 0.25f * vt * er.x) * fz;
    dvt.y += (vt * et.y - 0.25f * vt * er.y) * fz;
    swt = fmaxf(swt, vmx * 0.4f * expf(-rt * rt / (4.0f * Re * Re)) * fz);
    if (B.z > 0.0f && zc < B.x) {
      const float rd = fmaxf(1.4f * A.z, 150.0f);
      zdeb += db2lin(B.y) * fminf(1.0f, rd * rd / (Re * Re)) * expf(-rt * rt / (Re * Re)) *
              (1.0f - sstep(0.55f * B.x, B.x, zc)) * B.z;
    }
  }
  const float ztot = zmet + zdeb;
  if (ztot <= 1e-6f) return G;
  const float dbz = 10.0f * logf(ztot) / 2.302585f;
  if (dbz < R.phys.y) return G;  // echo floor: below it the gate is empty
  G.zref = code(dbz, 2.0f, 66.0f);
  if (gi >= (int)R.gates.y) return G;
  // radial velocity: model wind along the beam plus the tornadoes, minus the echo's fall speed; aliasing
  const float ce = cosf(el), se = sinf(el);
  const float vtf = s0.vt * zmet / ztot;
  float vr = ((s0.u.x + dvt.x) * e.x + (s0.u.y + dvt.y) * e.y) * ce + (s0.u.z - vtf) * se;
  const float nyq = R.nyq.x;
  vr = vr - 2.0f * nyq * floorf((vr + nyq) / (2.0f * nyq));
  G.vel = code(vr, 2.0f, 129.0f);
  // spectrum width: SGS turbulence, shear across the beam, fall-speed spread, tornado
  const float vru = (s2.u.x * e.x + s2.u.y * e.y) * ce, vrl = (s1.u.x * e.x + s1.u.y * e.y) * ce;
  const float swv = sqrtf(1.0f + sq_(clampf_(2.5f * sqrtf(fmaxf(s0.nu, 0.0f) / 1000.0f), 0.0f, 8.0f)) + sq_((vru - vrl) / 2.5f) +
                          sq_(0.15f * s0.vt * se) + sq_(swt));
  G.sw = code(fminf(swv, 19.5f), 2.0f, 129.0f);
  // dual polarisation from the hydrometeor mix
  const float d0 = 3.67f / s0.lr * 1000.0f;  // rain median volume diameter (mm)
  const bool melt = s0.T > 273.15f;
  const float zdrR = clampf_(0.2f + 1.15f * (d0 - 0.5f), 0.1f, 4.8f);
  const float zdrS = melt ? 1.3f : 0.3f;
  const float zdrG = melt ? (zg > 3.2e5f ? 0.0f : 0.7f) : 0.15f;
  const float zv = zr / db2lin(zdrR) + zs / db2lin(zdrS) + zg / db2lin(zdrG) + zdeb / db2lin(0.2f);
  const float zdr = 10.0f * logf(fmaxf(ztot, 1e-9f) / fmaxf(zv, 1e-9f)) / 2.302585f;
  G.zdr = code(clampf_(zdr, -7.9f, 7.9f), 16.0f, 128.0f);
  const float fr = zr / ztot, fs = zs / ztot, fg = zg / ztot, fd = zdeb / ztot;
  float cc = fr * (d0 > 2.5f ? 0.986f : 0.995f) + fs * (melt ? 0.93f : 0.985f) +
             fg * (melt ? (zg > 3.2e5f ? 0.90f : 0.94f) : 0.985f) + fd * 0.55f;
  cc -= 0.24f * fr * fg;  // hail mixed with rain decorrelates the channels
  G.rho = code(clampf_(cc, 0.2f, 1.05f), 300.0f, -60.5f);
  // KDP from rain (S band): R = 0.017 Z^0.714, KDP = (R / 44)^(1 / 0.822)
  G.kdp = pw(0.017f * pw(fmaxf(zr, 0.0f), 0.714f) / 44.0f, 1.2165f);
  return G;
}
}  // namespace

// one work-item per gate, consecutive lanes along the beam: at one instruction the wave's samples
// lie 250 m apart, a cell or two, so their texels share cache lines. Four lanes make a word; the
// radial's gate count is a multiple of 4, so a word never straddles two radials.
__device__ __forceinline__ uint32_t pack4(uint32_t b) {
  return b | (__shfl_down(b, 1) << 8) | (__shfl_down(b, 2) << 16) | (__shfl_down(b, 3) << 24);
}

__global__ void __launch_bounds__(64) k_radar_cut(RadarU R, RadarArgs a) {
  const int ngr = (int)R.gates.x, nq = ngr / 4;
  const int nrad = (int)R.cut.w;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < nrad * ngr;  // whole 4-lane groups are live or not
  const int nx = (int)R.mdl.x, ny = (int)R.mdl.y, nz = (int)R.mdl.z;
  const RadarCtx C{R, a.BS, Tex{reinterpret_cast<const float*>(a.V), nx + 1, ny + 1, nz + 1, 4},
                   Tex{reinterpret_cast<const float*>(a.Q1), nx, ny, nz, 4}, Tex{reinterpret_cast<const float*>(a.Q2), nx, ny, nz, 4},
                   Tex{reinterpret_cast<const float*>(a.aux), nx, ny, nz, 2}};
  const int ri = live ? t / ngr : 0, gi = live ? t % ngr : 0;
  const int ng = (int)R.gates.y, ndq = ng / 4;
  Gate G{0u, 0u, 0u, 0u, 0u, -1.0f};
  if (live) G = gate(C, a.tor, ri, gi);
  if (live && gi < ng) a.kdp[ri * ng + gi] = G.kdp;
  // all lanes take part in the shuffles
  const uint32_t pr = pack4(G.zref), pv = pack4(G.vel), ps = pack4(G.sw), pz = pack4(G.zdr), ph = pack4(G.rho);
  if (!live || (gi & 3)) return;
  const int q = gi >> 2;
  const uint32_t mask = (uint32_t)R.offB.w;
  a.out[(uint32_t)R.offA.x + (uint32_t)(ri * nq + q)] = pr;
  if (q < ndq) {
    const uint32_t o = (uint32_t)(ri * ndq + q);
    if (mask & 2u) a.out[(uint32_t)R.offA.y + o] = pv;
    if (mask & 4u) a.out[(uint32_t)R.offA.z + o] = ps;
    if (mask & 8u) a.out[(uint32_t)R.offA.w + o] = pz;
    if (mask & 32u) a.out[(uint32_t)R.offB.y + o] = ph;
  }
}

// PHIDP accumulated along each radial from KDP (two-way), from the system offset: one wave per
// radial. The running sum keeps the reference's serial FP32 order (one add per gate, in gate order,
// each increment rounded before the add): a chunk of 64 gates is loaded coalesced, the increments are
// formed lane-parallel, and the chain walks the lanes with readlane; the codes are formed
// lane-parallel again and packed four to a word.
__global__ void __launch_bounds__(64) k_radar_phi(RadarU R, RadarArgs a) {
  const int ri = blockIdx.x;
  const int lane = threadIdx.x;
  const int ng = (int)R.gates.y;
  const float gw = R.gates.w;
  const float* kr = a.kdp + (size_t)ri * (size_t)ng;
  uint32_t* dst = a.out + (uint32_t)R.offB.x + (uint32_t)(ri * (ng / 4));
  float acc = R.nyq.w;
  for (int base = 0; base < ng; base += 64) {
    const int g = base + lane;
    const float kd = g < ng ? kr[g] : -1.0f;
    const float inc = kd >= 0.0f ? 2.0f * kd * gw * 0.001f : 0.0f;  // + 0 leaves acc (>= 25) unchanged
    float mine = 0.0f;
#pragma unroll
    for (int l = 0; l < 64; l++) {
      acc += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(inc), l));
      mine = lane == l ? acc : mine;
    }
    uint32_t b = 0u;
    if (kd >= 0.0f) {
      const float phi = mine - 360.0f * floorf(mine / 360.0f);
      b = (uint32_t)clampf_(rintf(phi * 253.0f / 360.0f + 2.0f), 2.0f, 255.0f);
    }
    const uint32_t w = b | (__shfl_down(b, 1) << 8) | (__shfl_down(b, 2) << 16) | (__shfl_down(b, 3) << 24);
    if ((lane & 3) == 0 && g < ng) dst[g >> 2] = w;  // ng is a multiple of 4: the whole word is in range
  }
}

// the geometry of listed gates (rg: n (radial, gate) pairs) as the cut kernel computes it: 7 floats per gate
// (x, y, ex, ey, zc, hl, hu)
__global__ void __launch_bounds__(64) k_radar_geom(RadarU R, int n, const int* rg, float* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const RadarCtx C{R, nullptr, Tex{nullptr, 0, 0, 0, 0}, Tex{nullptr, 0, 0, 0, 0}, Tex{nullptr, 0, 0, 0, 0},
                   Tex{nullptr, 0, 0, 0, 0}};
  const GateGeom g = gate_geom(C, rg[2 * q], rg[2 * q + 1]);
  float* o = out + (size_t)q * 7;
  o[0] = g.p.x; o[1] = g.p.y; o[2] = g.e.x; o[3] = g.e.y; o[4] = g.zc; o[5] = g.hl; o[6] = g.hu;
}

void launch_radar_geom(hipStream_t s, const RadarU& R, int n, const int* rg, float* out) {
  if (n > 0) hipLaunchKernelGGL(k_radar_geom, dim3((n + 63) / 64), dim3(64), 0, s, R, n, rg, out);
}

void launch_radar_cut(hipStream_t s, const RadarU& R, const RadarArgs& a) {
  const int nthreads = (int)R.cut.w * (int)R.gates.x;
  hipLaunchKernelGGL(k_radar_cut, dim3((nthreads + 63) / 64), dim3(64), 0, s, R, a);
  if (((uint32_t)R.offB.w) & 16u) hipLaunchKernelGGL(k_radar_phi, dim3((int)R.cut.w), dim3(64), 0, s, R, a);
}

}  // namespace obl
