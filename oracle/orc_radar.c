/*
 * CPU oracle of the virtual radar cuts (radar.js:41-268) -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 * FP32 throughout, built with -ffp-contract=off; the state is sampled trilinearly, clamp-to-edge, at
 * texel coordinate u*S - 1/2 with exact FP32 weights (the reference's linear sampler; the hardware
 * filter's sub-texel precision is implementation-defined). WGSL round() is half-to-even: rintf.
 *
 * Pins (tests/test_oracle_radar.py): the beam height against the 4/3-earth closed form and the
 * exact geometry of a ray; the great-circle position of a gate against an independent
 * spherical-trigonometry computation; echo-free gates stay 0; a uniform rain field gives its
 * Marshall-Palmer / power-law dBZ code; a uniform wind gives its radial velocity with aliasing;
 * PHIDP is the system offset plus twice the running sum of KDP.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

typedef struct { float x, y; } F2;
typedef struct { float x, y, z; } F3;
typedef struct { float x, y, z, w; } F4;
typedef struct { F4 site, geo, cut, gates, nyq, offA, offB, mdl, mdl2, phys, phys2; } RU;

static const float RE_ = 6371000.0f;
static float clampf_(float e, float lo, float hi) { float m = e > lo ? e : lo; return m < hi ? m : hi; }
static int clampi_(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static float sstep(float a, float b, float x) { float t = clampf_((x - a) / (b - a), 0.0f, 1.0f); return t * t * (3.0f - 2.0f * t); }
static float sq_(float x) { return x * x; }
static float db2lin(float d) { return powf(10.0f, 0.1f * d); }
static uint32_t code(float v, float sc, float ofs) { return (uint32_t)clampf_(rintf(v * sc + ofs), 2.0f, 255.0f); }

typedef struct {
  const float* p;
  int W, H, D, stride;
} Tex;
static float tat(const Tex* t, int x, int y, int z, int ch) {
  return t->p[(size_t)t->stride * ((size_t)x + (size_t)t->W * ((size_t)y + (size_t)t->H * (size_t)z)) + ch];
}
static float lerp_(float a, float b, float t) { return a + t * (b - a); }
static float tsamp(const Tex* t, float tx, float ty, float tz, int ch) {
  float fx = floorf(tx), fy = floorf(ty), fz = floorf(tz);
  float ax = tx - fx, ay = ty - fy, az = tz - fz;
  int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
  int x1 = clampi_(x0 + 1, 0, t->W - 1), y1 = clampi_(y0 + 1, 0, t->H - 1), z1 = clampi_(z0 + 1, 0, t->D - 1);
  x0 = clampi_(x0, 0, t->W - 1); y0 = clampi_(y0, 0, t->H - 1); z0 = clampi_(z0, 0, t->D - 1);
  float c00 = lerp_(tat(t, x0, y0, z0, ch), tat(t, x1, y0, z0, ch), ax), c10 = lerp_(tat(t, x0, y1, z0, ch), tat(t, x1, y1, z0, ch), ax);
  float c01 = lerp_(tat(t, x0, y0, z1, ch), tat(t, x1, y0, z1, ch), ax), c11 = lerp_(tat(t, x0, y1, z1, ch), tat(t, x1, y1, z1, ch), ax);
  return lerp_(lerp_(c00, c10, ay), lerp_(c01, c11, ay), az);
}
static float tsamp_uv(const Tex* t, F3 uc, int ch) {
  return tsamp(t, uc.x * (float)t->W - 0.5f, uc.y * (float)t->H - 0.5f, uc.z * (float)t->D - 0.5f, ch);
}

typedef struct {
  const RU* R;
  const float* BS;
  Tex V, Q1, Q2, aux;
} Ctx;
static F4 bsA_(const Ctx* c, int k) { const float* p = c->BS + clampi_(k, 0, (int)c->R->mdl.z - 1) * 8; F4 r = {p[0], p[1], p[2], p[3]}; return r; }
static F4 bsB_(const Ctx* c, int k) { const float* p = c->BS + clampi_(k, 0, (int)c->R->mdl.z - 1) * 8 + 4; F4 r = {p[0], p[1], p[2], p[3]}; return r; }

/* radar.js:57-69 */
static F2 gatePos(const Ctx* c, float a, float s) {
  const RU* R = c->R;
  float p1 = R->geo.z, l1 = R->geo.w, d = s / RE_;
  float sp2 = sinf(p1) * cosf(d) + cosf(p1) * sinf(d) * cosf(a);
  float p2 = asinf(clampf_(sp2, -1.0f, 1.0f));
  float l2 = l1 + atan2f(sinf(a) * sinf(d) * cosf(p1), cosf(d) - sinf(p1) * sp2);
  float p0 = R->geo.x, dl = l2 - R->geo.y;
  float cc = clampf_(sinf(p0) * sinf(p2) + cosf(p0) * cosf(p2) * cosf(dl), -1.0f, 1.0f);
  float cA = acosf(cc);
  float k = cA < 1e-6f ? 1.0f : cA / sinf(cA);
  F2 r = {RE_ * k * cosf(p2) * sinf(dl) - R->mdl2.z, RE_ * k * (cosf(p0) * sinf(p2) - sinf(p0) * cosf(p2) * cosf(dl)) - R->mdl2.w};
  return r;
}
/* radar.js:71-77 */
static F2 beam(const Ctx* c, float r, float el) {
  float a = c->R->site.w;
  float q = r * r + 2.0f * r * a * sinf(el);
  float h = q / (sqrtf(q + a * a) + a);
  float s = a * asinf(clampf_(r * cosf(el) / (a + h), -1.0f, 1.0f));
  F2 o = {h, s};
  return o;
}

/* a gate's geometry (radar.js:146-156): position, radial direction, the three beam heights */
typedef struct { F2 p, e; float zc, hl, hu; } Geom;
static Geom geom(const Ctx* c, int ri, int gi) {
  const RU* R = c->R;
  Geom g;
  float r = R->gates.z + R->gates.w * (float)gi;
  float el = R->cut.x, bw = 0.0165f;
  float az = (R->cut.y + R->cut.z * (float)ri) * 0.017453293f;
  F2 hc = beam(c, r, el);
  g.p = gatePos(c, az, hc.y);
  F2 p2 = gatePos(c, az, hc.y + 200.0f);
  g.e.x = p2.x - g.p.x; g.e.y = p2.y - g.p.y;
  float en = sqrtf(g.e.x * g.e.x + g.e.y * g.e.y);
  g.e.x /= en; g.e.y /= en;
  float zAnt = R->site.z - R->mdl2.y;
  g.zc = zAnt + hc.x;
  g.hl = beam(c, r, el - 0.4f * bw).x + zAnt;
  g.hu = beam(c, r, el + 0.4f * bw).x + zAnt;
  return g;
}

typedef struct { float zr, zs, zg, vt; F3 u; float T, qr, qg, rho, nu, inside, lr; } Smp;
/* radar.js:95-139 */
static Smp sample(const Ctx* c, F2 xy, float z) {
  const RU* R = c->R;
  Smp o;
  F3 n = {R->mdl.x, R->mdl.y, R->mdl.z};
  float dx = R->mdl.w, dz = R->mdl2.x;
  F3 g = {xy.x / dx, xy.y / dx, z / dz};
  o.inside = (g.x >= 0 && g.y >= 0 && g.z >= 0 && g.x <= n.x && g.y <= n.y && g.z <= n.z && z >= 0) ? 1.0f : 0.0f;
  F3 gc = {clampf_(g.x, 0, n.x), clampf_(g.y, 0, n.y), clampf_(g.z, 0, n.z)};
  F3 uc = {gc.x / n.x, gc.y / n.y, gc.z / n.z};
  float q1w = tsamp_uv(&c->Q1, uc, 3);
  float q2x = tsamp_uv(&c->Q2, uc, 0), q2y = tsamp_uv(&c->Q2, uc, 1), q2z = tsamp_uv(&c->Q2, uc, 2);
  F3 N1 = {n.x + 1.0f, n.y + 1.0f, n.z + 1.0f};
  F3 cu = {(gc.x + 0.5f) / N1.x, gc.y / N1.y, gc.z / N1.z}, cv = {gc.x / N1.x, (gc.y + 0.5f) / N1.y, gc.z / N1.z};
  F3 cw = {gc.x / N1.x, gc.y / N1.y, (gc.z + 0.5f) / N1.z}, ct = {gc.x / N1.x, gc.y / N1.y, gc.z / N1.z};
  float u = tsamp_uv(&c->V, cu, 0), v = tsamp_uv(&c->V, cv, 1), w = tsamp_uv(&c->V, cw, 2), th = tsamp_uv(&c->V, ct, 3);
  int k = (int)clampf_(floorf(gc.z), 0.0f, n.z - 1.0f);
  F4 A = bsA_(c, k), B = bsB_(c, k);
  float rho = A.z, T = th * B.x;
  float qr = fmaxf(q1w, 0), qs = fmaxf(q2x, 0), qg = fmaxf(q2y, 0);
  o.zr = 3.63e9f * powf(rho * qr, 1.75f);
  o.zs = 9.80e8f * powf(rho * qs, 1.75f) * 0.224f * 4.0f;
  o.zg = (T < 273.15f ? 1.2e10f : 4.33e10f) * powf(rho * qg, 1.75f);
  float sq = sqrtf(R->phys.x / rho);
  float vr = fminf(36.34f * powf(fmaxf(0.001f * rho * qr, 0.0f), 0.1364f) * sq, 10.0f * sq);
  float vs = clampf_(1.1f * powf(fmaxf(1000.0f * rho * qs, 1e-6f), 0.11f), 0.3f, 2.2f) * sq;
  float vg = clampf_(9.0f * powf(fmaxf(1000.0f * rho * qg, 1e-6f), 0.42f), 1.0f, 55.0f) * sq;
  float rr = qg / 0.009f;
  vg = fminf(vg * (1.0f + rr * rr), 65.0f);
  if (R->phys2.x > 0.5f && R->phys.z < 0.5f) {
    o.zg = (T < 273.15f ? 1.3e9f : 1.0e10f) * powf(rho * qg, 1.75f);
    vg = clampf_(5.0f * powf(fmaxf(1000.0f * rho * qg, 1e-6f), 0.25f), 1.0f, 9.0f) * sq;
  }
  o.lr = powf(3.14159f * 1000.0f * 8e6f / fmaxf(rho * qr, 1e-9f), 0.25f);
  if (R->phys.z > 0.5f) {
    float nr = fmaxf(q2z, 0.0f);
    if (qr > 1e-9f) {
      float l = clampf_(powf(3.14159f * 1000.0f * fmaxf(nr, 1e-3f) / qr, 0.33333f), 1000.0f, 1.0e5f);
      float nb = qr * l * l * l / 3141.59f;
      o.zr = 720e18f * rho * nb / powf(l, 6.0f);
      o.lr = l;
      vr = fminf(3949.0f * powf(l, -0.8f), 10.0f) * sq;
    } else {
      o.zr = 0.0f;
    }
    if (qg > 1e-9f) {
      float n0 = clampf_(200.0f / (rho * qg), 1e4f, 3e6f);
      float l = powf(3.14159f * 500.0f * n0 / (rho * qg), 0.25f);
      o.zg = 720e18f * n0 / powf(l, 7.0f) * (T < 273.15f ? 0.0666f : 0.24f);
      vg = fminf(2483.0f * powf(l, -0.89f), 30.0f) * sq;
    } else {
      o.zg = 0.0f;
    }
  }
  float zt = o.zr + o.zs + o.zg;
  o.vt = zt > 1e-6f ? (o.zr * vr + o.zs * vs + o.zg * vg) / zt : 0.0f;
  o.u.x = u; o.u.y = v; o.u.z = w;
  o.T = T; o.qr = qr; o.qg = qg; o.rho = rho;
  o.nu = tsamp_uv(&c->aux, uc, 0);
  return o;
}

/* a gate's codes and KDP; for orc_radar_vals also each moment's value before rounding and the margin of every
 * discrete decision the gate went through (how far its input lay from the switch) */
typedef struct {
  uint32_t zref, vel, sw, zdr, rho;
  float kdp;
  float val[5];  /* REF, VEL, SW, ZDR, RHO: v * scale + offset before rounding (NAN where not formed) */
  float mgeo;    /* m: distance of the beam centre from the model top and of s0 from the domain's edges */
  float mecho;   /* dB: dbz - threshold (the echo decision), or log10(ztot / 1e-6) * 10 below it */
  float mwrap;   /* m/s: distance of the radial velocity from a Nyquist wrap */
  float mT;      /* K: least |T - 273.15| over the three beam samples (the melting / graupel switches) */
  float mzg;     /* relative distance of zg from the 3.2e5 switch */
  float md0;     /* mm: distance of the median drop size from the 2.5 mm switch */
  float mq;      /* relative distance of qr and qg from the 1e-9 switches of the Thompson sizes */
} Gate;
static float relm(float v, float t) { return fabsf(v - t) / t; }
/* radar.js:144-214 */
/* gg: null, or the gate's geometry to use in place of its own (Geom's p, e, zc, hl, hu as 7 floats;
   orc_radar_gates) */
static Gate gate(const Ctx* c, const float* tor, int ri, int gi, const float* gg) {
  const RU* R = c->R;
  Gate G = {0, 0, 0, 0, 0, -1.0f, {NAN, NAN, NAN, NAN, NAN}, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY,
            INFINITY, INFINITY};
  float r = R->gates.z + R->gates.w * (float)gi;
  float el = R->cut.x;
  float bw = 0.0165f;
  Geom gm = geom(c, ri, gi);
  if (gg) {
    gm.p.x = gg[0]; gm.p.y = gg[1]; gm.e.x = gg[2]; gm.e.y = gg[3]; gm.zc = gg[4]; gm.hl = gg[5]; gm.hu = gg[6];
  }
  F2 p = gm.p, e = gm.e;
  float zc = gm.zc;
  G.mgeo = fabsf(R->mdl.z * R->mdl2.x - 200.0f - zc);
  {
    float ex = fminf(fminf(p.x, R->mdl.x * R->mdl.w - p.x), fminf(p.y, R->mdl.y * R->mdl.w - p.y));
    G.mgeo = fminf(G.mgeo, fabsf(ex));
  }
  if (zc > R->mdl.z * R->mdl2.x - 200.0f) return G;
  float hl = gm.hl, hu = gm.hu;
  F2 pp = p;
  Smp s0 = sample(c, pp, fmaxf(zc, 5.0f)), s1 = sample(c, pp, fmaxf(hl, 5.0f)), s2 = sample(c, pp, fmaxf(hu, 5.0f));
  if (s0.inside < 0.5f) return G;
  G.mT = fminf(fabsf(s0.T - 273.15f), fminf(fabsf(s1.T - 273.15f), fabsf(s2.T - 273.15f)));
  G.mq = fminf(fminf(relm(s0.qr, 1e-9f), relm(s1.qr, 1e-9f)), relm(s2.qr, 1e-9f));
  G.mq = fminf(G.mq, fminf(fminf(relm(s0.qg, 1e-9f), relm(s1.qg, 1e-9f)), relm(s2.qg, 1e-9f)));
  float zr = 0.5f * s0.zr + 0.25f * (s1.zr + s2.zr), zs = 0.5f * s0.zs + 0.25f * (s1.zs + s2.zs);
  float zg = 0.5f * s0.zg + 0.25f * (s1.zg + s2.zg);
  float zmet = zr + zs + zg;
  float zdeb = 0.0f, swt = 0.0f, dvx = 0.0f, dvy = 0.0f;
  float bwm = 0.5f * r * bw;
  for (int q = 0; q < (int)R->phys.w; q++) {
    const float* A = tor + 8 * q;
    const float* B = A + 4;
    float dxx = p.x - A[0], dyy = p.y - A[1];
    float rt = sqrtf(dxx * dxx + dyy * dyy);
    float Re = sqrtf(A[2] * A[2] + bwm * bwm + 125.0f * 125.0f);
    if (rt > 8.0f * Re) continue;
    float vmx = A[3] * A[2] / Re;
    float vt = vmx * (rt < Re ? rt / Re : Re / fmaxf(rt, 1.0f));
    float etx = -dyy / fmaxf(rt, 1.0f), ety = dxx / fmaxf(rt, 1.0f), erx = dxx / fmaxf(rt, 1.0f), ery = dyy / fmaxf(rt, 1.0f);
    float fz = 1.0f - sstep(1200.0f, 3500.0f, zc);
    dvx += (vt * etx - 0.25f * vt * erx) * fz;
    dvy += (vt * ety - 0.25f * vt * ery) * fz;
    swt = fmaxf(swt, vmx * 0.4f * expf(-rt * rt / (4.0f * Re * Re)) * fz);
    if (B[2] > 0.0f && zc < B[0]) {
      float rd = fmaxf(1.4f * A[2], 150.0f);
      zdeb += db2lin(B[1]) * fminf(1.0f, rd * rd / (Re * Re)) * expf(-rt * rt / (Re * Re)) * (1.0f - sstep(0.55f * B[0], B[0], zc)) * B[2];
    }
  }
  float ztot = zmet + zdeb;
  G.mecho = ztot > 0.0f ? fabsf(10.0f * log10f(ztot / 1e-6f)) : INFINITY;
  if (ztot <= 1e-6f) return G;
  float dbz = 10.0f * logf(ztot) / 2.302585f;
  G.mecho = fminf(G.mecho, fabsf(dbz - R->phys.y));
  if (dbz < R->phys.y) return G;
  G.zref = code(dbz, 2.0f, 66.0f);
  G.val[0] = dbz * 2.0f + 66.0f;
  if (gi >= (int)R->gates.y) return G;
  float ce = cosf(el), se = sinf(el);
  float vtf = s0.vt * zmet / ztot;
  float vr = ((s0.u.x + dvx) * e.x + (s0.u.y + dvy) * e.y) * ce + (s0.u.z - vtf) * se;
  float nyq = R->nyq.x;
  {
    float w = (vr + nyq) / (2.0f * nyq);
    G.mwrap = fabsf(w - rintf(w)) * 2.0f * nyq;
  }
  vr = vr - 2.0f * nyq * floorf((vr + nyq) / (2.0f * nyq));
  G.vel = code(vr, 2.0f, 129.0f);
  G.val[1] = vr * 2.0f + 129.0f;
  float vru = (s2.u.x * e.x + s2.u.y * e.y) * ce, vrl = (s1.u.x * e.x + s1.u.y * e.y) * ce;
  float swv = sqrtf(1.0f + sq_(clampf_(2.5f * sqrtf(fmaxf(s0.nu, 0.0f) / 1000.0f), 0.0f, 8.0f)) + sq_((vru - vrl) / 2.5f) +
                    sq_(0.15f * s0.vt * se) + sq_(swt));
  G.sw = code(fminf(swv, 19.5f), 2.0f, 129.0f);
  G.val[2] = fminf(swv, 19.5f) * 2.0f + 129.0f;
  float d0 = 3.67f / s0.lr * 1000.0f;
  int melt = s0.T > 273.15f;
  float zdrR = clampf_(0.2f + 1.15f * (d0 - 0.5f), 0.1f, 4.8f);
  float zdrS = melt ? 1.3f : 0.3f;
  float zdrG = melt ? (zg > 3.2e5f ? 0.0f : 0.7f) : 0.15f;
  float zv = zr / db2lin(zdrR) + zs / db2lin(zdrS) + zg / db2lin(zdrG) + zdeb / db2lin(0.2f);
  float zdr = 10.0f * logf(fmaxf(ztot, 1e-9f) / fmaxf(zv, 1e-9f)) / 2.302585f;
  G.zdr = code(clampf_(zdr, -7.9f, 7.9f), 16.0f, 128.0f);
  G.val[3] = clampf_(zdr, -7.9f, 7.9f) * 16.0f + 128.0f;
  G.mzg = relm(zg, 3.2e5f);
  G.md0 = fabsf(d0 - 2.5f);
  float fr = zr / ztot, fs = zs / ztot, fg = zg / ztot, fd = zdeb / ztot;
  float cc = fr * (d0 > 2.5f ? 0.986f : 0.995f) + fs * (melt ? 0.93f : 0.985f) + fg * (melt ? (zg > 3.2e5f ? 0.90f : 0.94f) : 0.985f) + fd * 0.55f;
  cc -= 0.24f * fr * fg;
  G.rho = code(clampf_(cc, 0.2f, 1.05f), 300.0f, -60.5f);
  G.val[4] = clampf_(cc, 0.2f, 1.05f) * 300.0f - 60.5f;
  G.kdp = powf(0.017f * powf(fmaxf(zr, 0.0f), 0.714f) / 44.0f, 1.2165f);
  return G;
}

/* one cut (both passes, radar.js:215-240 and 243-267): ru = the 44-float RADAR_FIELDS block; out
   and kdp as the GPU's buffers */
void orc_radar_cut(const float* ru, const float* BS, const float* V, const float* Q1, const float* Q2, const float* aux,
                   const float* tor, uint32_t* out, float* kdp) {
  RU R;
  memcpy(&R, ru, sizeof(R));
  int nx = (int)R.mdl.x, ny = (int)R.mdl.y, nz = (int)R.mdl.z;
  Ctx C = {&R, BS, {V, nx + 1, ny + 1, nz + 1, 4}, {Q1, nx, ny, nz, 4}, {Q2, nx, ny, nz, 4}, {aux, nx, ny, nz, 2}};
  int nq = (int)R.gates.x / 4, nrad = (int)R.cut.w, ng = (int)R.gates.y, ndq = ng / 4;
  unsigned mask = (unsigned)R.offB.w;
#pragma omp parallel for schedule(dynamic, 4)
  for (int ri = 0; ri < nrad; ri++)
    for (int q = 0; q < nq; q++) {
      uint32_t pr = 0, pv = 0, ps = 0, pz = 0, ph = 0;
      for (int k = 0; k < 4; k++) {
        int gi = q * 4 + k;
        Gate G = gate(&C, tor, ri, gi, NULL);
        pr |= G.zref << (8u * k); pv |= G.vel << (8u * k); ps |= G.sw << (8u * k);
        pz |= G.zdr << (8u * k); ph |= G.rho << (8u * k);
        if (gi < ng) kdp[ri * ng + gi] = G.kdp;
      }
      out[(uint32_t)R.offA.x + (uint32_t)(ri * nq + q)] = pr;
      if (q < ndq) {
        uint32_t o = (uint32_t)(ri * ndq + q);
        if (mask & 2u) out[(uint32_t)R.offA.y + o] = pv;
        if (mask & 4u) out[(uint32_t)R.offA.z + o] = ps;
        if (mask & 8u) out[(uint32_t)R.offA.w + o] = pz;
        if (mask & 32u) out[(uint32_t)R.offB.y + o] = ph;
      }
    }
  if (!(mask & 16u)) return;
  /* radar.js:243-267 */
  for (int ri = 0; ri < nrad; ri++) {
    float acc = R.nyq.w;
    for (int q = 0; q < ng / 4; q++) {
      uint32_t w = 0;
      for (int k = 0; k < 4; k++) {
        int gi = q * 4 + k;
        float kd = kdp[ri * ng + gi];
        if (kd >= 0.0f) {
          acc += 2.0f * kd * R.gates.w * 0.001f;
          float phi = acc - 360.0f * floorf(acc / 360.0f);
          w |= (uint32_t)clampf_(rintf(phi * 253.0f / 360.0f + 2.0f), 2.0f, 255.0f) << (8u * k);
        }
      }
      out[(uint32_t)R.offB.x + (uint32_t)(ri * (ng / 4) + q)] = w;
    }
  }
}

/* the geometry alone, for the pins: beam (h, s) and gate position (x, y) of (ri, gi) */
void orc_radar_geom(const float* ru, int ri, int gi, float* hs_xy) {
  RU R;
  memcpy(&R, ru, sizeof(R));
  Ctx C;
  memset(&C, 0, sizeof(C));
  C.R = &R;
  float r = R.gates.z + R.gates.w * (float)gi;
  float az = (R.cut.y + R.cut.z * (float)ri) * 0.017453293f;
  F2 hc = beam(&C, r, R.cut.x);
  F2 p = gatePos(&C, az, hc.y);
  hs_xy[0] = hc.x; hs_xy[1] = hc.y; hs_xy[2] = p.x; hs_xy[3] = p.y;
}

#define ORC_NGATEV 26
/* listed gates' values and decision margins, for the test that names the gates whose codes may differ between the
   GPU and this oracle. rg = n (radial, gate) pairs; gg = null, or the GPU's geometry of each listed gate (7 floats:
   x, y, ex, ey, zc, hl, hu). out[n][26] = val[0..4] (Gate.val), mgeo, mecho, mwrap, mT, mzg, md0, mq, then the
   five values again at the GPU's geometry (NaN without gg), then this oracle's own geometry (7 floats), then the
   gate's KDP (the PHI pass's input, -1 where no echo) at its own geometry and at the GPU's (NaN without gg). */
void orc_radar_gates(const float* ru, const float* BS, const float* V, const float* Q1, const float* Q2,
                     const float* aux, const float* tor, int n, const int* rg, const float* gg, float* out) {
  RU R;
  memcpy(&R, ru, sizeof(R));
  int nx = (int)R.mdl.x, ny = (int)R.mdl.y, nz = (int)R.mdl.z;
  Ctx C = {&R, BS, {V, nx + 1, ny + 1, nz + 1, 4}, {Q1, nx, ny, nz, 4}, {Q2, nx, ny, nz, 4}, {aux, nx, ny, nz, 2}};
#pragma omp parallel for schedule(dynamic, 16)
  for (int q = 0; q < n; q++) {
    int ri = rg[2 * q], gi = rg[2 * q + 1];
    float* o = out + (size_t)q * ORC_NGATEV;
    Gate G = gate(&C, tor, ri, gi, NULL);
    for (int k = 0; k < 5; k++) o[k] = G.val[k];
    o[5] = G.mgeo; o[6] = G.mecho; o[7] = G.mwrap; o[8] = G.mT; o[9] = G.mzg; o[10] = G.md0; o[11] = G.mq;
    for (int k = 0; k < 5; k++) o[12 + k] = NAN;
    o[24] = G.kdp;
    o[25] = NAN;
    if (gg) {
      Gate H = gate(&C, tor, ri, gi, gg + (size_t)q * 7);
      for (int k = 0; k < 5; k++) o[12 + k] = H.val[k];
      o[25] = H.kdp;
    }
    Geom g = geom(&C, ri, gi);
    o[17] = g.p.x; o[18] = g.p.y; o[19] = g.e.x; o[20] = g.e.y; o[21] = g.zc; o[22] = g.hl; o[23] = g.hu;
  }
}
