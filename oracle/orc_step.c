/*
 * CPU oracle of the model step kernels -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Compiled twice: -DORC_F64 gives the FP64 oracle (suffix _f64), otherwise the
 * FP32 restatement of the reference arithmetic (suffix _f32).  Build with
 * -ffp-contract=off so that FP32 products and sums round as written.
 *
 * Each function follows the WGSL kernel cited beside it line by line:
 * the same texel conventions (clamped textureLoad; trilinear, clamp-to-edge
 * samples at texel coordinate u*S - 1/2), the same evaluation order of sums,
 * the same branches.  The hardware texture filter's sub-texel weight precision
 * and WGSL's transcendental ULPs are implementation-defined; here weights are
 * exact FP32 and exp/pow/sin are the C library's.
 *
 * Parity pins (tests/test_oracle_*.py; each would fail on a plausible transcription mistake):
 *   hash/rnd3      integer-exact against the published PCG-RXS-M-XS hash
 *   lsf/ci         closed forms for no boundaries / no events, a straight front and one tower;
 *                  phT the downward trapezoid integral of the target buoyancy
 *   init           target state + bounded noise, qv <= 0.95 qvs
 *   nut            zero for uniform flow; linear shear gives the closed-form nu
 *   adv*           identity at zero wind; constants preserved; integer-Courant
 *                  shifts exact; no new extrema; brute-force trilinear
 *   phys forces    the exact Laplacians of quadratics (Prandtl 1/3 for theta and vapour); lateral and
 *                  top sponge rates at known positions and condensate removal; air-mass nudging to
 *                  the smoothstep target profile; f-plane Coriolis on the departure wind; the target
 *                  phi gradient; semi-slip log-law drag; buoyancy of theta', qv' and each condensate;
 *                  the one-sided CI push; speed cap and NaN guard
 *   fall speeds    Klemp-Wilhelmson rain; Thompson rain and graupel slopes and fall speeds as moments
 *                  of the exponential size distribution (gamma functions); power-law exponents,
 *                  coefficients and bounds of snow, Lin and maritime graupel; Heymsfield-Donner ice
 *   microphysics   each process isolated on hand-built points: Kessler autoconversion and accretion,
 *                  new-drop number, Thompson accretion by its size distribution, Lin ice
 *                  autoconversion, collection and riming, rain freezing (Bigg) and collection by
 *                  graupel, snow-to-graupel, melting (Lin and Thompson Lin-1983 form with drop
 *                  numbers), Klemp-Wilhelmson and Thompson evaporation, Seifert-Beheng
 *                  self-collection and breakup with the drop bounds, sublimation, mixed-phase
 *                  saturation adjustment; total-water conservation and positivity throughout
 *   div/proj       div(proj(V)) ~ 0 with the solved phi
 *   maps           zero wind gives zero vorticity, helicity, convergence; solid-body rotation
 *                  gives zeta = 2 Omega; a uniform divergence gives its convergence; a cloud-free
 *                  column's brightness temperature is the lowest level's; one opaque layer's is its
 *                  own; the Lin rain reflectivity law in closed form
 *   solve          scipy DST-II, manufactured solution, sparse LU
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#ifdef ORC_F64
typedef double R;
#define S_(n) n##_f64
#define EXP exp
#define POW pow
#define SQRT sqrt
#define FLOOR floor
#define SIN sin
#define COS cos
#define ATAN2 atan2
#define FABS fabs
#define LOG log
#else
typedef float R;
#define S_(n) n##_f32
#define EXP expf
#define POW powf
#define SQRT sqrtf
#define FLOOR floorf
#define SIN sinf
#define COS cosf
#define ATAN2 atan2f
#define FABS fabsf
#define LOG logf
#endif

typedef struct { R x, y, z, w; } R4;
typedef struct { R x, y; } R2;
typedef struct { int x, y, z; } I3;

/* ---------------- uniform block (MODEL_FIELDS, model.js:23-37) ---------------- */
enum {
  U_DIMS = 0, U_DZV = 4, U_STEP = 8, U_SPONGE = 12, U_TURB = 16, U_MICRO = 20, U_MISC = 24, U_COR = 28,
  U_CI = 32, U_PAD = 36, U_MICRO2 = 40, U_HUR = 44, U_HUR2 = 48, U_HUR3 = 52, U_HUR4 = 56, U_HUR5 = 60,
  U_VTX = 64, U_VTX2 = 68, U_CN = 72
};

typedef struct {
  const float* U;
  const float* BS;
  int nx, ny, nz;
  R dx, dz, dt;
} Ctx;

static Ctx mkctx(const float* U, const float* BS) {
  Ctx c;
  c.U = U; c.BS = BS;
  c.nx = (int)U[U_DIMS]; c.ny = (int)U[U_DIMS + 1]; c.nz = (int)U[U_DIMS + 2];
  c.dx = (R)U[U_DIMS + 3]; c.dz = (R)U[U_DZV]; c.dt = (R)U[U_STEP];
  return c;
}
#define UF(c, i) ((R)(c)->U[i])

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline R minr(R a, R b) { return a < b ? a : b; }
static inline R maxr(R a, R b) { return a > b ? a : b; }
/* WGSL clamp(e, lo, hi) = min(max(e, lo), hi) */
static inline R clampr(R e, R lo, R hi) { return minr(maxr(e, lo), hi); }
/* WGSL mix(a, b, t) = a (1 - t) + b t */
static inline R mixr(R a, R b, R t) { return a * ((R)1 - t) + b * t; }
static inline R fractr(R x) { return x - FLOOR(x); }
static R sstep(R a, R b, R x) { R t = clampr((x - a) / (b - a), 0, 1); return t * t * ((R)3 - (R)2 * t); }

/* BS table: per level k, A = (theta0, qv0, rho_c, rho_w bottom), B = (pi0, p0, u0, v0) (model.js:47-53) */
static inline R4 bsA(const Ctx* c, int k) {
  const float* p = c->BS + clampi(k, 0, c->nz - 1) * 8;
  R4 r = {p[0], p[1], p[2], p[3]};
  return r;
}
static inline R4 bsB(const Ctx* c, int k) {
  const float* p = c->BS + clampi(k, 0, c->nz - 1) * 8 + 4;
  R4 r = {p[0], p[1], p[2], p[3]};
  return r;
}
static inline R rhoW(const Ctx* c, int k) { return (R)c->BS[clampi(k, 0, c->nz) * 8 + 3]; }

/* ---------------- field access ---------------- */
static inline size_t vidx(const Ctx* c, int x, int y, int z) {
  return (size_t)x + (size_t)(c->nx + 1) * ((size_t)y + (size_t)(c->ny + 1) * (size_t)z);
}
static inline size_t cidx(const Ctx* c, int x, int y, int z) {
  return (size_t)x + (size_t)c->nx * ((size_t)y + (size_t)c->ny * (size_t)z);
}
static inline R4 get4(const R* a, size_t i) { R4 r = {a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]}; return r; }
static inline void put4(R* a, size_t i, R4 v) { a[4 * i] = v.x; a[4 * i + 1] = v.y; a[4 * i + 2] = v.z; a[4 * i + 3] = v.w; }
static inline R ch4(R4 v, int ch) { return ch == 0 ? v.x : ch == 1 ? v.y : ch == 2 ? v.z : v.w; }

/* textureLoad(V, clamp(c, 0, N)) on the (n+1)^3 face grid */
static inline R4 ldV(const Ctx* c, const R* V, int x, int y, int z) {
  return get4(V, vidx(c, clampi(x, 0, c->nx), clampi(y, 0, c->ny), clampi(z, 0, c->nz)));
}
/* textureLoad(Q, clamp(c, 0, N-1)) on the cell grid */
static inline R4 ldQ(const Ctx* c, const R* Q, int x, int y, int z) {
  return get4(Q, cidx(c, clampi(x, 0, c->nx - 1), clampi(y, 0, c->ny - 1), clampi(z, 0, c->nz - 1)));
}
static inline R4 ld2d(const Ctx* c, const R* M, int i, int j) {
  i = clampi(i, 0, c->nx - 1); j = clampi(j, 0, c->ny - 1);
  return get4(M, (size_t)i + (size_t)c->nx * (size_t)j);
}

/* Trilinear sample, clamp-to-edge, of a W x H x D float4 texture at texel
 * coordinates (tx, ty, tz) = u*S - 1/2 (textureSampleLevel with the linear
 * sampler of model.js:1167). Lerps x, then y, then z as a + t (b - a). */
static inline R4 lerp4(R4 a, R4 b, R t) {
  R4 r = {a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z), a.w + t * (b.w - a.w)};
  return r;
}
static R4 sample3(const R* T, int W, int H, int D, R tx, R ty, R tz) {
  R fx = FLOOR(tx), fy = FLOOR(ty), fz = FLOOR(tz);
  R ax = tx - fx, ay = ty - fy, az = tz - fz;
  int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
  int x1 = clampi(x0 + 1, 0, W - 1), y1 = clampi(y0 + 1, 0, H - 1), z1 = clampi(z0 + 1, 0, D - 1);
  x0 = clampi(x0, 0, W - 1); y0 = clampi(y0, 0, H - 1); z0 = clampi(z0, 0, D - 1);
#define TX(x, y, z) get4(T, (size_t)(x) + (size_t)W * ((size_t)(y) + (size_t)H * (size_t)(z)))
  R4 c00 = lerp4(TX(x0, y0, z0), TX(x1, y0, z0), ax);
  R4 c10 = lerp4(TX(x0, y1, z0), TX(x1, y1, z0), ax);
  R4 c01 = lerp4(TX(x0, y0, z1), TX(x1, y0, z1), ax);
  R4 c11 = lerp4(TX(x0, y1, z1), TX(x1, y1, z1), ax);
#undef TX
  return lerp4(lerp4(c00, c10, ay), lerp4(c01, c11, ay), az);
}

/* ---------------- noise (noise.js:5-30) ---------------- */
static uint32_t pcg_(uint32_t v) {
  uint32_t s = v * 747796405u + 2891336453u;
  uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (w >> 22u) ^ w;
}
static uint32_t hash2i_(int32_t x, int32_t y) { return pcg_((uint32_t)x ^ pcg_((uint32_t)y + 0x9e3779b9u)); }
static uint32_t hash3i_(int32_t x, int32_t y, int32_t z) {
  return pcg_((uint32_t)x ^ pcg_((uint32_t)y ^ pcg_((uint32_t)z + 0x7f4a7c15u)));
}
#ifndef ORC_F64
uint32_t orc_pcg(uint32_t v) { return pcg_(v); }
uint32_t orc_hash3i(int32_t x, int32_t y, int32_t z) { return hash3i_(x, y, z); }
#endif
static R rnd3(int x, int y, int z) { return (R)(hash3i_(x, y, z) >> 8u) * (R)(1.0 / 16777216.0); }
static R2 grad2(int x, int y) {
  R a = (R)(hash2i_(x, y) >> 8u) * (R)(6.2831853 / 16777216.0);
  R2 g = {COS(a), SIN(a)};
  return g;
}
static R fade(R t) { return t * t * t * (t * (t * (R)6 - (R)15) + (R)10); }
static R gnoise2(R px, R py) {
  R fx0 = FLOOR(px), fy0 = FLOOR(py);
  int ix = (int)fx0, iy = (int)fy0;
  R fx = fractr(px), fy = fractr(py);
  R ux = fade(fx), uy = fade(fy);
  R2 g;
  g = grad2(ix, iy);         R a = g.x * fx + g.y * fy;
  g = grad2(ix + 1, iy);     R b = g.x * (fx - 1) + g.y * fy;
  g = grad2(ix, iy + 1);     R c = g.x * fx + g.y * (fy - 1);
  g = grad2(ix + 1, iy + 1); R d = g.x * (fx - 1) + g.y * (fy - 1);
  return mixr(mixr(a, b, ux), mixr(c, d, ux), uy) * (R)1.4142;
}
static R fbm2(R px, R py, int oct) {
  R a = (R)0.5, s = 0;
  for (int i = 0; i < oct; i++) {
    s += a * gnoise2(px, py);
    /* mat2x2f(1.6, 1.2, -1.2, 1.6) * p: columns (1.6, 1.2) and (-1.2, 1.6) */
    R nx = (R)1.6 * px + (R)-1.2 * py + (R)17.1;
    R ny = (R)1.2 * px + (R)1.6 * py + (R)3.7;
    px = nx; py = ny;
    a *= (R)0.5;
  }
  return s;
}

/* ---------------- thermodynamics and fall speeds (model.js:59-106) ---------------- */
#define T0K ((R)273.15)
#define CPD ((R)1005.7)
#define RVC ((R)461.5)
#define LVC ((R)2.501e6)
#define LFC ((R)3.337e5)
#define LSC ((R)2.834e6)
static R esl(R T) { return (R)611.2 * EXP((R)17.67 * (T - T0K) / (T - (R)29.65)); }
static R esi(R T) { return (R)611.2 * EXP((R)21.8745584 * (T - (R)273.16) / (T - (R)7.66)); }
static R fIce(R T) { return clampr((T0K - T) / (R)40, 0, 1); }
static R qvsat(R T, R p, R f) { R e = mixr(esl(T), esi(T), f); return (R)0.622 * e / maxr(p - e, (R)100); }
static int thompson(const Ctx* c) { return c->U[U_MICRO2 + 2] > 0.5f; }
static R vtR(const Ctx* c, R q, R rho) {
  R f = SQRT(UF(c, U_STEP + 2) / rho);
  return minr((R)36.34 * POW(maxr((R)0.001 * rho * q, 0), (R)0.1364) * f, (R)10 * f);
}
static R vtS(const Ctx* c, R q, R rho) {
  return clampr((R)1.1 * POW(maxr((R)1000 * rho * q, (R)1e-6), (R)0.11), (R)0.3, (R)2.2) * SQRT(UF(c, U_STEP + 2) / rho);
}
static R vtG(const Ctx* c, R q, R rho) {
  R v = clampr((R)9 * POW(maxr((R)1000 * rho * q, (R)1e-6), (R)0.42), (R)1, (R)55) * SQRT(UF(c, U_STEP + 2) / rho);
  R r = q / (R)0.009;
  return minr(v * ((R)1 + r * r), (R)65);
}
static R lamR(R q, R n) {
  return clampr(POW((R)3.14159 * (R)1000 * maxr(n, (R)1e-3) / maxr(q, (R)1e-12), (R)0.33333), (R)1000, (R)1.0e5);
}
static R2 vtRT(const Ctx* c, R q, R n, R rho) {
  R l = POW(lamR(q, n), (R)-0.8);
  R s = SQRT(UF(c, U_STEP + 2) / rho);
  R2 v = {minr((R)2503 * l, (R)10) * s, minr((R)784 * l, (R)10) * s};
  return v;
}
static R n0G(R q, R rho) { return clampr((R)200 / maxr(rho * q, (R)1e-9), (R)1e4, (R)3e6); }
static R lamG(R q, R rho) { return POW((R)3.14159 * (R)500 * n0G(q, rho) / maxr(rho * q, (R)1e-12), (R)0.25); }
static R vtGT(const Ctx* c, R q, R rho) {
  return minr((R)1503 * POW(lamG(q, rho), (R)-0.89), (R)25) * SQRT(UF(c, U_STEP + 2) / rho);
}
static R2 vtRain(const Ctx* c, R qr, R nr, R rho) {
  if (thompson(c)) return vtRT(c, qr, nr, rho);
  R v = vtR(c, qr, rho);
  R2 r = {v, v};
  return r;
}
/* outbreak mode: maritime() is false (hur3.w only set in hurricane mode) */
/* hurricane mode (model.js:88-119): a warm sea under the domain */
static int hurOn(const Ctx* c) { return c->U[U_HUR] > 0.5f; }
static int maritime(const Ctx* c) { return c->U[U_HUR3 + 3] > 0.5f && !thompson(c); }
static R vtGM(const Ctx* c, R q, R rho) {
  return clampr((R)5 * POW(maxr((R)1000 * rho * q, (R)1e-6), (R)0.25), 1, 9) * SQRT(UF(c, U_STEP + 2) / rho);
}
static R vtGrau(const Ctx* c, R qg, R rho) {
  if (maritime(c)) return vtGM(c, qg, rho);
  return thompson(c) ? vtGT(c, qg, rho) : vtG(c, qg, rho);
}
static R vtI(const Ctx* c, R q, R rho) {
  return (hurOn(c) && q > (R)1e-9) ? minr((R)3.29 * POW(maxr(rho * q, (R)1e-12), (R)0.16), (R)1.5) : (R)0;
}
static R cdSfc(const Ctx* c, R u1, R land) {
  R f = UF(c, U_HUR2 + 3);
  R cs = minr(((R)0.75 + (R)0.067 * f * u1) * (R)1e-3, UF(c, U_HUR5)) * f * f;
  return mixr(cs, UF(c, U_HUR3), clampr(land, 0, 1));
}
/* the sea surface map (land fraction, SST K, 0, 0) per column; NULL = no sea (outbreak mode) */
static R sfcAt(const Ctx* c, const float* sfc, int i, int j, int ch) {
  if (!sfc) return 0;
  return (R)sfc[4 * ((size_t)clampi(i, 0, c->nx - 1) + (size_t)c->nx * clampi(j, 0, c->ny - 1)) + ch];
}

/* The fall speeds and size-distribution slopes of one point (model.js:59-106), for their pins:
 * out = (vtR, vtS, vtG, vtRT mass-weighted, vtRT number-weighted, vtGT, vtGM, vtI, lamR, lamG, n0G);
 * q is the mixing ratio of the species (kg/kg), n Thompson's rain number (1/kg), rho the air density. */
void S_(orc_fall)(const float* U, double q, double n, double rho, double* out) {
  Ctx C = mkctx(U, NULL), *c = &C;
  R q_ = (R)q, n_ = (R)n, r_ = (R)rho;
  R2 rt = vtRT(c, q_, n_, r_);
  out[0] = vtR(c, q_, r_); out[1] = vtS(c, q_, r_); out[2] = vtG(c, q_, r_);
  out[3] = rt.x; out[4] = rt.y; out[5] = vtGT(c, q_, r_); out[6] = vtGM(c, q_, r_); out[7] = vtI(c, q_, r_);
  out[8] = lamR(q_, n_); out[9] = lamG(q_, r_); out[10] = n0G(q_, r_);
}

/* ---------------- large-scale target (model.js:125-134) ---------------- */
static R tprof(R z, R H) { return (R)1 - sstep((R)0.7 * H, (R)1.15 * H, z); }
static R tgtTh(const Ctx* c, R4 A, int k) {
  R z = ((R)k + (R)0.5) * c->dz;
  return bsA(c, k).x + A.x / bsB(c, k).x * tprof(z, A.y);
}
static R tgtQv(const Ctx* c, R4 A, int k) {
  R z = ((R)k + (R)0.5) * c->dz;
  return bsA(c, k).y * ((R)1 - A.z * tprof(z, A.w));
}
static R tgtU(const Ctx* c, const R* lsfB, int x, int y, int z) {
  R zz = ((R)z + (R)0.5) * c->dz;
  R4 a = ld2d(c, lsfB, x - 1, y), b = ld2d(c, lsfB, x, y);
  return bsB(c, z).z + (R)0.5 * (a.x * tprof(zz, a.z) + b.x * tprof(zz, b.z));
}
static R tgtV(const Ctx* c, const R* lsfB, int x, int y, int z) {
  R zz = ((R)z + (R)0.5) * c->dz;
  R4 a = ld2d(c, lsfB, x, y - 1), b = ld2d(c, lsfB, x, y);
  return bsB(c, z).w + (R)0.5 * (a.y * tprof(zz, a.z) + b.y * tprof(zz, b.z));
}

/* ====================== LSF (model.js:142-244) ====================== */
/* boundary record: 148 floats = p0..p4 (5 vec4) + pts[32] (vec4) */
static void bpt(const Ctx* c, const float* b, int i, R* x, R* y) {
  const float* q = b + 20 + 4 * i;
  R t = UF(c, U_STEP + 1);
  *x = (R)q[0] + (R)q[2] * t;
  *y = (R)q[1] + (R)q[3] * t;
}
static void sdist(const Ctx* c, const float* b, R px, R py, R* sd, R* over_, R* along_) {
  int np = (int)b[0];
  R best = (R)1e30, sg = 1, over = 0, along = 0, s0 = 0;
  for (int i = 0; i < np - 1; i++) {
    R ax, ay, ex, ey;
    bpt(c, b, i, &ax, &ay);
    bpt(c, b, i + 1, &ex, &ey);
    R abx = ex - ax, aby = ey - ay;
    R la = maxr(SQRT(abx * abx + aby * aby), (R)1);
    R t0 = ((px - ax) * abx + (py - ay) * aby) / (la * la);
    R t = t0;
    if (i > 0) t = maxr(t, 0);
    if (i < np - 2) t = minr(t, 1);
    R qx = px - (ax + t * abx), qy = py - (ay + t * aby);
    R d = SQRT(qx * qx + qy * qy);
    if (d < best) {
      best = d;
      sg = (abx * (py - ay) - aby * (px - ax) > 0) ? (R)1 : (R)-1;
      over = (i == 0 ? maxr(-t0, 0) * la : (R)0) + (i == np - 2 ? maxr(t0 - (R)1, 0) * la : (R)0);
      along = s0 + t * la;
    }
    s0 += la;
  }
  *sd = best * sg * (R)b[15];
  *over_ = over;
  *along_ = along;
}
static R alongVar(const Ctx* c, const float* b, R s) {
  const float* P4 = b + 16;
  if (P4[0] <= 0.0f) return 0;
  return clampr(fbm2(s / (R)P4[2], (R)P4[1] + UF(c, U_STEP + 1) / (R)5400, 2) * (R)1.6, -1, 1);
}
static R hotSpots(const Ctx* c, const float* b, R s) {
  R L = (R)b[18];
  int sd = (int)((R)b[17] * (R)1000);
  R ph = UF(c, U_STEP + 1) / (R)4800;
  int e = (int)FLOOR(ph);
  R fe = sstep(0, 1, fractr(ph));
  int k0 = (int)FLOOR(s / L);
  R acc = 0;
  for (int k = k0 - 1; k <= k0 + 1; k++) {
    for (int q = 0; q < 2; q++) {
      uint32_t h = hash3i_(k, e + q, sd);
      R r1 = (R)(h & 0xffffu) / (R)65535;
      R r2 = (R)((h >> 16u) & 0xffffu) / (R)65535;
      R z = (s - ((R)k + (R)0.5 + (R)0.7 * (r1 - (R)0.5)) * L) / ((R)0.17 * L);
      acc += (q == 1 ? fe : (R)1 - fe) * ((R)0.35 + (R)0.65 * r2) * EXP(-z * z);
    }
  }
  return clampr(acc, 0, 1);
}
void S_(orc_lsf)(const float* U, const float* BS, const float* BD, R* lsfA, R* lsfB, R* phT) {
  Ctx C = mkctx(U, BS), *c = &C;
  int nb = (int)U[U_COR + 2];
#pragma omp parallel for schedule(static)
  for (int j = 0; j < c->ny; j++)
    for (int i = 0; i < c->nx; i++) {
      R px = ((R)i + (R)0.5) * c->dx, py = ((R)j + (R)0.5) * c->dx;
      R sT = 0, sH = 0, sQ = 0, sHq = 0, sU = 0, sV = 0, sHw = 0, sw = 0, nud = 0;
      for (int bi = 0; bi < nb; bi++) {
        const float* b = BD + 148 * bi;
        R d, over, along;
        sdist(c, b, px, py, &d, &over, &along);
        const float *P0 = b, *P1 = b + 4, *P2 = b + 8, *P3 = b + 12;
        R w = sstep(-(R)P0[3], (R)P0[3], d) * ((R)1 - sstep(0, (R)60000, over));
        if (w <= 0) continue;
        R vr = (R)b[16];
        R fadev = (R)1 - sstep((R)30000, (R)90000, d);
        R av = vr * alongVar(c, b, along) * fadev;
        R lift = vr > 0 ? mixr(1, mixr((R)1 - (R)0.85 * vr, (R)1 + (R)1.3 * vr, hotSpots(c, b, along)), fadev) : (R)1;
        R Hmin = (R)P1[3] * clampr(lift, (R)0.15, (R)3);
        R H = clampr((R)P1[2] * maxr(d + (R)P0[3], 0) * ((R)1 + (R)0.4 * av), Hmin, (R)P1[1]);
        sT += w * (R)P1[0] * EXP((R)0.5 * av);
        sQ += w * (R)P2[0];
        sU += w * (R)P2[2] * lift;
        sV += w * (R)P2[3] * lift;
        sH += w * H;
        sHq += w * H * (R)P2[1];
        sHw += w * H * (R)P3[0];
        sw += w;
        nud = maxr(nud, (R)P3[1] * sstep((R)P3[2], (R)P3[2] + (R)40000, d) * ((R)1 - sstep(0, (R)60000, over)));
      }
      R iw = (R)1 / maxr(sw, (R)1e-6);
      R4 A = {clampr(sT, -20, 10), maxr(sH * iw, 100), clampr(sQ, 0, (R)0.92), maxr(sHq * iw, 100)};
      R4 B = {sU, sV, maxr(sHw * iw, 100), nud};
      put4(lsfA, (size_t)i + (size_t)c->nx * j, A);
      put4(lsfB, (size_t)i + (size_t)c->nx * j, B);
      R dz = c->dz, ph = 0, bprev = 0;
      for (int k = c->nz - 1; k >= 0; k--) {
        R z = ((R)k + (R)0.5) * dz;
        R4 a = bsA(c, k), bb = bsB(c, k);
        R dth = A.x / bb.x * tprof(z, A.y);
        R dqv = -a.y * A.z * tprof(z, A.w);
        R bt = (R)9.81 * (dth / a.x + (R)0.608 * dqv);
        if (k == c->nz - 1) ph = -bt * (R)0.5 * dz;
        else ph -= (R)0.5 * (bt + bprev) * dz;
        bprev = bt;
        phT[cidx(c, i, j, k)] = ph;
      }
    }
}

/* ====================== CI map (model.js:249-281) ====================== */
void S_(orc_ci)(const float* U, const float* BS, const float* CI, R* ciMap) {
  Ctx C = mkctx(U, BS), *c = &C;
  R t = UF(c, U_STEP + 1);
  int ne = (int)U[U_CI];
#pragma omp parallel for schedule(static)
  for (int j = 0; j < c->ny; j++)
    for (int i = 0; i < c->nx; i++) {
      R px = ((R)i + (R)0.5) * c->dx, py = ((R)j + (R)0.5) * c->dx;
      R4 best = {0, 0, 1500, 1500};
      for (int e = 0; e < ne; e++) {
        const float* E = CI + 12 * e; /* a = E[0..3], b = E[4..7], c = E[8..11] */
        if (t < (R)E[6] || t > (R)E[7]) continue;
        R dxv = px - ((R)E[0] + (R)E[8] * (t - (R)E[6]));
        R dyv = py - ((R)E[1] + (R)E[9] * (t - (R)E[6]));
        R ang = ATAN2(dyv, dxv);
        R wob = (R)1 + (R)0.25 * gnoise2(COS(ang) * (R)1.6 + (R)E[10], SIN(ang) * (R)1.6 + (R)3.7 * (R)E[10]);
        R g = sstep((R)E[6], (R)E[6] + maxr((R)E[11], (R)60), t);
        R r = SQRT(dxv * dxv + dyv * dyv) / ((R)E[3] * wob * ((R)0.35 + (R)0.65 * g));
        if (r >= 1) continue;
        R Lx = (R)c->nx * c->dx, Ly = (R)c->ny * c->dx;
        R edge = minr(minr(px, Lx - px), minr(py, Ly - py));
        R h = sstep((R)E[6], (R)E[6] + (R)60, t) * ((R)1 - sstep((R)E[7] - (R)180, (R)E[7], t)) *
              sstep(UF(c, U_SPONGE), UF(c, U_SPONGE) + (R)10000, edge);
        R wt = (R)E[5] * ((R)0.4 + (R)0.6 * g) * COS((R)1.5707963 * r);
        if (h * wt > best.x * best.y) { best.x = h; best.y = wt; best.z = (R)E[2]; best.w = (R)E[4]; }
      }
      put4(ciMap, (size_t)i + (size_t)c->nx * j, best);
    }
}

/* ====================== INIT (model.js:284-339) ====================== */
/* vxAt (model.js:292-297): the vortex table's (tangential wind, theta') at radius r, level k */
static R2 vxAt(const Ctx* c, const float* VX, R r, int k) {
  R2 z0 = {0, 0};
  int nr = (int)c->U[U_VTX + 3];
  R fr = r / UF(c, U_VTX + 2);
  int i0 = (int)FLOOR(fr);
  if (!VX || i0 >= nr - 1) return z0;
  size_t o = (size_t)clampi(k, 0, c->nz - 1) * (size_t)nr + (size_t)i0;
  R t = fr - (R)i0;
  R2 v = {mixr((R)VX[4 * o], (R)VX[4 * (o + 1)], t), mixr((R)VX[4 * o + 1], (R)VX[4 * (o + 1) + 1], t)};
  return v;
}
void S_(orc_init2)(const float* U, const float* BS, const R* lsfA, const R* lsfB, const float* VX, R* V, R* Q1, R* Q2,
                   int seed) {
  Ctx C = mkctx(U, BS), *c = &C;
  int nx = c->nx, ny = c->ny, nz = c->nz;
  int vortex = c->U[U_VTX2 + 2] > 0.5f;
  R xc = UF(c, U_VTX), yc = UF(c, U_VTX + 1), dx = c->dx;
#pragma omp parallel for schedule(static)
  for (int z = 0; z <= nz; z++)
    for (int y = 0; y <= ny; y++)
      for (int x = 0; x <= nx; x++) {
        int k = z < nz - 1 ? z : nz - 1;
        R4 A = ld2d(c, lsfA, x, y);
        R zz = ((R)k + (R)0.5) * c->dz;
        R nz0 = sstep((R)800, (R)1500, zz);
        if (vortex) {
          R px = ((R)x + (R)0.5) * dx - xc, py = ((R)y + (R)0.5) * dx - yc;
          nz0 = SQRT(px * px + py * py) < UF(c, U_VTX2) ? sstep((R)6000, (R)8000, zz) : (R)1;
        }
        R noise = UF(c, U_MISC + 2) * (rnd3(x + 17, y + 3 + 65536 * seed, z + 5) - (R)0.5) * (R)2 * ((R)1 - nz0);
        R th = tgtTh(c, A, k) + noise;
        R u = (y < ny && z < nz) ? tgtU(c, lsfB, x, y, k) : (R)0;
        R v = (x < nx && z < nz) ? tgtV(c, lsfB, x, y, k) : (R)0;
        if (vortex) {
          /* the balanced warm-core vortex built on the host (setVortex, model.js:1318-1365) */
          R pux = (R)x * dx - xc, puy = ((R)y + (R)0.5) * dx - yc, ru = maxr(SQRT(pux * pux + puy * puy), 1);
          R pvx = ((R)x + (R)0.5) * dx - xc, pvy = (R)y * dx - yc, rv = maxr(SQRT(pvx * pvx + pvy * pvy), 1);
          R pcx = ((R)x + (R)0.5) * dx - xc, pcy = ((R)y + (R)0.5) * dx - yc, rc = SQRT(pcx * pcx + pcy * pcy);
          u -= vxAt(c, VX, ru, k).x * puy / ru;
          v += vxAt(c, VX, rv, k).x * pvx / rv;
          th += vxAt(c, VX, rc, k).y;
        }
        R T = th * bsB(c, k).x;
        R qv = minr(tgtQv(c, A, k), (R)0.95 * qvsat(T, bsB(c, k).y, fIce(T)));
        if (vortex) {
          /* a nearly saturated core (model.js:328-336) */
          R pcx = ((R)x + (R)0.5) * dx - xc, pcy = ((R)y + (R)0.5) * dx - yc;
          R f = ((R)1 - sstep((R)0.6, (R)1, SQRT(pcx * pcx + pcy * pcy) / UF(c, U_VTX2))) * sstep((R)500, (R)1500, zz) *
                ((R)1 - sstep((R)7000, (R)10000, zz));
          R q2 = maxr(qv, f * UF(c, U_VTX2 + 1) * qvsat(T, bsB(c, k).y, fIce(T)));
          R dqv = q2 - qv;
          qv = q2;
          th -= (R)0.608 * dqv * bsA(c, k).x;
        }
        R4 o = {u, v, 0, th};
        put4(V, vidx(c, x, y, z), o);
        if (x < nx && y < ny && z < nz) {
          R4 q1 = {qv, 0, 0, 0}, q2 = {0, 0, 0, 0};
          put4(Q1, cidx(c, x, y, z), q1);
          put4(Q2, cidx(c, x, y, z), q2);
        }
      }
}
void S_(orc_init)(const float* U, const float* BS, const R* lsfA, const R* lsfB, R* V, R* Q1, R* Q2, int seed) {
  S_(orc_init2)(U, BS, lsfA, lsfB, NULL, V, Q1, Q2, seed);
}

/* Build-defined warm bubble (no reference counterpart): theta += dth cos^2(pi r / 2),
 * r^2 = ((x-x0)/rh)^2 + ((y-y0)/rh)^2 + ((z-z0)/rz)^2, at every texel's cell centre. */
void S_(orc_bubble)(const float* U, R* V, double x0, double y0, double z0, double rh, double rz, double dth) {
  Ctx C = mkctx(U, NULL), *c = &C;
#pragma omp parallel for schedule(static)
  for (int z = 0; z <= c->nz; z++)
    for (int y = 0; y <= c->ny; y++)
      for (int x = 0; x <= c->nx; x++) {
        double px = (x + 0.5) * (double)c->dx, py = (y + 0.5) * (double)c->dx, pz = (z + 0.5) * (double)c->dz;
        double a = (px - x0) / rh, b = (py - y0) / rh, d = (pz - z0) / rz;
        double r = sqrt(a * a + b * b + d * d);
        if (r < 1.0) {
          double cc = cos(1.5707963267948966 * r);
          V[4 * vidx(c, x, y, z) + 3] += (R)(dth * cc * cc);
        }
      }
}

/* ====================== NUT (model.js:342-389), outbreak mode ====================== */
static R nut_uc(const Ctx* c, const R* V, int x, int y, int z) {
  int k = clampi(z, 0, c->nz - 1);
  return (R)0.5 * (ldV(c, V, x, y, z).x + ldV(c, V, x + 1, y, z).x) - bsB(c, k).z;
}
static R nut_vc(const Ctx* c, const R* V, int x, int y, int z) {
  int k = clampi(z, 0, c->nz - 1);
  return (R)0.5 * (ldV(c, V, x, y, z).y + ldV(c, V, x, y + 1, z).y) - bsB(c, k).w;
}
static R nut_wc(const Ctx* c, const R* V, int x, int y, int z) {
  return (R)0.5 * (ldV(c, V, x, y, z).z + ldV(c, V, x, y, z + 1).z);
}
void S_(orc_nut2)(const float* U, const float* BS, const R* V, const R* Q1, const float* sfc, R* aux) {
  Ctx C = mkctx(U, BS), *c = &C;
  R dx = c->dx, dz = c->dz;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < c->nz; z++)
    for (int y = 0; y < c->ny; y++)
      for (int x = 0; x < c->nx; x++) {
        R4 v0 = ldV(c, V, x, y, z);
        R dudx = (ldV(c, V, x + 1, y, z).x - v0.x) / dx;
        R dvdy = (ldV(c, V, x, y + 1, z).y - v0.y) / dx;
        R dwdz = (ldV(c, V, x, y, z + 1).z - v0.z) / dz;
        R dudy = (nut_uc(c, V, x, y + 1, z) - nut_uc(c, V, x, y - 1, z)) / ((R)2 * dx);
        R dudz = (nut_uc(c, V, x, y, z + 1) - nut_uc(c, V, x, y, z - 1)) / ((R)2 * dz);
        R dvdx = (nut_vc(c, V, x + 1, y, z) - nut_vc(c, V, x - 1, y, z)) / ((R)2 * dx);
        R dvdz = (nut_vc(c, V, x, y, z + 1) - nut_vc(c, V, x, y, z - 1)) / ((R)2 * dz);
        R dwdx = (nut_wc(c, V, x + 1, y, z) - nut_wc(c, V, x - 1, y, z)) / ((R)2 * dx);
        R dwdy = (nut_wc(c, V, x, y + 1, z) - nut_wc(c, V, x, y - 1, z)) / ((R)2 * dx);
        R s12 = (R)0.5 * (dudy + dvdx), s13 = (R)0.5 * (dudz + dwdx), s23 = (R)0.5 * (dvdz + dwdy);
        R S2 = (R)2 * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + (R)4 * (s12 * s12 + s13 * s13 + s23 * s23);
        int kz0 = z - 1 > 0 ? z - 1 : 0, kz1 = z + 1 < c->nz - 1 ? z + 1 : c->nz - 1;
        R N2 = (R)9.81 / bsA(c, z).x * (ldV(c, V, x, y, kz1).w - ldV(c, V, x, y, kz0).w) / ((R)(kz1 - kz0) * dz);
        R4 q1 = get4(Q1, cidx(c, x, y, z));
        if (c->U[U_COR + 3] > 0.5f && q1.y + q1.z > (R)1e-5) N2 = minr(N2, 0);
        R Ri = N2 / maxr(S2, (R)1e-10);
        R fst = SQRT(maxr(0, (R)1 - Ri / (R)0.33));
        R sS = SQRT(maxr(S2, 0));
        R lh = UF(c, U_TURB) * dx, lv = (R)0.18 * dz;
        R nuh = minr(lh * lh * sS * fst, (R)0.08 * dx * dx / maxr(c->dt, (R)0.1));
        R nuv = minr(lv * lv * sS * fst, (R)0.15 * dz * dz / maxr(c->dt, (R)0.1));
        /* hurricane mode: the K-profile boundary layer, K = kappa u* z (1 - z/h)^2 (model.js:377-386) */
        R zk = ((R)z + (R)0.5) * dz, hb = UF(c, U_HUR2 + 2);
        if (hurOn(c) && zk < hb) {
          R4 s0 = ldV(c, V, x, y, 0);
          R ua = (R)0.5 * (s0.x + ldV(c, V, x + 1, y, 0).x), va = (R)0.5 * (s0.y + ldV(c, V, x, y + 1, 0).y);
          R u1 = SQRT(ua * ua + va * va);
          R ust = SQRT(cdSfc(c, u1, sfcAt(c, sfc, x, y, 0))) * maxr(u1, UF(c, U_HUR3 + 1));
          R q = (R)1 - zk / hb;
          nuv = maxr(nuv, minr((R)0.4 * ust * zk * q * q, (R)0.15 * dz * dz / maxr(c->dt, (R)0.1)));
        }
        size_t i = cidx(c, x, y, z);
        aux[2 * i] = nuh;
        aux[2 * i + 1] = nuv;
      }
}

void S_(orc_nut)(const float* U, const float* BS, const R* V, const R* Q1, R* aux) {
  S_(orc_nut2)(U, BS, V, Q1, NULL, aux);
}

/* ====================== advection of V (model.js:392-481) ====================== */
typedef struct { R x, y, z; } R3;
typedef struct { R3 vu, vv, vw, vc; } Loc;
static Loc localVel(const Ctx* c, const R* V, int x, int y, int z) {
  R4 p = ldV(c, V, x, y, z);
  R4 xm = ldV(c, V, x - 1, y, z), xp = ldV(c, V, x + 1, y, z);
  R4 ym = ldV(c, V, x, y - 1, z), yp = ldV(c, V, x, y + 1, z);
  R4 zm = ldV(c, V, x, y, z - 1), zp = ldV(c, V, x, y, z + 1);
  R4 xmyp = ldV(c, V, x - 1, y + 1, z), xmzp = ldV(c, V, x - 1, y, z + 1);
  R4 xpym = ldV(c, V, x + 1, y - 1, z), ymzp = ldV(c, V, x, y - 1, z + 1);
  R4 xpzm = ldV(c, V, x + 1, y, z - 1), ypzm = ldV(c, V, x, y + 1, z - 1);
  Loc L;
  L.vu.x = p.x; L.vu.y = (R)0.25 * (xm.y + p.y + xmyp.y + yp.y); L.vu.z = (R)0.25 * (xm.z + p.z + xmzp.z + zp.z);
  L.vv.x = (R)0.25 * (ym.x + p.x + xpym.x + xp.x); L.vv.y = p.y; L.vv.z = (R)0.25 * (ym.z + p.z + ymzp.z + zp.z);
  L.vw.x = (R)0.25 * (zm.x + p.x + xpzm.x + xp.x); L.vw.y = (R)0.25 * (zm.y + p.y + ypzm.y + yp.y); L.vw.z = p.z;
  L.vc.x = (R)0.5 * (p.x + xp.x); L.vc.y = (R)0.5 * (p.y + yp.y); L.vc.z = (R)0.5 * (p.z + zp.z);
  return L;
}
static R3 kvec(const Ctx* c) { R3 k = {c->dt / c->dx, c->dt / c->dx, c->dt / c->dz}; return k; }
#define VSAMP(T, tx, ty, tz) sample3(T, c->nx + 1, c->ny + 1, c->nz + 1, tx, ty, tz)
/* texU/texVv/texW/texC (model.js:411-414) in texel coordinates */
static R sampU(const Ctx* c, const R* T, R3 p) { return VSAMP(T, p.x, p.y - (R)0.5, p.z - (R)0.5).x; }
static R sampV(const Ctx* c, const R* T, R3 p) { return VSAMP(T, p.x - (R)0.5, p.y, p.z - (R)0.5).y; }
static R sampW(const Ctx* c, const R* T, R3 p) { return VSAMP(T, p.x - (R)0.5, p.y - (R)0.5, p.z).z; }
static R sampC(const Ctx* c, const R* T, R3 p) { return VSAMP(T, p.x - (R)0.5, p.y - (R)0.5, p.z - (R)0.5).w; }
/* velAt / velQ (model.js:416-418, 488-492): the face winds sampled at p */
static R3 velAt(const Ctx* c, const R* V, R3 p) { R3 v = {sampU(c, V, p), sampV(c, V, p), sampW(c, V, p)}; return v; }
static R3 clampX(const Ctx* c, R3 p) {
  R3 r = {clampr(p.x, 0, (R)c->nx), clampr(p.y, 0, (R)c->ny), clampr(p.z, 0, (R)c->nz)};
  return r;
}
/* departure (s = 1) / forward (s = -1) point, clamped to [0, N] (model.js:420-430): first order, or with the wind
 * at the trajectory's midpoint when cn.w is set (hurricane mode) */
static R3 trace(const Ctx* c, const R* V, R3 p, R3 v0, R s) {
  R3 k = kvec(c), r;
  if (c->U[U_CN + 3] < 0.5f) {
    r.x = p.x - s * k.x * v0.x; r.y = p.y - s * k.y * v0.y; r.z = p.z - s * k.z * v0.z;
    return clampX(c, r);
  }
  R3 m = {p.x - (R)0.5 * s * k.x * v0.x, p.y - (R)0.5 * s * k.y * v0.y, p.z - (R)0.5 * s * k.z * v0.z};
  R3 vm = velAt(c, V, clampX(c, m));
  r.x = p.x - s * k.x * vm.x; r.y = p.y - s * k.y * vm.y; r.z = p.z - s * k.z * vm.z;
  return clampX(c, r);
}

/* V carries the trajectories; Vval the values carried (hurricane mode: Vin less 40 % of the last pressure force,
 * WGSL_M_PREP; otherwise V itself) */
void S_(orc_advv1_2)(const float* U, const float* BS, const R* V, const R* Vval, R* Vt) {
  Ctx C = mkctx(U, BS), *c = &C;
#pragma omp parallel for schedule(static)
  for (int z = 0; z <= c->nz; z++)
    for (int y = 0; y <= c->ny; y++)
      for (int x = 0; x <= c->nx; x++) {
        Loc L = localVel(c, V, x, y, z);
        R3 pu = {(R)x, (R)y + (R)0.5, (R)z + (R)0.5}, pv = {(R)x + (R)0.5, (R)y, (R)z + (R)0.5};
        R3 pw = {(R)x + (R)0.5, (R)y + (R)0.5, (R)z}, pc = {(R)x + (R)0.5, (R)y + (R)0.5, (R)z + (R)0.5};
        R4 o;
        o.x = sampU(c, Vval, trace(c, V, pu, L.vu, 1));
        o.y = sampV(c, Vval, trace(c, V, pv, L.vv, 1));
        o.z = sampW(c, Vval, trace(c, V, pw, L.vw, 1));
        o.w = sampC(c, Vval, trace(c, V, pc, L.vc, 1));
        put4(Vt, vidx(c, x, y, z), o);
      }
}
void S_(orc_advv1)(const float* U, const float* BS, const R* V, R* Vt) { S_(orc_advv1_2)(U, BS, V, V, Vt); }
static void mmV(const Ctx* c, const R* T, R3 p, R ox, R oy, R oz, int ch, R* lo, R* hi) {
  int i0 = (int)FLOOR(p.x - ox), j0 = (int)FLOOR(p.y - oy), k0 = (int)FLOOR(p.z - oz);
  R a = (R)1e30, b = (R)-1e30;
  for (int q = 0; q < 8; q++) {
    R s = ch4(ldV(c, T, i0 + (q & 1), j0 + ((q >> 1) & 1), k0 + (q >> 2)), ch);
    a = minr(a, s);
    b = maxr(b, s);
  }
  *lo = a;
  *hi = b;
}
void S_(orc_advv2_2)(const float* U, const float* BS, const R* V, const R* Vval, const R* Vt, R* Vout) {
  Ctx C = mkctx(U, BS), *c = &C;
  R3 k = kvec(c);
  R hmc = (R)0.5 * UF(c, U_MISC + 1);
#pragma omp parallel for schedule(static)
  for (int z = 0; z <= c->nz; z++)
    for (int y = 0; y <= c->ny; y++)
      for (int x = 0; x <= c->nx; x++) {
        Loc L = localVel(c, V, x, y, z);
        R3 pu = {(R)x, (R)y + (R)0.5, (R)z + (R)0.5}, pv = {(R)x + (R)0.5, (R)y, (R)z + (R)0.5};
        R3 pw = {(R)x + (R)0.5, (R)y + (R)0.5, (R)z}, pc = {(R)x + (R)0.5, (R)y + (R)0.5, (R)z + (R)0.5};
        R4 hat = get4(Vt, vidx(c, x, y, z));
        R4 old = ldV(c, Vval, x, y, z);
        R4 b;
        b.x = sampU(c, Vt, trace(c, V, pu, L.vu, -1));
        b.y = sampV(c, Vt, trace(c, V, pv, L.vv, -1));
        b.z = sampW(c, Vt, trace(c, V, pw, L.vw, -1));
        b.w = sampC(c, Vt, trace(c, V, pc, L.vc, -1));
        R cx = L.vc.x * k.x, cy = L.vc.y * k.y, cz = L.vc.z * k.z;
        R cour = SQRT(cx * cx + cy * cy + cz * cz);
        R sm = hmc * clampr(cour / (R)0.35, UF(c, U_DZV + 3), 1);
        R4 r = {hat.x + sm * (old.x - b.x), hat.y + sm * (old.y - b.y), hat.z + sm * (old.z - b.z), hat.w + sm * (old.w - b.w)};
        R lo, hi;
        mmV(c, Vval, trace(c, V, pu, L.vu, 1), 0, (R)0.5, (R)0.5, 0, &lo, &hi); r.x = clampr(r.x, lo, hi);
        mmV(c, Vval, trace(c, V, pv, L.vv, 1), (R)0.5, 0, (R)0.5, 1, &lo, &hi); r.y = clampr(r.y, lo, hi);
        mmV(c, Vval, trace(c, V, pw, L.vw, 1), (R)0.5, (R)0.5, 0, 2, &lo, &hi); r.z = clampr(r.z, lo, hi);
        mmV(c, Vval, trace(c, V, pc, L.vc, 1), (R)0.5, (R)0.5, (R)0.5, 3, &lo, &hi); r.w = clampr(r.w, lo, hi);
        put4(Vout, vidx(c, x, y, z), r);
      }
}
void S_(orc_advv2)(const float* U, const float* BS, const R* V, const R* Vt, R* Vout) {
  S_(orc_advv2_2)(U, BS, V, V, Vt, Vout);
}

/* ====================== advection of moisture (model.js:482-576) ====================== */
static R3 cellVel(const Ctx* c, const R* V, int x, int y, int z) {
  R4 p = ldV(c, V, x, y, z);
  R3 v = {(R)0.5 * (p.x + ldV(c, V, x + 1, y, z).x), (R)0.5 * (p.y + ldV(c, V, x, y + 1, z).y),
          (R)0.5 * (p.z + ldV(c, V, x, y, z + 1).z)};
  return v;
}
/* trace of something falling at `fall`, clamped to [1/2, N - 1/2] (model.js:494-499): first order, or (cn.w set,
 * hurricane mode) with the wind sampled at the midpoint, clamped to [0, N] */
static R3 traceQ(const Ctx* c, const R* V, R3 p, R3 v0, R fall, R s) {
  R3 k = kvec(c), r, vm = {v0.x, v0.y, v0.z - fall};
  if (c->U[U_CN + 3] > 0.5f) {
    R3 m = {p.x - (R)0.5 * s * k.x * vm.x, p.y - (R)0.5 * s * k.y * vm.y, p.z - (R)0.5 * s * k.z * vm.z};
    R3 w = velAt(c, V, clampX(c, m));
    vm.x = w.x; vm.y = w.y; vm.z = w.z - fall;
  }
  r.x = clampr(p.x - s * k.x * vm.x, (R)0.5, (R)c->nx - (R)0.5);
  r.y = clampr(p.y - s * k.y * vm.y, (R)0.5, (R)c->ny - (R)0.5);
  r.z = clampr(p.z - s * k.z * vm.z, (R)0.5, (R)c->nz - (R)0.5);
  return r;
}
/* texS(x) = x / N: texel coordinate x - 1/2 */
static R4 sampQ(const Ctx* c, const R* T, R3 p) {
  return sample3(T, c->nx, c->ny, c->nz, p.x - (R)0.5, p.y - (R)0.5, p.z - (R)0.5);
}
static void mmQ(const Ctx* c, const R* T, R3 p, R4* lo, R4* hi) {
  int i0 = (int)FLOOR(p.x - (R)0.5), j0 = (int)FLOOR(p.y - (R)0.5), k0 = (int)FLOOR(p.z - (R)0.5);
  R4 a = {(R)1e30, (R)1e30, (R)1e30, (R)1e30}, b = {(R)-1e30, (R)-1e30, (R)-1e30, (R)-1e30};
  for (int q = 0; q < 8; q++) {
    R4 s = ldQ(c, T, i0 + (q & 1), j0 + ((q >> 1) & 1), k0 + (q >> 2));
    a.x = minr(a.x, s.x); a.y = minr(a.y, s.y); a.z = minr(a.z, s.z); a.w = minr(a.w, s.w);
    b.x = maxr(b.x, s.x); b.y = maxr(b.y, s.y); b.z = maxr(b.z, s.z); b.w = maxr(b.w, s.w);
  }
  *lo = a;
  *hi = b;
}
void S_(orc_advq1)(const float* U, const float* BS, const R* V, const R* Q1, const R* Q2, R* Q1t, R* Q2t) {
  Ctx C = mkctx(U, BS), *c = &C;
  int thm = thompson(c);
#pragma omp parallel for schedule(static)
  for (int z = 0; z < c->nz; z++)
    for (int y = 0; y < c->ny; y++)
      for (int x = 0; x < c->nx; x++) {
        R3 v = cellVel(c, V, x, y, z);
        R3 p = {(R)x + (R)0.5, (R)y + (R)0.5, (R)z + (R)0.5};
        size_t i = cidx(c, x, y, z);
        R4 q1 = get4(Q1, i), q2 = get4(Q2, i);
        R rho = bsA(c, z).z;
        R2 fv = vtRain(c, q1.w, q2.z, rho);
        R fs = vtS(c, q2.x, rho), fg = vtGrau(c, q2.y, rho), fi = vtI(c, q1.z, rho);
        R4 a = sampQ(c, Q1, traceQ(c, V, p, v, 0, 1));
        if (fi > 0) a.z = sampQ(c, Q1, traceQ(c, V, p, v, fi, 1)).z;
        R r = sampQ(c, Q1, traceQ(c, V, p, v, fv.x, 1)).w;
        R4 b = sampQ(c, Q2, traceQ(c, V, p, v, 0, 1));
        R s = sampQ(c, Q2, traceQ(c, V, p, v, fs, 1)).x;
        R g = sampQ(c, Q2, traceQ(c, V, p, v, fg, 1)).y;
        R nr = b.z;
        if (thm) nr = sampQ(c, Q2, traceQ(c, V, p, v, fv.y, 1)).z;
        R4 o1 = {a.x, a.y, a.z, r}, o2 = {s, g, nr, b.w};
        put4(Q1t, i, o1);
        put4(Q2t, i, o2);
      }
}
void S_(orc_advq2)(const float* U, const float* BS, const R* V, const R* Q1, const R* Q2, const R* Q1t,
                   const R* Q2t, R* Q1o, R* Q2o) {
  Ctx C = mkctx(U, BS), *c = &C;
  int thm = thompson(c);
  R3 k = kvec(c);
  R hmc = (R)0.5 * UF(c, U_MISC + 1);
#pragma omp parallel for schedule(static)
  for (int z = 0; z < c->nz; z++)
    for (int y = 0; y < c->ny; y++)
      for (int x = 0; x < c->nx; x++) {
        R3 v = cellVel(c, V, x, y, z);
        R3 p = {(R)x + (R)0.5, (R)y + (R)0.5, (R)z + (R)0.5};
        size_t i = cidx(c, x, y, z);
        R4 q1 = get4(Q1, i), q2 = get4(Q2, i);
        R rho = bsA(c, z).z;
        R2 fv = vtRain(c, q1.w, q2.z, rho);
        R fr = fv.x, fs = vtS(c, q2.x, rho), fg = vtGrau(c, q2.y, rho), fi = vtI(c, q1.z, rho);
        R4 h1 = get4(Q1t, i), h2 = get4(Q2t, i);
        R cx = v.x * k.x, cy = v.y * k.y, cz = v.z * k.z;
        R cour = SQRT(cx * cx + cy * cy + cz * cz);
        R sm = hmc * clampr(cour / (R)0.35, UF(c, U_DZV + 3), 1);
        R4 b1 = sampQ(c, Q1t, traceQ(c, V, p, v, 0, -1));
        R br = sampQ(c, Q1t, traceQ(c, V, p, v, fr, -1)).w;
        R4 b2 = sampQ(c, Q2t, traceQ(c, V, p, v, 0, -1));
        R bs = sampQ(c, Q2t, traceQ(c, V, p, v, fs, -1)).x;
        R bg = sampQ(c, Q2t, traceQ(c, V, p, v, fg, -1)).y;
        R fn = thm ? fv.y : (R)0;
        R bn = sampQ(c, Q2t, traceQ(c, V, p, v, fn, -1)).z;
        R bi = fi > 0 ? sampQ(c, Q1t, traceQ(c, V, p, v, fi, -1)).z : b1.z;
        R4 r1 = {h1.x + sm * (q1.x - b1.x), h1.y + sm * (q1.y - b1.y), h1.z + sm * (q1.z - bi), h1.w + sm * (q1.w - br)};
        R4 r2 = {h2.x + sm * (q2.x - bs), h2.y + sm * (q2.y - bg), h2.z + sm * (q2.z - bn), h2.w + sm * (q2.w - b2.w)};
        R4 l1, u1, lr, ur, l2, u2, ls, us, lg, ug, ln, un;
        mmQ(c, Q1, traceQ(c, V, p, v, 0, 1), &l1, &u1);
        mmQ(c, Q1, traceQ(c, V, p, v, fr, 1), &lr, &ur);
        mmQ(c, Q2, traceQ(c, V, p, v, 0, 1), &l2, &u2);
        mmQ(c, Q2, traceQ(c, V, p, v, fs, 1), &ls, &us);
        mmQ(c, Q2, traceQ(c, V, p, v, fg, 1), &lg, &ug);
        R4 li, ui;
        mmQ(c, Q1, traceQ(c, V, p, v, fi, 1), &li, &ui);
        r1.x = clampr(r1.x, l1.x, u1.x);
        r1.y = clampr(r1.y, l1.y, u1.y);
        r1.z = clampr(r1.z, li.z, ui.z);
        r1.w = clampr(r1.w, lr.w, ur.w);
        mmQ(c, Q2, traceQ(c, V, p, v, fn, 1), &ln, &un);
        r2.x = clampr(r2.x, ls.x, us.x);
        r2.y = clampr(r2.y, lg.y, ug.y);
        r2.z = clampr(r2.z, ln.z, un.z);
        r2.w = clampr(r2.w, l2.w, u2.w);
        R4 o1 = {maxr(r1.x, 0), maxr(r1.y, 0), maxr(r1.z, 0), maxr(r1.w, 0)};
        R4 o2 = {maxr(r2.x, 0), maxr(r2.y, 0), maxr(r2.z, 0), maxr(r2.w, 0)};
        put4(Q1o, i, o1);
        put4(Q2o, i, o2);
      }
}

/* ====================== forces + microphysics (model.js:579-907) ====================== */
static R lapH(const Ctx* c, const R* V, int x, int y, int z, int ch, R me) {
  return ch4(ldV(c, V, x + 1, y, z), ch) + ch4(ldV(c, V, x - 1, y, z), ch) + ch4(ldV(c, V, x, y + 1, z), ch) +
         ch4(ldV(c, V, x, y - 1, z), ch) - (R)4 * me;
}
static R lapZ(const Ctx* c, const R* V, int x, int y, int z, int ch, R me, int lo, int hi) {
  R a = lo ? me : ch4(ldV(c, V, x, y, z - 1), ch);
  R b = hi ? me : ch4(ldV(c, V, x, y, z + 1), ch);
  return a + b - (R)2 * me;
}
static R bsV(const Ctx* c, int k, int ch) {
  int kk = clampi(k, 0, c->nz - 1);
  if (ch == 0) return bsB(c, kk).z;
  if (ch == 1) return bsB(c, kk).w;
  if (ch == 3) return bsA(c, kk).x;
  return 0;
}
static R lapZp(const Ctx* c, const R* V, int x, int y, int z, int ch, R me, int lo, int hi) {
  R m0 = me - bsV(c, z, ch);
  R a = lo ? m0 : ch4(ldV(c, V, x, y, z - 1), ch) - bsV(c, z - 1, ch);
  R b = hi ? m0 : ch4(ldV(c, V, x, y, z + 1), ch) - bsV(c, z + 1, ch);
  return a + b - (R)2 * m0;
}
static R spongeLat(const Ctx* c, R xm, R ym) {
  R Lx = (R)c->nx * c->dx, Ly = (R)c->ny * c->dx;
  R e = minr(minr(xm, Lx - xm), minr(ym, Ly - ym));
  R d = clampr((R)1 - e / UF(c, U_SPONGE), 0, 1);
  return d * ((R)0.3 + (R)0.7 * d) / UF(c, U_SPONGE + 1);
}
static R spongeTop(const Ctx* c, R z) {
  R ztop = (R)c->nz * c->dz;
  if (z < UF(c, U_SPONGE + 2)) return 0;
  R s = SIN((R)1.5707963 * (z - UF(c, U_SPONGE + 2)) / (ztop - UF(c, U_SPONGE + 2)));
  return s * s / UF(c, U_SPONGE + 3);
}
static R buoy(const Ctx* c, const R* V, const R* Q1, const R* Q2, int x, int y, int z) {
  R4 A = bsA(c, z), v = ldV(c, V, x, y, z), q1 = ldQ(c, Q1, x, y, z), q2 = ldQ(c, Q2, x, y, z);
  return (R)9.81 * ((v.w - A.x) / A.x + (R)0.608 * (q1.x - A.y) - (q1.y + q1.z + q1.w + q2.x + q2.y));
}
static R2 ldA(const Ctx* c, const R* aux, int x, int y, int z) {
  size_t i = cidx(c, clampi(x, 0, c->nx - 1), clampi(y, 0, c->ny - 1), clampi(z, 0, c->nz - 1));
  R2 r = {aux[2 * i], aux[2 * i + 1]};
  return r;
}
static R ldP(const Ctx* c, const R* phT, int x, int y, int z) {
  return phT[cidx(c, clampi(x, 0, c->nx - 1), clampi(y, 0, c->ny - 1), clampi(z, 0, c->nz - 1))];
}

/* The microphysics block (model.js:775-899), from the clamped mixing ratios to the
 * saturation adjustment; st = (T in, then th out), qv, qc, qi, qr, qs, qg, nr. */
typedef struct { R T, qv, qc, qi, qr, qs, qg, nr; } MState;
static void micro_block(const Ctx* c, int k, MState* m) {
  R dt = c->dt;
  R4 A = bsA(c, k), B = bsB(c, k);
  R p0 = B.y, rho = A.z;
  int thm = thompson(c);
  R T = m->T, qv = m->qv, qc = m->qc, qi = m->qi, qr = m->qr, qs = m->qs, qg = m->qg, nr = m->nr;
  R rq = (R)1000 * rho;
  R rs = SQRT(UF(c, U_STEP + 2) / rho);
  R pauto = UF(c, U_MICRO) * maxr(qc - UF(c, U_MICRO + 1), 0);
  R accr = (R)2.2 * qc * POW(qr, (R)0.875);
  if (thm && qr > (R)1e-9) {
    R l = lamR(qr, nr);
    accr = (R)3100 * rho * nr * POW(l, (R)-2.8) * qc * rs;
  }
  R da = minr(pauto * dt, qc);
  R dc = minr(accr * dt, qc - da);
  qc -= da + dc;
  qr += da + dc;
  nr += da / (R)2.68e-10;
  R dTc = T - T0K;
  if (dTc < 0) {
    R psaut = (R)1e-3 * EXP((R)0.025 * dTc) * maxr(qi - (R)6e-4, 0);
    R psacw = (R)5.4e-3 * qc * POW(maxr(rq * qs, 0), (R)0.81) * rs;
    R psaci = (R)5.4e-3 * qi * POW(maxr(rq * qs, 0), (R)0.81) * EXP((R)0.025 * dTc) * rs;
    R ewet = (R)1 / ((R)1 + qg / (R)0.008);
    R pgacw = (R)3.1e-3 * qc * POW(maxr(rq * qg, 0), (R)0.875) * rs * ewet;
    R pgaci = (R)3.0e-4 * qi * POW(maxr(rq * qg, 0), (R)0.875) * rs;
    R pgacr = (R)3.1e-3 * qr * POW(maxr(rq * qg, 0), (R)0.875) * rs * ewet;
    R pgfr = minr((R)2e-4 * UF(c, U_MICRO + 2) * (EXP((R)0.66 * minr(-dTc, (R)20)) - (R)1), (R)0.08) * qr;
    R pgaut = (R)1e-3 * EXP((R)0.09 * dTc) * maxr(qs - (R)6e-4, 0) * (thm ? (R)0.25 : (R)1);
    R a1 = minr((psaut + psaci + pgaci) * dt, qi);
    R fs1 = (psaut + psaci) / maxr(psaut + psaci + pgaci, (R)1e-20);
    qi -= a1;
    qs += a1 * fs1;
    qg += a1 * ((R)1 - fs1);
    R a2 = minr((psacw + pgacw) * dt, qc);
    R fs2 = psacw / maxr(psacw + pgacw, (R)1e-20);
    R rimeS = thm ? (R)0.95 : (R)0.7;
    qc -= a2;
    qs += a2 * fs2 * rimeS;
    qg += a2 * ((R)1 - fs2 * rimeS);
    R a3 = minr((pgacr + pgfr) * dt, qr);
    if (T < (R)235) a3 = qr;
    if (qr > (R)1e-12) nr *= (qr - a3) / qr;
    qr -= a3;
    qg += a3;
    R a4 = minr(pgaut * dt, qs);
    qs -= a4;
    qg += a4;
    T += LFC / CPD * (a2 + a3);
  } else {
    R m1 = minr((R)0.012 * dTc * qs * dt, qs);
    R m2 = minr((R)0.0022 * dTc * qg * dt, qg);
    if (thm) {
      if (qg > (R)1e-9) {
        R l = lamG(qg, rho), n0 = n0G(qg, rho);
        R vent = (R)0.78 / (l * l) + (R)0.31 * (R)0.843 * (R)1.89 * (R)5430 * POW(UF(c, U_STEP + 2) / rho, (R)0.25) *
                                         POW(l, (R)-2.945);
        R heat = (R)2.4e-2 * dTc - LVC * (R)2.26e-5 * rho * (qvsat(T0K, p0, 0) - qv);
        m2 = clampr((R)6.2832 / (rho * LFC) * heat * n0 * vent * dt, 0, qg);
        nr += m2 / qg * n0 / (rho * l);
      }
      if (qs > (R)1e-9) {
        R ls = POW((R)3.14159 * (R)100 * (R)2e6 / (rho * qs), (R)0.25);
        nr += m1 / qs * (R)2e6 / (rho * ls);
      }
    }
    qs -= m1;
    qg -= m2;
    qr += m1 + m2;
    T -= LFC / CPD * (m1 + m2);
  }
  R qvsl = qvsat(T, p0, 0);
  if (qr > 0 && qv < qvsl) {
    R rc = (R)0.001 * rho * qr;
    R ern = (((R)1.6 + (R)124.9 * POW(rc, (R)0.2046)) * POW(rc, (R)0.525) / ((R)2.55e8 / (p0 * qvsl) + (R)5.4e5)) *
            ((qvsl - qv) / ((R)0.001 * rho * qvsl));
    R srt = UF(c, U_MICRO2 + 1) > 0 ? SQRT(rc / (rc + (R)0.001 * rho * UF(c, U_MICRO2 + 1))) : (R)1;
    R e = ern * srt * dt * UF(c, U_TURB + 2);
    if (thm) {
      R l = lamR(qr, nr), n0 = rho * nr * l;
      R Ap = LVC * LVC / ((R)2.4e-2 * RVC * T * T), Bp = (R)1 / (rho * qvsl * (R)2.26e-5);
      R vent = (R)0.78 / (l * l) + (R)0.31 * (R)0.843 * (R)1.827 * (R)7492 * POW(UF(c, U_STEP + 2) / rho, (R)0.25) *
                                       POW(l, (R)-2.9);
      e = (R)6.2832 * ((R)1 - qv / qvsl) * n0 * vent / (rho * (Ap + Bp)) * dt * UF(c, U_MICRO2 + 3);
    }
    e = clampr(e, 0, minr(qr, qvsl - qv));
    if (thm && qr > (R)1e-12) nr *= (qr - e) / qr;
    qr -= e;
    qv += e;
    T -= LVC / CPD * e;
  }
  if (thm) {
    if (qr > (R)1e-9) {
      R dm = (R)1.817 / lamR(qr, nr);
      R fb = clampr((R)1 - EXP((R)2300 * (dm - (R)1.1e-3)), -2, 1);
      nr *= EXP((R)-7.12 * rho * qr * fb * dt);
      nr = clampr(nr, qr * (R)3.183e5, qr * (R)3.183e11);
    } else {
      nr = 0;
    }
  }
  R qvsi = qvsat(T, p0, 1);
  if (dTc < 0 && qv < qvsi) {
    R sub = maxr((R)1 - qv / qvsi, 0);
    R e = minr(((R)2e-3 * qs + (R)5e-4 * qg) * sub * dt, (qvsi - qv) * (R)0.5);
    R fsn = qs / maxr(qs + qg, (R)1e-20);
    e = minr(e, qs + qg);
    qs -= e * fsn;
    qg -= e * ((R)1 - fsn);
    qv += e;
    T -= LSC / CPD * e;
  }
  R qi0 = qi, qcl = qc + qi, dsum = 0;
  for (int it = 0; it < 3; it++) {
    R f = fIce(T);
    R qs_ = qvsat(T, p0, f);
    R L = mixr(LVC, LSC, f);
    R d = (qv - qs_) / ((R)1 + L * L * qs_ / (CPD * RVC * T * T));
    d = maxr(d, -qcl);
    qv -= d;
    qcl += d;
    T += L / CPD * d;
    dsum += d;
  }
  R f = fIce(T);
  qi = f * qcl;
  qc = qcl - qi;
  T += LFC / CPD * (qi - qi0 - f * dsum);
  m->T = T; m->qv = qv; m->qc = qc; m->qi = qi; m->qr = qr; m->qs = qs; m->qg = qg; m->nr = nr;
}

/* One point of the microphysics block: st = (theta, qv, qc, qi, qr, qs, qg, nr) in and out. */
void S_(orc_micro_point)(const float* U, const float* BS, int k, R* st) {
  Ctx C = mkctx(U, BS), *c = &C;
  R pi0 = bsB(c, k).x;
  MState m = {st[0] * pi0, st[1], st[2], st[3], st[4], st[5], st[6], st[7]};
  micro_block(c, k, &m);
  st[0] = m.T / pi0; st[1] = m.qv; st[2] = m.qc; st[3] = m.qi; st[4] = m.qr; st[5] = m.qs; st[6] = m.qg; st[7] = m.nr;
}

static size_t pidx(const Ctx* c, int x, int y, int z);
/* sfc: the sea surface map (NULL outside hurricane mode); phi: the last solve's padded phi (read by the sea
 * fluxes' surface pressure in hurricane mode only) */
void S_(orc_phys2)(const float* U, const float* BS, const R* Vin, const R* Q1in, const R* Q2in, const R* aux,
                   const R* lsfA, const R* lsfB, const R* phT, const R* ciMap, const float* sfc, const R* phi, R* Vout,
                   R* Q1o, R* Q2o) {
  Ctx C = mkctx(U, BS), *c = &C;
  int hur = hurOn(c);
  int nx = c->nx, ny = c->ny, nz = c->nz;
  R dx = c->dx, dz = c->dz, dt = c->dt;
  R ix2 = (R)1 / (dx * dx), iz2 = (R)1 / (dz * dz);
  R fc = UF(c, U_COR), nudR = (R)1 / UF(c, U_COR + 1);
  R nu0 = UF(c, U_TURB + 1);
  int thm = thompson(c);
#pragma omp parallel for schedule(static)
  for (int z = 0; z <= nz; z++)
    for (int y = 0; y <= ny; y++)
      for (int x = 0; x <= nx; x++) {
        R4 V = ldV(c, Vin, x, y, z), o = V;
        /* ---- u ---- */
        if (y < ny && z < nz) {
          R xmx = (R)x * dx, xmy = ((R)y + (R)0.5) * dx, zz = ((R)z + (R)0.5) * dz;
          R2 a0 = ldA(c, aux, x - 1, y, z), a1 = ldA(c, aux, x, y, z);
          R nux = (R)0.5 * (a0.x + a1.x) + nu0, nuy = (R)0.5 * (a0.y + a1.y) + nu0;
          R du = nux * lapH(c, Vin, x, y, z, 0, V.x) * ix2 + nuy * lapZp(c, Vin, x, y, z, 0, V.x, z == 0, z == nz - 1) * iz2;
          R ut = tgtU(c, lsfB, x, y, z);
          R vbar = (R)0.25 * (ldV(c, Vin, x - 1, y, z).y + V.y + ldV(c, Vin, x - 1, y + 1, z).y + ldV(c, Vin, x, y + 1, z).y);
          R vtb = (R)0.5 * (tgtV(c, lsfB, x - 1, y, z) + tgtV(c, lsfB, x, y + 1, z));
          R nw = (R)0.5 * (ld2d(c, lsfB, x - 1, y).w + ld2d(c, lsfB, x, y).w);
          du += (spongeLat(c, xmx, xmy) + spongeTop(c, zz) + nw * nudR) * (ut - V.x);
          du += fc * (vbar - vtb);
          if (x > 0 && x < nx) du -= (ldP(c, phT, x, y, z) - ldP(c, phT, x - 1, y, z)) / dx;
          if (z == 0 && c->U[U_TURB + 3] > 0.5f) {
            R lg = SQRT(V.x * V.x + vbar * vbar), lg0 = SQRT(ut * ut + vtb * vtb);
            R cd = hur ? cdSfc(c, lg, (R)0.5 * (sfcAt(c, sfc, x - 1 > 0 ? x - 1 : 0, y, 0) + sfcAt(c, sfc, x, y, 0)))
                       : UF(c, U_MICRO + 3);
            du += cd * (-lg * V.x + lg0 * ut) / dz;
          }
          o.x = V.x + dt * du;
        } else {
          o.x = 0;
        }
        /* ---- v ---- */
        if (x < nx && z < nz) {
          R xmx = ((R)x + (R)0.5) * dx, xmy = (R)y * dx, zz = ((R)z + (R)0.5) * dz;
          R2 a0 = ldA(c, aux, x, y - 1, z), a1 = ldA(c, aux, x, y, z);
          R nux = (R)0.5 * (a0.x + a1.x) + nu0, nuy = (R)0.5 * (a0.y + a1.y) + nu0;
          R dv = nux * lapH(c, Vin, x, y, z, 1, V.y) * ix2 + nuy * lapZp(c, Vin, x, y, z, 1, V.y, z == 0, z == nz - 1) * iz2;
          R vt = tgtV(c, lsfB, x, y, z);
          R ubar = (R)0.25 * (ldV(c, Vin, x, y - 1, z).x + V.x + ldV(c, Vin, x + 1, y - 1, z).x + ldV(c, Vin, x + 1, y, z).x);
          R utb = (R)0.5 * (tgtU(c, lsfB, x, y - 1, z) + tgtU(c, lsfB, x + 1, y, z));
          R nw = (R)0.5 * (ld2d(c, lsfB, x, y - 1).w + ld2d(c, lsfB, x, y).w);
          dv += (spongeLat(c, xmx, xmy) + spongeTop(c, zz) + nw * nudR) * (vt - V.y);
          dv -= fc * (ubar - utb);
          if (y > 0 && y < ny) dv -= (ldP(c, phT, x, y, z) - ldP(c, phT, x, y - 1, z)) / dx;
          if (z == 0 && c->U[U_TURB + 3] > 0.5f) {
            R lg = SQRT(ubar * ubar + V.y * V.y), lg0 = SQRT(utb * utb + vt * vt);
            R cd = hur ? cdSfc(c, lg, (R)0.5 * (sfcAt(c, sfc, x, y - 1 > 0 ? y - 1 : 0, 0) + sfcAt(c, sfc, x, y, 0)))
                       : UF(c, U_MICRO + 3);
            dv += cd * (-lg * V.y + lg0 * vt) / dz;
          }
          o.y = V.y + dt * dv;
        } else {
          o.y = 0;
        }
        /* ---- w ---- */
        if (x < nx && y < ny && z > 0 && z < nz) {
          R xmx = ((R)x + (R)0.5) * dx, xmy = ((R)y + (R)0.5) * dx, zz = (R)z * dz;
          R2 a0 = ldA(c, aux, x, y, z - 1), a1 = ldA(c, aux, x, y, z);
          R nux = (R)0.5 * (a0.x + a1.x) + nu0, nuy = (R)0.5 * (a0.y + a1.y) + nu0;
          R dw = nux * lapH(c, Vin, x, y, z, 2, V.z) * ix2 + nuy * lapZ(c, Vin, x, y, z, 2, V.z, 0, 0) * iz2;
          dw += (R)0.5 * (buoy(c, Vin, Q1in, Q2in, x, y, z - 1) + buoy(c, Vin, Q1in, Q2in, x, y, z)) -
                (ldP(c, phT, x, y, z) - ldP(c, phT, x, y, z - 1)) / dz;
          dw += (spongeLat(c, xmx, xmy) + spongeTop(c, zz)) * ((R)0 - V.z);
          R4 ci = get4(ciMap, (size_t)x + (size_t)nx * y);
          if (ci.x > 0 && !hur) {
            R zr = FABS(zz - ci.z) / ci.w;
            if (zr < 1) {
              R cv = COS((R)1.5707963 * zr);
              dw += UF(c, U_MISC + 3) * ci.x * cv * maxr(ci.y * cv - V.z, 0);
            }
          }
          o.z = V.z + dt * dw;
        } else {
          o.z = 0;
        }
        /* ---- theta + microphysics ---- */
        if (x < nx && y < ny && z < nz) {
          R xmx = ((R)x + (R)0.5) * dx, xmy = ((R)y + (R)0.5) * dx, zz = ((R)z + (R)0.5) * dz;
          R4 A = bsA(c, z), B = bsB(c, z);
          R4 LA = ld2d(c, lsfA, x, y);
          R nw = ld2d(c, lsfB, x, y).w;
          R tth = tgtTh(c, LA, z), tqv = tgtQv(c, LA, z);
          R relax = spongeLat(c, xmx, xmy) + nw * nudR;
          R2 a1 = ldA(c, aux, x, y, z);
          R nux = a1.x * (R)3 + nu0, nuy = a1.y * (R)3 + nu0;
          R th = V.w + dt * (nux * lapH(c, Vin, x, y, z, 3, V.w) * ix2 +
                             nuy * lapZp(c, Vin, x, y, z, 3, V.w, z == 0, z == nz - 1) * iz2 +
                             (relax + spongeTop(c, zz)) * (tth - V.w));
          R4 q1 = ldQ(c, Q1in, x, y, z), q2 = ldQ(c, Q2in, x, y, z);
          R4 e1 = ldQ(c, Q1in, x + 1, y, z), w1 = ldQ(c, Q1in, x - 1, y, z);
          R4 n1 = ldQ(c, Q1in, x, y + 1, z), s1 = ldQ(c, Q1in, x, y - 1, z);
          R4 l1 = {e1.x + w1.x + n1.x + s1.x - (R)4 * q1.x, e1.y + w1.y + n1.y + s1.y - (R)4 * q1.y,
                   e1.z + w1.z + n1.z + s1.z - (R)4 * q1.z, 0};
          R4 dn = z == 0 ? q1 : ldQ(c, Q1in, x, y, z - 1);
          R4 up = z == nz - 1 ? q1 : ldQ(c, Q1in, x, y, z + 1);
          R4 l1z = {dn.x + up.x - (R)2 * q1.x, dn.y + up.y - (R)2 * q1.y, dn.z + up.z - (R)2 * q1.z, 0};
          R qvb = (z == 0 ? A.y : bsA(c, z - 1).y) + (z == nz - 1 ? A.y : bsA(c, z + 1).y) - (R)2 * A.y;
          l1z.x -= qvb;
          q1.x = q1.x + dt * (nux * l1.x * ix2 + nuy * l1z.x * iz2);
          q1.y = q1.y + dt * (nux * l1.y * ix2 + nuy * l1z.y * iz2);
          q1.z = q1.z + dt * (nux * l1.z * ix2 + nuy * l1z.z * iz2);
          q1.x += dt * relax * (tqv - q1.x);
          if (hur) {
            /* hurricane mode (model.js:713-774) */
            th += dt * clampr((tth - V.w) / UF(c, U_HUR2), -UF(c, U_HUR2 + 1), UF(c, U_HUR2 + 1));
            R4 cs = get4(ciMap, (size_t)x + (size_t)nx * y);
            if (cs.x > 0) {
              R zr = FABS(zz - cs.z) / cs.w;
              if (zr < 1) {
                R f = cs.x * COS((R)1.5707963 * zr) * cs.y / (R)8;
                th += dt * (R)1.2e-3 * f;
                q1.x += dt * f / (R)600 * maxr((R)0.97 * qvsat(V.w * B.x, B.y, 0) - q1.x, 0);
              }
            }
            R ex = xmx - UF(c, U_HUR4), ey = xmy - UF(c, U_HUR4 + 1);
            R rS = SQRT(ex * ex + ey * ey);
            R wEnv = sstep(UF(c, U_HUR4 + 3), (R)1.5 * UF(c, U_HUR4 + 3), rS) / (R)3600;
            th += dt * wEnv * (tth - V.w);
            q1.x += dt * wEnv * (tqv - q1.x);
            if (zz < UF(c, U_HUR2 + 2)) {
              R4 s0 = ldV(c, Vin, x, y, 0);
              R ua = (R)0.5 * (s0.x + ldV(c, Vin, x + 1, y, 0).x), va = (R)0.5 * (s0.y + ldV(c, Vin, x, y + 1, 0).y);
              R u1 = SQRT(ua * ua + va * va);
              R e = UF(c, U_STEP + 1) / (R)900, e0 = FLOOR(e), qx = xmx / (R)6000, qy = xmy / (R)6000;
              R ga = gnoise2(qx + (R)37.1 * e0, qy + (R)11.3 * e0);
              R gb = gnoise2(qx + (R)37.1 * (e0 + (R)1), qy + (R)11.3 * (e0 + (R)1));
              R rr = (R)0.5 * mixr(ga, gb, sstep(0, 1, fractr(e)));
              R fu = minr(u1 * u1 / (R)225, 1) * ((R)1 - sstep((R)20, (R)32, u1)) *
                     ((R)1 - sstep(UF(c, U_HUR4 + 2), (R)1.5 * UF(c, U_HUR4 + 2), rS));
              th += dt * UF(c, U_HUR3 + 2) * (R)2 * rr * fu * ((R)1 - zz / UF(c, U_HUR2 + 2));
            }
            if (z == 0) {
              R land = clampr(sfcAt(c, sfc, x, y, 0), 0, 1), sst = sfcAt(c, sfc, x, y, 1);
              R ua = (R)0.5 * (V.x + ldV(c, Vin, x + 1, y, z).x), va = (R)0.5 * (V.y + ldV(c, Vin, x, y + 1, z).y);
              R u1 = SQRT(ua * ua + va * va);
              R ch = UF(c, U_HUR + 1) * UF(c, U_HUR2 + 3) * maxr(u1, UF(c, U_HUR3 + 1)) / dz;
              R ps = maxr(UF(c, U_HUR + 3) + A.z * UF(c, U_CN + 1) * phi[pidx(c, x, y, z)], (R)50000);
              R ths = sst * POW((R)1.0e5 / ps, (R)0.2854);
              th += dt * ch * ((R)1 - land) * (ths - V.w);
              if (c->U[U_HUR5 + 1] > 0.5f) {
                R cdh = cdSfc(c, u1, land);
                th += dt * cdh * u1 * u1 * u1 / (CPD * dz) / B.x;
              }
              R qsS = qvsat(sst, ps, 0), qsL = qvsat(V.w * B.x, ps, 0);
              q1.x += dt * ch * (((R)1 - land) * (qsS - q1.x) + land * UF(c, U_HUR + 2) * maxr(qsL - q1.x, 0));
            }
          }
          q1.x = maxr(q1.x, 0); q1.y = maxr(q1.y, 0); q1.z = maxr(q1.z, 0); q1.w = maxr(q1.w, 0);
          q2.x = maxr(q2.x, 0); q2.y = maxr(q2.y, 0); q2.z = maxr(q2.z, 0); q2.w = maxr(q2.w, 0);
          R pi0 = B.x;
          MState m = {th * pi0, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, maxr(q2.z, 0)};
          micro_block(c, z, &m);
          th = m.T / pi0;
          R envDecay = 0;
          if (hur) {
            R ex = xmx - UF(c, U_HUR4), ey = xmy - UF(c, U_HUR4 + 1);
            envDecay = sstep(UF(c, U_HUR4 + 3), (R)1.5 * UF(c, U_HUR4 + 3), SQRT(ex * ex + ey * ey)) / (R)1800;
          }
          R keep = (R)1 / ((R)1 + (R)2 * dt * spongeLat(c, xmx, xmy) + dt * envDecay);
          m.qc *= keep; m.qi *= keep; m.qr *= keep; m.qs *= keep; m.qg *= keep; m.nr *= keep;
          size_t ii = cidx(c, x, y, z);
          R4 o1 = {maxr(m.qv, 0), maxr(m.qc, 0), maxr(m.qi, 0), maxr(m.qr, 0)};
          R4 o2 = {maxr(m.qs, 0), maxr(m.qg, 0), thm ? maxr(m.nr, 0) : q2.z, q2.w};
          put4(Q1o, ii, o1);
          put4(Q2o, ii, o2);
          o.w = th;
        }
        R sp = SQRT(o.x * o.x + o.y * o.y + o.z * o.z);
        R vcap = UF(c, U_MISC);
        if (sp > vcap) {
          R f = vcap / sp;
          o.x *= f; o.y *= f; o.z *= f;
        }
        if (o.x != o.x || o.y != o.y || o.z != o.z || o.w != o.w) {
          R4 Bz = bsB(c, z);
          o.x = Bz.z; o.y = Bz.w; o.z = 0; o.w = bsA(c, z).x;
        }
        put4(Vout, vidx(c, x, y, z), o);
      }
}

void S_(orc_phys)(const float* U, const float* BS, const R* Vin, const R* Q1in, const R* Q2in, const R* aux,
                  const R* lsfA, const R* lsfB, const R* phT, const R* ciMap, R* Vout, R* Q1o, R* Q2o) {
  S_(orc_phys2)(U, BS, Vin, Q1in, Q2in, aux, lsfA, lsfB, phT, ciMap, NULL, NULL, Vout, Q1o, Q2o);
}

/* ====================== DIV (model.js:910-925) ====================== */
static size_t pidx(const Ctx* c, int x, int y, int z) {
  int p = (int)c->U[U_PAD], nxp = (int)c->U[U_PAD + 1], nyp = (int)c->U[U_PAD + 2];
  return (size_t)(x + p) + (size_t)nxp * ((size_t)(y + p) + (size_t)nyp * (size_t)z);
}
void S_(orc_div)(const float* U, const float* BS, const R* V, R* rhs) {
  Ctx C = mkctx(U, BS), *c = &C;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < c->nz; z++)
    for (int y = 0; y < c->ny; y++)
      for (int x = 0; x < c->nx; x++) {
        R4 p = ldV(c, V, x, y, z);
        R rc = bsA(c, z).z;
        R wt = z == c->nz - 1 ? (R)0 : ldV(c, V, x, y, z + 1).z;
        R wb = z == 0 ? (R)0 : p.z;
        R d = rc * (ldV(c, V, x + 1, y, z).x - p.x + ldV(c, V, x, y + 1, z).y - p.y) / c->dx +
              (rhoW(c, z + 1) * wt - rhoW(c, z) * wb) / c->dz;
        rhs[pidx(c, x, y, z)] = d / c->dt;
      }
}

/* ====================== PROJ (model.js:928-946) ====================== */
static R Pphi(const Ctx* c, const R* phi, int x, int y, int z) {
  int p = (int)c->U[U_PAD];
  return phi[pidx(c, clampi(x, -p, c->nx + p - 1), clampi(y, -p, c->ny + p - 1), clampi(z, 0, c->nz - 1))];
}
void S_(orc_proj)(const float* U, const float* BS, const R* V, const R* phi, R* Vout) {
  Ctx C = mkctx(U, BS), *c = &C;
  R kx = c->dt / c->dx, kz = c->dt / c->dz;
#pragma omp parallel for schedule(static)
  for (int z = 0; z <= c->nz; z++)
    for (int y = 0; y <= c->ny; y++)
      for (int x = 0; x <= c->nx; x++) {
        R4 Vv = ldV(c, V, x, y, z), o = Vv;
        R pc = Pphi(c, phi, x, y, z);
        if (y < c->ny && z < c->nz) o.x = Vv.x - kx * (pc - Pphi(c, phi, x - 1, y, z));
        if (x < c->nx && z < c->nz) o.y = Vv.y - kx * (pc - Pphi(c, phi, x, y - 1, z));
        if (x < c->nx && y < c->ny && z > 0 && z < c->nz) o.z = Vv.z - kz * (pc - Pphi(c, phi, x, y, z - 1));
        else o.z = 0;
        put4(Vout, vidx(c, x, y, z), o);
      }
}

/* ====================== 2-D maps (model.js:1028-1100) ======================
 * reflectivity helpers zRain / zGrau (model.js:94-106) */
static R zRain(const Ctx* c, R qr, R nr, R rho) {
  if (!thompson(c)) return (R)3.63e9 * POW(maxr(rho * qr, 0), (R)1.75);
  if (qr < (R)1e-9) return 0;
  R l = lamR(qr, nr);
  R nb = qr * l * l * l / (R)3141.59;
  return (R)720e18 * rho * nb / (l * l * l * l * l * l);
}
static R zGrau(const Ctx* c, R qg, R rho, R T) {
  if (maritime(c)) return (T < T0K ? (R)1.3e9 : (R)1.0e10) * POW(maxr(rho * qg, 0), (R)1.75);
  if (!thompson(c)) return (T < T0K ? (R)1.2e10 : (R)4.33e10) * POW(maxr(rho * qg, 0), (R)1.75);
  if (qg < (R)1e-9) return 0;
  R l = lamG(qg, rho), l2 = l * l;
  return (R)720e18 * n0G(qg, rho) / (l2 * l2 * l2 * l) * (T < T0K ? (R)0.0666 : (R)0.24);
}
typedef struct { R x, y, z; } MV3;
static MV3 map_cvel(const Ctx* c, const R* V, int i, int j, int k) {
  R4 p = ldV(c, V, i, j, k);
  MV3 r = {(R)0.5 * (p.x + ldV(c, V, i + 1, j, k).x), (R)0.5 * (p.y + ldV(c, V, i, j + 1, k).y),
           (R)0.5 * (p.z + ldV(c, V, i, j, k + 1).z)};
  return r;
}
static R map_dbz(const Ctx* c, const R* V, const R* Q1, const R* Q2, int i, int j, int k) {
  R4 q1 = get4(Q1, cidx(c, i, j, k)), q2 = get4(Q2, cidx(c, i, j, k));
  R rho = bsA(c, k).z, T = ldV(c, V, i, j, k).w * bsB(c, k).x;
  R Z = zRain(c, q1.w, q2.z, rho) + (R)9.80e8 * POW(maxr(rho * q2.x, 0), (R)1.75) * (R)0.224 * (R)4 + zGrau(c, q2.y, rho, T);
  return Z > (R)1e-3 ? (R)10 * LOG(Z) / (R)2.302585 : (R)-30;
}
static R map_zeta(const Ctx* c, const R* V, int i, int j, int k) {
  return (map_cvel(c, V, i + 1, j, k).y - map_cvel(c, V, i - 1, j, k).y - map_cvel(c, V, i, j + 1, k).x +
          map_cvel(c, V, i, j - 1, k).x) / ((R)2 * c->dx);
}
/* maps: 4 planar maps of nx*ny x 4 (m0..m3 as the reference's map buffer); phi padded as the solve's; sfc the
 * sea surface map or NULL */
void S_(orc_maps2)(const float* U, const float* BS, const R* V, const R* Q1, const R* Q2, const R* phi, const float* sfc,
                   R* maps) {
  Ctx C = mkctx(U, BS), *c = &C;
  const size_t n2 = (size_t)c->nx * c->ny;
  const R dz = c->dz;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < c->ny; j++)
    for (int i = 0; i < c->nx; i++) {
      R comp = -30, wmx = 0, uh = 0;
      int k12 = (int)((R)12000 / dz);
      if (k12 > c->nz - 1) k12 = c->nz - 1;
      for (int k = 0; k < k12; k++) {
        R w = map_cvel(c, V, i, j, k).z;
        wmx = maxr(wmx, w);
        comp = maxr(comp, map_dbz(c, V, Q1, Q2, i, j, k));
        R z = ((R)k + (R)0.5) * dz;
        if (z > (R)2000 && z < (R)5000 && w > 0) uh += w * maxr(map_zeta(c, V, i, j, k), 0) * dz;
      }
      int k1 = clampi((int)((R)1000 / dz), 1, c->nz - 2);
      MV3 v0 = map_cvel(c, V, i, j, 0);
      R4 s0 = ldV(c, V, i, j, 0);
      R conv = -((ldV(c, V, i + 1, j, 0).x - s0.x + ldV(c, V, i, j + 1, 0).y - s0.y) / c->dx) * (R)1000;
      size_t idx = (size_t)i + (size_t)c->nx * j;
      R spd = SQRT(v0.x * v0.x + v0.y * v0.y);
      R4 m0 = {comp, map_dbz(c, V, Q1, Q2, i, j, k1), s0.w - bsA(c, 0).x, map_zeta(c, V, i, j, k1)};
      R4 m1 = {wmx, uh, conv, spd};
      MV3 v1 = map_cvel(c, V, i, j, k1);
      R4 m2 = {v0.x, v0.y, v1.x, v1.y};
      R tr = 1, bt = 0;
      for (int k = c->nz - 1; k >= 0; k--) {
        R4 q1 = get4(Q1, cidx(c, i, j, k)), q2 = get4(Q2, cidx(c, i, j, k));
        R dtau = ((R)80 * q1.y + (R)60 * q1.z + (R)25 * q2.x + (R)4 * q2.y + (R)1 * q1.w) * bsA(c, k).z * dz;
        R a = (R)1 - EXP(-dtau);
        bt += tr * a * ldV(c, V, i, j, k).w * bsB(c, k).x;
        tr *= (R)1 - a;
        if (tr < (R)1e-3) break;
      }
      R sst = sfcAt(c, sfc, i, j, 1);
      bt += tr * (sst > (R)200 ? sst : s0.w * bsB(c, 0).x);
      int p = (int)c->U[U_PAD], nxp = (int)c->U[U_PAD + 1];
      R ps = bsA(c, 0).z * UF(c, U_CN + 2) * phi[(size_t)(i + p) + (size_t)nxp * (size_t)(j + p)] / (R)100;
      int k5 = clampi((int)((R)5000 / dz), 1, c->nz - 2);
      R4 m3 = {ps, bt, UF(c, U_HUR2 + 3) * spd, map_zeta(c, V, i, j, k5)};
      put4(maps, idx, m0);
      put4(maps, idx + n2, m1);
      put4(maps, idx + 2 * n2, m2);
      put4(maps, idx + 3 * n2, m3);
    }
}
void S_(orc_maps)(const float* U, const float* BS, const R* V, const R* Q1, const R* Q2, const R* phi, R* maps) {
  S_(orc_maps2)(U, BS, V, Q1, Q2, phi, NULL, maps);
}

/* ====================== hurricane mode: PREP (model.js:996-1026) ======================
 * the winds the advection carries: V less cn.x of the last step's pressure force (phi scaled by cn.y) */
void S_(orc_prep)(const float* U, const float* BS, const R* V, const R* phi, R* Vp) {
  Ctx C = mkctx(U, BS), *c = &C;
  R f = UF(c, U_CN) * UF(c, U_CN + 1);
  R kx = f * c->dt / c->dx, kz = f * c->dt / c->dz;
#pragma omp parallel for schedule(static)
  for (int z = 0; z <= c->nz; z++)
    for (int y = 0; y <= c->ny; y++)
      for (int x = 0; x <= c->nx; x++) {
        R4 Vv = ldV(c, V, x, y, z), o = Vv;
        R pc = Pphi(c, phi, x, y, z);
        if (y < c->ny && z < c->nz) o.x = Vv.x - kx * (pc - Pphi(c, phi, x - 1, y, z));
        if (x < c->nx && z < c->nz) o.y = Vv.y - kx * (pc - Pphi(c, phi, x, y - 1, z));
        if (x < c->nx && y < c->ny && z > 0 && z < c->nz) o.z = Vv.z - kz * (pc - Pphi(c, phi, x, y, z - 1));
        put4(Vp, vidx(c, x, y, z), o);
      }
}

/* ====================== hurricane mode: the 2-dx filter FILT (model.js:948-994) ====================== */
static R4 bih(R4 c0, R4 xp, R4 xm, R4 yp, R4 ym, R4 pp, R4 pm, R4 mp, R4 mm, R4 x2p, R4 x2m, R4 y2p, R4 y2m) {
  R4 r;
#define BIH(ch) ((R)20 * c0.ch - (R)8 * (xp.ch + xm.ch + yp.ch + ym.ch) + (R)2 * (pp.ch + pm.ch + mp.ch + mm.ch) + x2p.ch + \
                 x2m.ch + y2p.ch + y2m.ch)
  r.x = BIH(x); r.y = BIH(y); r.z = BIH(z); r.w = BIH(w);
#undef BIH
  return r;
}
void S_(orc_filt)(const float* U, const float* BS, const R* V, const R* Q1, const R* Q2, R* Vo, R* Q1o, R* Q2o) {
  Ctx C = mkctx(U, BS), *c = &C;
  R a = UF(c, U_HUR5 + 2);
#pragma omp parallel for schedule(static)
  for (int z = 0; z <= c->nz; z++)
    for (int y = 0; y <= c->ny; y++)
      for (int x = 0; x <= c->nx; x++) {
#define NB(L, F) L(c, F, x, y, z), L(c, F, x + 1, y, z), L(c, F, x - 1, y, z), L(c, F, x, y + 1, z), L(c, F, x, y - 1, z), \
                 L(c, F, x + 1, y + 1, z), L(c, F, x + 1, y - 1, z), L(c, F, x - 1, y + 1, z), L(c, F, x - 1, y - 1, z),      \
                 L(c, F, x + 2, y, z), L(c, F, x - 2, y, z), L(c, F, x, y + 2, z), L(c, F, x, y - 2, z)
        R4 v = ldV(c, V, x, y, z), b = bih(NB(ldV, V));
        R4 o = {v.x - a * b.x, v.y - a * b.y, v.z - a * b.z, v.w - a * b.w};
        if (z == 0 || z >= c->nz) o.z = 0;
        put4(Vo, vidx(c, x, y, z), o);
        if (x < c->nx && y < c->ny && z < c->nz) {
          R4 q = ldQ(c, Q1, x, y, z), bq = bih(NB(ldQ, Q1));
          R4 o1 = {maxr(q.x - a * bq.x, 0), maxr(q.y - a * bq.y, 0), maxr(q.z - a * bq.z, 0), maxr(q.w - a * bq.w, 0)};
          R4 r = ldQ(c, Q2, x, y, z), br = bih(NB(ldQ, Q2));
          R4 o2 = {maxr(r.x - a * br.x, 0), maxr(r.y - a * br.y, 0), maxr(r.z - a * br.z, 0), maxr(r.w - a * br.w, 0)};
          put4(Q1o, cidx(c, x, y, z), o1);
          put4(Q2o, cidx(c, x, y, z), o2);
        }
#undef NB
      }
}
