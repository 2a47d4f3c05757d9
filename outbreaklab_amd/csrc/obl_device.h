// Device-side helpers for the model-step kernels (gfx950).
//
// The uniform block is the reference's 19-vec4f MODEL_FIELDS block
// (model.js:23-37), passed to every kernel by value (304 B of kernel
// arguments), so a step needs no uniform upload and no synchronisation.
//
// Field layouts in HBM (x fastest, as the reference's 3-D textures):
//   V       float4 (nx+1)(ny+1)(nz+1)   u at x-face, v at y-face, w at z-face, theta
//   Q1, Q2  float4 nx*ny*nz             qv qc qi qr / qs qg Nr spare
//   aux     float2 cells                nu_h, nu_v
//   phT     float cells
//   lsfA/B, ciMap  float4 nx*ny
//   rhs/phi float on the padded solve grid, pidx = (x+p) + nxp((y+p) + nyp z)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace obl {

enum UIdx {
  U_DIMS = 0, U_DZV, U_STEP, U_SPONGE, U_TURB, U_MICRO, U_MISC, U_COR, U_CI, U_PAD, U_MICRO2,
  U_HUR, U_HUR2, U_HUR3, U_HUR4, U_HUR5, U_VTX, U_VTX2, U_CN, U_COUNT
};

struct MU {
  float4 v[U_COUNT];
};

struct Dims {
  int nx, ny, nz;
  float dx, dz, dt;
  // computed once per thread by dims(): the Courant factors dt / dx, dt / dz, and the reciprocals that the loops
  // multiply by (rmul). x / b with the WGSL-accuracy division of these kernels is x * rcp(b) rounded (for
  // |b| <= 2^96, any grid), so rmul(x, rcp(b)) gives the same bits without recomputing rcp(b) and its range check
  // at every use.
  float kx, kz, rdx, rdz, r2dx, r2dz, rdt;
};

// x * r rounded on its own, as a quotient is: the empty asm keeps the product from being fused into a following
// add (which would round once for both)
__device__ __forceinline__ float rmul(float x, float r) {
  float p = x * r;
  asm("" : "+v"(p));
  return p;
}

// a wave-uniform float moved to a scalar register
__device__ __forceinline__ float sreg(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

__device__ __forceinline__ Dims dims(const MU& u) {
  Dims d;
  d.nx = (int)u.v[U_DIMS].x; d.ny = (int)u.v[U_DIMS].y; d.nz = (int)u.v[U_DIMS].z;
  d.dx = u.v[U_DIMS].w; d.dz = u.v[U_DZV].x; d.dt = u.v[U_STEP].x;
  // (kept in scalar registers: uniform values made by vector instructions would otherwise hold a vector register
  // each for the whole kernel)
  d.kx = sreg(d.dt / d.dx); d.kz = sreg(d.dt / d.dz);
  d.rdx = sreg(1.0f / d.dx); d.rdz = sreg(1.0f / d.dz); d.r2dx = sreg(1.0f / (2.0f * d.dx));
  d.r2dz = sreg(1.0f / (2.0f * d.dz)); d.rdt = sreg(1.0f / d.dt);
  return d;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }
// WGSL clamp(e, lo, hi) = min(max(e, lo), hi); mix(a, b, t) = a (1 - t) + b t
__device__ __forceinline__ float clampf_(float e, float lo, float hi) { return fminf(fmaxf(e, lo), hi); }
// Clamp of a finite coordinate in one v_med3_f32 (the trace points of the advection kernels).
__device__ __forceinline__ float clampt_(float e, float lo, float hi) { return __builtin_amdgcn_fmed3f(e, lo, hi); }
// Integer floor in one v_cvt_flr_i32_f32 (the compiler emits v_floor + v_cvt); with
// __builtin_amdgcn_fractf for the fraction, a texel coordinate splits in two instructions.
__device__ __forceinline__ int floori_(float t) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(t));
  return r;
}
// min / max of texels read from LDS, as the bare VALU instructions: fminf / fmaxf on loaded values
// first quiet possible signalling NaNs (one extra v_max_f32 per operand), which the limiter of the
// advection kernels, eight texels per channel, does not need.
__device__ __forceinline__ float vmin_(float a, float b) {
  float r;
  asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float vmax_(float a, float b) {
  float r;
  asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float vmin3_(float a, float b, float c) {
  float r;
  asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
__device__ __forceinline__ float vmax3_(float a, float b, float c) {
  float r;
  asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
// clamp to a limiter's [min, max] (lo <= hi): one v_med3_f32
__device__ __forceinline__ float vclamp_(float x, float2 m) {
  float r;
  asm("v_med3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(m.x), "v"(m.y));
  return r;
}
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float sstep(float a, float b, float x) {
  float t = clampf_((x - a) / (b - a), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

// Base-state table: per level (theta0, qv0, rho_c, rho_w bottom), (pi0, p0, u0, v0)
__device__ __forceinline__ float4 bsA(const float4* __restrict__ BS, int k, int nz) { return BS[clampi(k, 0, nz - 1) * 2]; }
__device__ __forceinline__ float4 bsB(const float4* __restrict__ BS, int k, int nz) { return BS[clampi(k, 0, nz - 1) * 2 + 1]; }
__device__ __forceinline__ float rhoW(const float4* __restrict__ BS, int k, int nz) { return BS[clampi(k, 0, nz) * 2].w; }

// Texel offsets are 32-bit: a field holds fewer than 2^28 texels and every factor of an offset
// is below 2^24 (checked when a model is created), so offsets are built with 24-bit multiplies and
// an access is a uniform base plus a 32-bit byte offset instead of 64-bit address arithmetic.
template <class T>
__device__ __forceinline__ T ldx(const T* p, uint32_t i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + i * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ void stx(T* p, uint32_t i, const T& v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + i * (uint32_t)sizeof(T)) = v;
}
// stx as a non-temporal store (streamed through the caches rather than kept in them): for outputs that the launch
// itself does not read again, so that its halo re-reads keep their L2 lines (the physics kernel 2.23 -> 2.10 ms)
template <class T>
__device__ __forceinline__ void stnt(T* p, uint32_t i, const T& v) {
  T* a = reinterpret_cast<T*>(reinterpret_cast<char*>(p) + i * (uint32_t)sizeof(T));
  if constexpr (sizeof(T) == sizeof(float)) {
    __builtin_nontemporal_store(v, a);
  } else {
    typedef float nv __attribute__((ext_vector_type(sizeof(T) / sizeof(float))));
    __builtin_nontemporal_store(__builtin_bit_cast(nv, v), reinterpret_cast<nv*>(a));
  }
}
// channel ch of float4 texel i
__device__ __forceinline__ float ldxc(const float4* p, uint32_t i, int ch) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + i * 16u + (uint32_t)ch * 4u);
}
__device__ __forceinline__ uint32_t off3(int W, int H, int x, int y, int z) {
  return (uint32_t)x + __umul24((uint32_t)W, (uint32_t)y + __umul24((uint32_t)H, (uint32_t)z));
}
__device__ __forceinline__ uint32_t vidx(const Dims& d, int x, int y, int z) { return off3(d.nx + 1, d.ny + 1, x, y, z); }
__device__ __forceinline__ uint32_t cidx(const Dims& d, int x, int y, int z) { return off3(d.nx, d.ny, x, y, z); }
__device__ __forceinline__ float4 ldV(const float4* __restrict__ V, const Dims& d, int x, int y, int z) {
  return ldx(V, vidx(d, clampi(x, 0, d.nx), clampi(y, 0, d.ny), clampi(z, 0, d.nz)));
}
__device__ __forceinline__ float ldVc(const float4* __restrict__ V, const Dims& d, int x, int y, int z, int ch) {
  return ldxc(V, vidx(d, clampi(x, 0, d.nx), clampi(y, 0, d.ny), clampi(z, 0, d.nz)), ch);
}
__device__ __forceinline__ float4 ldQ(const float4* __restrict__ Q, const Dims& d, int x, int y, int z) {
  return ldx(Q, cidx(d, clampi(x, 0, d.nx - 1), clampi(y, 0, d.ny - 1), clampi(z, 0, d.nz - 1)));
}
__device__ __forceinline__ float4 ld2d(const float4* __restrict__ M, const Dims& d, int i, int j) {
  return ldx(M, (uint32_t)clampi(i, 0, d.nx - 1) + __umul24((uint32_t)d.nx, (uint32_t)clampi(j, 0, d.ny - 1)));
}

__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float t) {
  return make_float4(lerpf(a.x, b.x, t), lerpf(a.y, b.y, t), lerpf(a.z, b.z, t), lerpf(a.w, b.w, t));
}

// Trilinear, clamp-to-edge sample of a W x H x D float4 texture at texel
// coordinates (u*S - 1/2); lerps x, then y, then z.
__device__ __forceinline__ float4 sample4(const float4* __restrict__ T, int W, int H, int D, float tx, float ty,
                                          float tz) {
  float fx = floorf(tx), fy = floorf(ty), fz = floorf(tz);
  float ax = tx - fx, ay = ty - fy, az = tz - fz;
  int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
  int x1 = clampi(x0 + 1, 0, W - 1), y1 = clampi(y0 + 1, 0, H - 1), z1 = clampi(z0 + 1, 0, D - 1);
  x0 = clampi(x0, 0, W - 1); y0 = clampi(y0, 0, H - 1); z0 = clampi(z0, 0, D - 1);
  size_t r00 = (size_t)W * ((size_t)y0 + (size_t)H * z0), r10 = (size_t)W * ((size_t)y1 + (size_t)H * z0);
  size_t r01 = (size_t)W * ((size_t)y0 + (size_t)H * z1), r11 = (size_t)W * ((size_t)y1 + (size_t)H * z1);
  float4 c00 = lerp4(T[r00 + x0], T[r00 + x1], ax);
  float4 c10 = lerp4(T[r10 + x0], T[r10 + x1], ax);
  float4 c01 = lerp4(T[r01 + x0], T[r01 + x1], ax);
  float4 c11 = lerp4(T[r11 + x0], T[r11 + x1], ax);
  return lerp4(lerp4(c00, c10, ay), lerp4(c01, c11, ay), az);
}
// one channel of the same sample (loads 8 scalars instead of 8 texels)
__device__ __forceinline__ float sample1(const float4* __restrict__ T4, int W, int H, int D, float tx, float ty,
                                         float tz, int ch) {
  const float* T = reinterpret_cast<const float*>(T4);
  float fx = floorf(tx), fy = floorf(ty), fz = floorf(tz);
  float ax = tx - fx, ay = ty - fy, az = tz - fz;
  int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
  int x1 = clampi(x0 + 1, 0, W - 1), y1 = clampi(y0 + 1, 0, H - 1), z1 = clampi(z0 + 1, 0, D - 1);
  x0 = clampi(x0, 0, W - 1); y0 = clampi(y0, 0, H - 1); z0 = clampi(z0, 0, D - 1);
  size_t r00 = (size_t)W * ((size_t)y0 + (size_t)H * z0), r10 = (size_t)W * ((size_t)y1 + (size_t)H * z0);
  size_t r01 = (size_t)W * ((size_t)y0 + (size_t)H * z1), r11 = (size_t)W * ((size_t)y1 + (size_t)H * z1);
  float c00 = lerpf(T[4 * (r00 + x0) + ch], T[4 * (r00 + x1) + ch], ax);
  float c10 = lerpf(T[4 * (r10 + x0) + ch], T[4 * (r10 + x1) + ch], ax);
  float c01 = lerpf(T[4 * (r01 + x0) + ch], T[4 * (r01 + x1) + ch], ax);
  float c11 = lerpf(T[4 * (r11 + x0) + ch], T[4 * (r11 + x1) + ch], ax);
  return lerpf(lerpf(c00, c10, ay), lerpf(c01, c11, ay), az);
}

// ---------------- noise (PCG hash, gradient noise) ----------------
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
  uint32_t s = v * 747796405u + 2891336453u;
  uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (w >> 22u) ^ w;
}
__device__ __forceinline__ uint32_t hash2i(int x, int y) { return pcg((uint32_t)x ^ pcg((uint32_t)y + 0x9e3779b9u)); }
__device__ __forceinline__ uint32_t hash3i(int x, int y, int z) {
  return pcg((uint32_t)x ^ pcg((uint32_t)y ^ pcg((uint32_t)z + 0x7f4a7c15u)));
}
__device__ __forceinline__ float rnd3(int x, int y, int z) {
  return (float)(hash3i(x, y, z) >> 8u) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ float fadef(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
__device__ __forceinline__ float gnoise2(float px, float py) {
  float fx0 = floorf(px), fy0 = floorf(py);
  int ix = (int)fx0, iy = (int)fy0;
  float fx = px - fx0, fy = py - fy0;
  float ux = fadef(fx), uy = fadef(fy);
  const float K = (float)(6.2831853 / 16777216.0);
  float a0 = (float)(hash2i(ix, iy) >> 8u) * K, a1 = (float)(hash2i(ix + 1, iy) >> 8u) * K;
  float a2 = (float)(hash2i(ix, iy + 1) >> 8u) * K, a3 = (float)(hash2i(ix + 1, iy + 1) >> 8u) * K;
  float a = cosf(a0) * fx + sinf(a0) * fy;
  float b = cosf(a1) * (fx - 1.0f) + sinf(a1) * fy;
  float c = cosf(a2) * fx + sinf(a2) * (fy - 1.0f);
  float dd = cosf(a3) * (fx - 1.0f) + sinf(a3) * (fy - 1.0f);
  return mixf(mixf(a, b, ux), mixf(c, dd, ux), uy) * 1.4142f;
}
__device__ __forceinline__ float fbm2(float px, float py, int oct) {
  float a = 0.5f, s = 0.0f;
  for (int i = 0; i < oct; i++) {
    s += a * gnoise2(px, py);
    float nx = 1.6f * px + -1.2f * py + 17.1f;
    float ny = 1.2f * px + 1.6f * py + 3.7f;
    px = nx; py = ny;
    a *= 0.5f;
  }
  return s;
}

// ---------------- transcendentals ----------------
// pow and exp on the hardware v_log_f32 / v_exp_f32 (log2, 2^x): pw(x, y) = 2^(y log2 x), which is
// how WGSL specifies pow's accuracy (inherited from exp2(y * log2(x))); ex(x) = 2^(x log2 e).
// Arguments here are >= 0 (pw) and moderate (ex); pw(0, y > 0) = 0.
__device__ __forceinline__ float pw(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
__device__ __forceinline__ float ex(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

// ---------------- thermodynamics and fall speeds ----------------
#define OBL_T0K 273.15f
#define OBL_CPD 1005.7f
#define OBL_RVC 461.5f
#define OBL_LVC 2.501e6f
#define OBL_LFC 3.337e5f
#define OBL_LSC 2.834e6f
__device__ __forceinline__ float esl(float T) { return 611.2f * ex(17.67f * (T - OBL_T0K) / (T - 29.65f)); }
__device__ __forceinline__ float esi(float T) { return 611.2f * ex(21.8745584f * (T - 273.16f) / (T - 7.66f)); }
__device__ __forceinline__ float fIce(float T) { return clampf_((OBL_T0K - T) / 40.0f, 0.0f, 1.0f); }
__device__ __forceinline__ float qvsat(float T, float p, float f) {
  float e = mixf(esl(T), esi(T), f);
  return 0.622f * e / fmaxf(p - e, 100.0f);
}
__device__ __forceinline__ float vtR(float q, float rho, float rsfc) {
  float f = sqrtf(rsfc / rho);
  return fminf(36.34f * pw(fmaxf(0.001f * rho * q, 0.0f), 0.1364f) * f, 10.0f * f);
}
__device__ __forceinline__ float vtS(float q, float rho, float rsfc) {
  return clampf_(1.1f * pw(fmaxf(1000.0f * rho * q, 1e-6f), 0.11f), 0.3f, 2.2f) * sqrtf(rsfc / rho);
}
__device__ __forceinline__ float vtG(float q, float rho, float rsfc) {
  float v = clampf_(9.0f * pw(fmaxf(1000.0f * rho * q, 1e-6f), 0.42f), 1.0f, 55.0f) * sqrtf(rsfc / rho);
  float r = q / 0.009f;
  return fminf(v * (1.0f + r * r), 65.0f);
}
__device__ __forceinline__ float lamR(float q, float n) {
  return clampf_(pw(3.14159f * 1000.0f * fmaxf(n, 1e-3f) / fmaxf(q, 1e-12f), 0.33333f), 1000.0f, 1.0e5f);
}
__device__ __forceinline__ float2 vtRT(float q, float n, float rho, float rsfc) {
  float l = pw(lamR(q, n), -0.8f);
  float s = sqrtf(rsfc / rho);
  return make_float2(fminf(2503.0f * l, 10.0f) * s, fminf(784.0f * l, 10.0f) * s);
}
__device__ __forceinline__ float n0G(float q, float rho) { return clampf_(200.0f / fmaxf(rho * q, 1e-9f), 1e4f, 3e6f); }
__device__ __forceinline__ float lamG(float q, float rho) {
  return pw(3.14159f * 500.0f * n0G(q, rho) / fmaxf(rho * q, 1e-12f), 0.25f);
}
__device__ __forceinline__ float vtGT(float q, float rho, float rsfc) {
  return fminf(1503.0f * pw(lamG(q, rho), -0.89f), 25.0f) * sqrtf(rsfc / rho);
}
__device__ __forceinline__ float2 vtRain(bool thm, float qr, float nr, float rho, float rsfc) {
  if (thm) return vtRT(qr, nr, rho, rsfc);
  float v = vtR(qr, rho, rsfc);
  return make_float2(v, v);
}
__device__ __forceinline__ float vtGrau(bool thm, float qg, float rho, float rsfc) {
  return thm ? vtGT(qg, rho, rsfc) : vtG(qg, rho, rsfc);
}
// ---- hurricane mode (model.js:88-119): a warm sea under the domain ----
__host__ __device__ __forceinline__ bool hur_on(const MU& u) { return u.v[U_HUR].x > 0.5f; }
// maritime (soft, low-density) graupel with the Lin scheme
__device__ __forceinline__ bool maritime_of(const MU& u) { return u.v[U_HUR3].w > 0.5f && !(u.v[U_MICRO2].z > 0.5f); }
__device__ __forceinline__ float vtGM(float q, float rho, float rsfc) {
  return clampf_(5.0f * pw(fmaxf(1000.0f * rho * q, 1e-6f), 0.25f), 1.0f, 9.0f) * sqrtf(rsfc / rho);
}
__device__ __forceinline__ float vtGrauM(bool thm, bool mar, float qg, float rho, float rsfc) {
  return mar ? vtGM(qg, rho, rsfc) : vtGrau(thm, qg, rho, rsfc);
}
// settling cloud ice (hurricane mode only)
__device__ __forceinline__ float vtI(float q, float rho) {
  return q > 1e-9f ? fminf(3.29f * pw(fmaxf(rho * q, 1e-12f), 0.16f), 1.5f) : 0.0f;
}
// drag coefficient of the lowest-level wind u1: over the sea rising with the wind up to a cap, over land the log law
__device__ __forceinline__ float cdSfc(const MU& u, float u1, float land) {
  float f = u.v[U_HUR2].w;
  float cs = fminf((0.75f + 0.067f * f * u1) * 1e-3f, u.v[U_HUR5].x) * f * f;
  float t = clampf_(land, 0.0f, 1.0f);
  return cs * (1.0f - t) + u.v[U_HUR3].x * t;
}

// ---------------- large-scale target ----------------
__device__ __forceinline__ float tprof(float z, float H) { return 1.0f - sstep(0.7f * H, 1.15f * H, z); }

}  // namespace obl
