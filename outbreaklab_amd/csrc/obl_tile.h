// LDS plane rings for the z-marching stencil kernels (gfx950).
//
// A block owns a TX x TY column of points and walks it from the ground to the
// lid. For every field it reads, the block keeps three z-planes (k-1, k, k+1)
// of a (TX+2) x (TY+2) tile -- its points plus a one-point halo -- in LDS,
// stored channel by channel (SoA) so that a wave's per-channel gathers hit
// distinct banks. Each plane is fetched from HBM once per block as whole
// float4 texels, with the reference's clamp-to-edge addressing applied at load
// time: a tile slot outside the field holds the clamped texel, which is what
// textureLoad(clamp(c)) and the clamp-to-edge sampler return.
//
// Reads inside the window come from LDS; anything outside it (a departure
// point more than one cell away, or a far neighbour) falls back to the same
// clamped read from global memory, so results never depend on the window.
#pragma once
#include "obl_device.h"

namespace obl {

constexpr int TX = 32, TY = 8, TPW = TX + 2, TPH = TY + 2, TPS = TPW * TPH;
constexpr int TILE_THREADS = TX * TY;
constexpr int RING_FLOATS = 3 * 4 * TPS;  // 3 planes x 4 channels

typedef __attribute__((address_space(3))) float lds_float;

// Reads of a RingW view are known to fall in the ring's window: no window tests, no fallback,
// and the plane offsets of k-1, k, k+1 are selects rather than a modulo.
struct RingW {
  const lds_float* s;
  int x0, y0, zc;
  int om, oc, op;  // float offsets of planes zc-1, zc, zc+1
  __device__ __forceinline__ int plane(int k) const { return k < zc ? om : (k == zc ? oc : op); }
  __device__ __forceinline__ int sidx(int i, int j, int k, int ch) const {
    return plane(k) + ch * TPS + __mul24(j - y0 + 1, TPW) + (i - x0 + 1);
  }
  __device__ __forceinline__ float ldc(int i, int j, int k, int ch) const { return s[sidx(i, j, k, ch)]; }
  __device__ __forceinline__ float4 ld4(int i, int j, int k) const {
    int b = sidx(i, j, k, 0);
    return make_float4(s[b], s[b + TPS], s[b + 2 * TPS], s[b + 3 * TPS]);
  }
};

// all lanes of the wave agree (wave-uniform branch condition)
__device__ __forceinline__ bool wave_all(bool p) { return __all((int)p) != 0; }

struct Ring {
  float* s;          // LDS, [3 planes][4 channels][TPH][TPW]
  const float4* g;   // the field in HBM
  int W, H, D;       // texel extents: W x H x D
  int x0, y0, zc;    // tile origin (first interior point) and the centre plane

  int om = 0, oc = 0, op = 0;  // float offsets of planes zc-1, zc, zc+1 (set by center)

  __device__ __forceinline__ static int slot(int k) { return (k + 3 * 4096) % 3; }
  __device__ __forceinline__ void center(int k) {
    zc = k;
    om = slot(k - 1) * 4 * TPS;
    oc = slot(k) * 4 * TPS;
    op = slot(k + 1) * 4 * TPS;
  }
  // view for reads known to lie within one point of the tile and of plane zc
  __device__ __forceinline__ RingW w() const { return RingW{(const lds_float*)s, x0, y0, zc, om, oc, op}; }
  __device__ __forceinline__ bool in(int i, int j, int k) const {
    return (unsigned)(i - x0 + 1) < (unsigned)TPW && (unsigned)(j - y0 + 1) < (unsigned)TPH &&
           (unsigned)(k - zc + 1) < 3u;
  }
  __device__ __forceinline__ uint32_t gidx(int i, int j, int k) const {
    return off3(W, H, clampi(i, 0, W - 1), clampi(j, 0, H - 1), clampi(k, 0, D - 1));
  }
  __device__ __forceinline__ int sidx(int i, int j, int k, int ch) const {
    return (slot(k) * 4 + ch) * TPS + __mul24(j - y0 + 1, TPW) + (i - x0 + 1);
  }
  __device__ __forceinline__ float ldc(int i, int j, int k, int ch) const {
    if (in(i, j, k)) return s[sidx(i, j, k, ch)];
    return ldxc(g, gidx(i, j, k), ch);
  }
  __device__ __forceinline__ float4 ld4(int i, int j, int k) const {
    if (in(i, j, k)) {
      int b = sidx(i, j, k, 0);
      return make_float4(s[b], s[b + TPS], s[b + 2 * TPS], s[b + 3 * TPS]);
    }
    return ldx(g, gidx(i, j, k));
  }
  // in-plane texel offset of halo slot t (clamped), and the offset of plane k (clamped)
  __device__ __forceinline__ uint32_t tofs(int t) const {
    return (uint32_t)clampi(x0 - 1 + t % TPW, 0, W - 1) + __umul24((uint32_t)W, (uint32_t)clampi(y0 - 1 + t / TPW, 0, H - 1));
  }
  __device__ __forceinline__ uint32_t pofs(int k) const {
    return __umul24((uint32_t)clampi(k, 0, D - 1), (uint32_t)(W * H));
  }
  // software-pipelined plane load: fetch plane k (clamped) into registers early, store it to
  // its slot after the compute that still reads the slot's old plane
  struct Pre {
    float4 v[(TPS + TILE_THREADS - 1) / TILE_THREADS];
  };
  // Every lane loads (lanes past the tile repeat its last texel): no divergent branch around the
  // loads, so the wait for a level's own reads need not wait for these.
  __device__ __forceinline__ void fetch(int k, int tid, Pre& r) const {
    const uint32_t po = pofs(k);
#pragma unroll
    for (int q = 0; q < (TPS + TILE_THREADS - 1) / TILE_THREADS; q++) {
      int t = min(tid + q * TILE_THREADS, TPS - 1);
      r.v[q] = ldx(g, tofs(t) + po);
    }
  }
  __device__ __forceinline__ void store(int k, int tid, const Pre& r) {
    int sl = slot(k);
#pragma unroll
    for (int q = 0; q < (TPS + TILE_THREADS - 1) / TILE_THREADS; q++) {
      int t = tid + q * TILE_THREADS;
      if (t < TPS) {
        float* p = s + sl * 4 * TPS + t;
        p[0] = r.v[q].x;
        p[TPS] = r.v[q].y;
        p[2 * TPS] = r.v[q].z;
        p[3 * TPS] = r.v[q].w;
      }
    }
  }
  // cooperative load of plane k (clamped) into its slot
  __device__ __forceinline__ void load(int k, int tid) {
    const uint32_t po = pofs(k);
    int sl = slot(k);
    for (int t = tid; t < TPS; t += TILE_THREADS) {
      float4 v = ldx(g, tofs(t) + po);
      float* p = s + sl * 4 * TPS + t;
      p[0] = v.x;
      p[TPS] = v.y;
      p[2 * TPS] = v.z;
      p[3 * TPS] = v.w;
    }
  }
};

// One z-plane of a (TX+2) x (TY+2) tile in LDS, for kernels that run one block per plane tile
// (a 3-D grid) and read x / y neighbours many times but z neighbours rarely. Contract: in-plane
// reads stay within one point of the block's tile (the halo), so whether a read hits LDS depends
// only on its z, which the compiler resolves statically at every call site: in-plane reads are
// LDS reads, the others clamped global reads.
struct Plane {
  lds_float* s;      // LDS, [4 channels][TPH][TPW]
  const float4* g;
  int W, H, D;
  int x0, y0, zc;
  __device__ __forceinline__ bool in(int, int, int k) const { return k == zc; }
  __device__ __forceinline__ uint32_t gidx(int i, int j, int k) const {
    return off3(W, H, clampi(i, 0, W - 1), clampi(j, 0, H - 1), clampi(k, 0, D - 1));
  }
  __device__ __forceinline__ int sidx(int i, int j, int, int ch) const {
    return ch * TPS + __mul24(j - y0 + 1, TPW) + (i - x0 + 1);
  }
  __device__ __forceinline__ float ldc(int i, int j, int k, int ch) const {
    if (in(i, j, k)) return s[sidx(i, j, k, ch)];
    return ldxc(g, gidx(i, j, k), ch);
  }
  __device__ __forceinline__ float4 ld4(int i, int j, int k) const {
    if (in(i, j, k)) {
      int b = sidx(i, j, k, 0);
      return make_float4(s[b], s[b + TPS], s[b + 2 * TPS], s[b + 3 * TPS]);
    }
    return ldx(g, gidx(i, j, k));
  }
  // cooperative load of the plane (clamped texels in the halo); the caller synchronises
  __device__ __forceinline__ void load(int tid) {
    int kk = clampi(zc, 0, D - 1);
    for (int t = tid; t < TPS; t += TILE_THREADS) {
      int ti = t % TPW, tj = t / TPW;
      float4 v = ldx(g, off3(W, H, clampi(x0 - 1 + ti, 0, W - 1), clampi(y0 - 1 + tj, 0, H - 1), kk));
      s[t] = v.x;
      s[TPS + t] = v.y;
      s[2 * TPS + t] = v.z;
      s[3 * TPS + t] = v.w;
    }
  }
};

// The same interface over HBM only (clamped reads, no window): for fields a kernel reads at a
// few points, and for the fallback path.
struct GAcc {
  const float4* g;
  int W, H, D;
  float* s;
  __device__ __forceinline__ bool in(int, int, int) const { return false; }
  __device__ __forceinline__ int sidx(int, int, int, int) const { return 0; }
  __device__ __forceinline__ uint32_t gidx(int i, int j, int k) const {
    return off3(W, H, clampi(i, 0, W - 1), clampi(j, 0, H - 1), clampi(k, 0, D - 1));
  }
  __device__ __forceinline__ float ldc(int i, int j, int k, int ch) const { return ldxc(g, gidx(i, j, k), ch); }
  __device__ __forceinline__ float4 ld4(int i, int j, int k) const { return ldx(g, gidx(i, j, k)); }
};

// Out-of-window paths, kept out of line: they are rare, and inlined they would hold registers
// across every sample of the fast path.
__device__ __noinline__ static float samp1_global(const float4* g4, int W, int H, int D, int i0, int j0, int k0, float ax,
                                                  float ay, float az, int ch) {
  int x0 = clampi(i0, 0, W - 1), x1 = clampi(i0 + 1, 0, W - 1);
  int y0 = clampi(j0, 0, H - 1), y1 = clampi(j0 + 1, 0, H - 1);
  int z0 = clampi(k0, 0, D - 1), z1 = clampi(k0 + 1, 0, D - 1);
  float c00 = lerpf(ldxc(g4, off3(W, H, x0, y0, z0), ch), ldxc(g4, off3(W, H, x1, y0, z0), ch), ax);
  float c10 = lerpf(ldxc(g4, off3(W, H, x0, y1, z0), ch), ldxc(g4, off3(W, H, x1, y1, z0), ch), ax);
  float c01 = lerpf(ldxc(g4, off3(W, H, x0, y0, z1), ch), ldxc(g4, off3(W, H, x1, y0, z1), ch), ax);
  float c11 = lerpf(ldxc(g4, off3(W, H, x0, y1, z1), ch), ldxc(g4, off3(W, H, x1, y1, z1), ch), ax);
  return lerpf(lerpf(c00, c10, ay), lerpf(c01, c11, ay), az);
}
__device__ __noinline__ static float2 minmax1_global(const float4* g4, int W, int H, int D, int i0, int j0, int k0,
                                                     int ch) {
  float lo = 1e30f, hi = -1e30f;
  for (int q = 0; q < 8; q++) {
    int i = clampi(i0 + (q & 1), 0, W - 1), j = clampi(j0 + ((q >> 1) & 1), 0, H - 1), k = clampi(k0 + (q >> 2), 0, D - 1);
    float s = ldxc(g4, off3(W, H, i, j, k), ch);
    lo = fminf(lo, s);
    hi = fmaxf(hi, s);
  }
  return make_float2(lo, hi);
}

// The LDS address of a texel box's first texel, made opaque to the compiler (an empty asm, no instruction): the
// box's other reads then carry their one-texel and one-row offsets in the read instructions themselves (two
// ds_read2_b32 per plane from one address), instead of one address add per pair of reads.
__device__ __forceinline__ const lds_float* lds_box(const lds_float* p) {
  asm("" : "+v"(p));
  return p;
}
// the eight texels of a box: rows of `row` floats, the upper plane dz floats above the lower
struct Box8 {
  float c000, c100, c010, c110, c001, c101, c011, c111;
};
__device__ __forceinline__ Box8 box8(const lds_float* s, int b, int dz, int row) {
  const lds_float* p = lds_box(s + b);
  const lds_float* q = p + dz;
  return Box8{p[0], p[1], p[row], p[row + 1], q[0], q[1], q[row], q[row + 1]};
}

// Window-guaranteed forms: the 2 x 2 x 2 texel box of the sample lies in the window, so
// k0 is zc - 1 or zc.
__device__ __forceinline__ float samp1(const RingW& a, float tx, float ty, float tz, int ch) {
  float ax = __builtin_amdgcn_fractf(tx), ay = __builtin_amdgcn_fractf(ty), az = __builtin_amdgcn_fractf(tz);
  int i0 = floori_(tx), j0 = floori_(ty), k0 = floori_(tz);
  bool lo = k0 < a.zc;
  int b = (lo ? a.om : a.oc) + ch * TPS + __mul24(j0 - a.y0 + 1, TPW) + (i0 - a.x0 + 1);
  int dz = lo ? a.oc - a.om : a.op - a.oc;
  const Box8 c = box8(a.s, b, dz, TPW);
  float c00 = lerpf(c.c000, c.c100, ax), c10 = lerpf(c.c010, c.c110, ax);
  float c01 = lerpf(c.c001, c.c101, ax), c11 = lerpf(c.c011, c.c111, ax);
  return lerpf(lerpf(c00, c10, ay), lerpf(c01, c11, ay), az);
}
__device__ __forceinline__ float2 minmax1(const RingW& a, int i0, int j0, int k0, int ch) {
  bool lo = k0 < a.zc;
  int b = (lo ? a.om : a.oc) + ch * TPS + __mul24(j0 - a.y0 + 1, TPW) + (i0 - a.x0 + 1);
  int dz = lo ? a.oc - a.om : a.op - a.oc;
  const Box8 c = box8(a.s, b, dz, TPW);
  return make_float2(vmin3_(vmin3_(c.c000, c.c100, c.c010), vmin3_(c.c110, c.c001, c.c101), vmin_(c.c011, c.c111)),
                     vmax3_(vmax3_(c.c000, c.c100, c.c010), vmax3_(c.c110, c.c001, c.c101), vmax_(c.c011, c.c111)));
}

// Trilinear, clamp-to-edge sample of one channel at texel coordinates (t - 1/2 convention
// applied by the caller); x, then y, then z lerps as a + t (b - a).
template <class A>
__device__ __forceinline__ float samp1(const A& a, float tx, float ty, float tz, int ch) {
  float ax = __builtin_amdgcn_fractf(tx), ay = __builtin_amdgcn_fractf(ty), az = __builtin_amdgcn_fractf(tz);
  int i0 = floori_(tx), j0 = floori_(ty), k0 = floori_(tz);
  if (!(a.in(i0, j0, k0) && a.in(i0 + 1, j0 + 1, k0 + 1))) return samp1_global(a.g, a.W, a.H, a.D, i0, j0, k0, ax, ay, az, ch);
  int b = a.sidx(i0, j0, k0, ch);
  int dz = a.sidx(i0, j0, k0 + 1, ch) - b;
  float c000 = a.s[b], c100 = a.s[b + 1], c010 = a.s[b + TPW], c110 = a.s[b + TPW + 1];
  float c001 = a.s[b + dz], c101 = a.s[b + dz + 1], c011 = a.s[b + dz + TPW], c111 = a.s[b + dz + TPW + 1];
  float c00 = lerpf(c000, c100, ax), c10 = lerpf(c010, c110, ax);
  float c01 = lerpf(c001, c101, ax), c11 = lerpf(c011, c111, ax);
  return lerpf(lerpf(c00, c10, ay), lerpf(c01, c11, ay), az);
}

template <class A>
__device__ __forceinline__ float4 samp4(const A& a, float tx, float ty, float tz) {
  return make_float4(samp1(a, tx, ty, tz, 0), samp1(a, tx, ty, tz, 1), samp1(a, tx, ty, tz, 2),
                     samp1(a, tx, ty, tz, 3));
}

// min / max of one channel over the 8 texels around a point (the MacCormack limiter)
template <class A>
__device__ __forceinline__ float2 minmax1(const A& a, int i0, int j0, int k0, int ch) {
  if (!(a.in(i0, j0, k0) && a.in(i0 + 1, j0 + 1, k0 + 1))) return minmax1_global(a.g, a.W, a.H, a.D, i0, j0, k0, ch);
  float lo = 1e30f, hi = -1e30f;
  int b = a.sidx(i0, j0, k0, ch);
  int dz = a.sidx(i0, j0, k0 + 1, ch) - b;
  const int off[8] = {0, 1, TPW, TPW + 1, dz, dz + 1, dz + TPW, dz + TPW + 1};
#pragma unroll
  for (int q = 0; q < 8; q++) {
    float s = a.s[b + off[q]];
    lo = fminf(lo, s);
    hi = fmaxf(hi, s);
  }
  return make_float2(lo, hi);
}

}  // namespace obl
