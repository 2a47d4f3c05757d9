// Exact anelastic pressure solve for gfx950 (replaces solver.js:12-103).
//
// The reference applies dense DST-II matrices (one GEMM per transform pass).
// Here every pass uses the even/odd symmetry of the DST-II basis,
// v_l(N-1-i) = (-1)^l v_l(i): the even modes see only S(i) = F(i) + F(N-1-i),
// the odd modes only D(i) = F(i) - F(N-1-i), i < N/2. Each pass is therefore
// two half-size GEMMs, computed together by one kernel on the MFMA units
// (v_mfma_f32_16x16x4f32), and the flops halve. Spectral arrays keep the even
// modes in the first half of each axis and the odd modes in the second half;
// the Thomas kernel uses eigenvalues permuted the same way.
//
// Passes (rows of the padded grid F[z][y][x], x fastest):
//   XF  G[r][s]      = sum_i S/D[r][i] W[s][i]            r = y + nyp z
//   YF  G[z][s][x]   = sum_j W[s][j] S/D[z][j][x]
//   XI  F[r][i], F[r][N-1-i] = E +- O, E = sum_s G[r][s] WT[i][s]
//   YI  F[z][j][x], F[z][N-1-j][x] = E +- O
// Thomas: one (sx, sy) mode per work-item, z swept in registers; FP32 with the
// reference's operation order, FP64, or FP64 only for the modes whose vertical
// system is near-singular.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "obl_kernels.h"

namespace obl {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));

constexpr int BM = 64, BN = 64, BK = 16;

// bf16 split of an fp32 value: hi = RNE(x), lo = RNE(x - hi)
__device__ __forceinline__ short bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return (short)(r >> 16);
}
__device__ __forceinline__ float bf16_to_f(short h) { return __uint_as_float(((uint32_t)(uint16_t)h) << 16); }

// Workgroup b of a launch runs on XCD b % 8. Renumber the tiles so that each XCD receives a
// contiguous range of them: the tiles that share a data block (consecutive in the tile order)
// then meet in one L2 instead of fetching the block into all eight.
__device__ __forceinline__ int xcd_remap(int b, int G) {
  int per = G >> 3, rem = G & 7, x = b & 7, j = b >> 3;
  return x < rem ? x * (per + 1) + j : rem * (per + 1) + (x - rem) * per + j;
}

// Tile coordinates of this workgroup. X passes: output-column tiles fastest (they share the data
// rows); Y passes: output-row tiles fastest (they share the data columns of one plane).
template <bool XMODE>
__device__ __forceinline__ void dst_tile(int& mt, int& nt, int& plane) {
  int gx = gridDim.x, gy = gridDim.y;
  int G = gx * gy * gridDim.z;
  int b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  int t = xcd_remap(b, G);
  if (XMODE) {
    nt = t % gx;
    t /= gx;
    mt = t % gy;
    plane = t / gy;
  } else {
    mt = t % gy;
    t /= gy;
    nt = t % gx;
    plane = t / gx;
  }
}

// LDS layout of the k-major tiles (rows of BK = 16 floats, no padding): the four 4-float chunks of
// row r are stored in the order q ^ ((r >> 2) & 3). A wave's float4 stores (4 threads per row, 16
// consecutive rows) and its fragment reads (16 rows, one chunk each) then fall in distinct banks.
__device__ __forceinline__ int swz(int r, int q) { return r * BK + 4 * (q ^ ((r >> 2) & 3)); }
constexpr int YLD = BN + 4;  // Y-mode data tile row ([k][n], n contiguous)

template <int MODE, int PREC>
__global__ void __launch_bounds__(256) k_dst(DstArgs a) {
  // one LDS stage; the global loads of stage s + 1 go to registers before the MFMAs of
  // stage s, so their latency overlaps this wave's own MFMAs
  constexpr bool XMODE = (MODE == DST_XF || MODE == DST_XI);
  __shared__ __attribute__((aligned(16))) float As[2][BM * BK];                 // k-major rows (swizzled)
  __shared__ __attribute__((aligned(16))) float Bs[2][XMODE ? BN * BK : BK * YLD];  // X: k-major; Y: [k][n]
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, kq = lane >> 4;
  int mt, nt, plane;
  dst_tile<XMODE>(mt, nt, plane);
  const int m0 = mt * BM, n0 = nt * BN;
  const int N = a.N, Nh = a.Nh;
  const size_t planeSz = (size_t)N * (size_t)a.inner;

  v4f acc[2][2][2];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[p][i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  float4 d0, d1, w0, w1;
  const int row = tid >> 2, q = tid & 3;   // matrix rows (and X-mode data rows)
  const int kk = tid >> 4, qq = tid & 15;  // Y-mode data: k row, 4 columns
  auto fetch = [&](int k0) {
    const size_t wo = (size_t)((XMODE ? n0 : m0) + row) * a.Kp + k0 + 4 * q;
    w0 = *reinterpret_cast<const float4*>(a.W0 + wo);
    w1 = *reinterpret_cast<const float4*>(a.W1 + wo);
    d0 = make_float4(0.f, 0.f, 0.f, 0.f);
    d1 = d0;
    if (XMODE) {
      int r = m0 + row, k = k0 + 4 * q;
      if (r < a.rows && k < Nh) {
        const float* src = a.in + (size_t)r * N;
        if (MODE == DST_XF) {
          float4 f = *reinterpret_cast<const float4*>(src + k);
          float4 g = *reinterpret_cast<const float4*>(src + (N - 4 - k));
          d0 = make_float4(f.x + g.w, f.y + g.z, f.z + g.y, f.w + g.x);
          d1 = make_float4(f.x - g.w, f.y - g.z, f.z - g.y, f.w - g.x);
        } else {
          d0 = *reinterpret_cast<const float4*>(src + k);
          d1 = *reinterpret_cast<const float4*>(src + Nh + k);
        }
      }
    } else {
      int k = k0 + kk, n = n0 + 4 * qq;
      if (k < Nh && n < a.inner) {
        const float* pl = a.in + (size_t)plane * planeSz;
        if (MODE == DST_YF) {
          float4 f = *reinterpret_cast<const float4*>(pl + (size_t)k * a.inner + n);
          float4 g = *reinterpret_cast<const float4*>(pl + (size_t)(N - 1 - k) * a.inner + n);
          d0 = make_float4(f.x + g.x, f.y + g.y, f.z + g.z, f.w + g.w);
          d1 = make_float4(f.x - g.x, f.y - g.y, f.z - g.z, f.w - g.w);
        } else {
          d0 = *reinterpret_cast<const float4*>(pl + (size_t)k * a.inner + n);
          d1 = *reinterpret_cast<const float4*>(pl + (size_t)(Nh + k) * a.inner + n);
        }
      }
    }
  };
  auto stash = [&]() {
    const int o = swz(row, q);
    if (XMODE) {
      *reinterpret_cast<float4*>(&As[0][o]) = d0;
      *reinterpret_cast<float4*>(&As[1][o]) = d1;
      *reinterpret_cast<float4*>(&Bs[0][o]) = w0;
      *reinterpret_cast<float4*>(&Bs[1][o]) = w1;
    } else {
      *reinterpret_cast<float4*>(&As[0][o]) = w0;
      *reinterpret_cast<float4*>(&As[1][o]) = w1;
      *reinterpret_cast<float4*>(&Bs[0][kk * YLD + 4 * qq]) = d0;
      *reinterpret_cast<float4*>(&Bs[1][kk * YLD + 4 * qq]) = d1;
    }
  };

  fetch(0);
  for (int k0 = 0; k0 < Nh; k0 += BK) {
    stash();
    __syncthreads();
    if (k0 + BK < Nh) fetch(k0 + BK);
    // ---------------- MFMA ----------------
    // Lane (lr, kq) feeds A[m = lr][k] and B[k][n = lr] for k = 4 kq + t at step t:
    // the same k permutation on both operands leaves the sum unchanged.
#pragma unroll
    for (int p = 0; p < 2; p++) {
      float4 af[2], bf[2];
#pragma unroll
      for (int i = 0; i < 2; i++) af[i] = *reinterpret_cast<const float4*>(&As[p][swz(wm * 32 + i * 16 + lr, kq)]);
#pragma unroll
      for (int j = 0; j < 2; j++) {
        if (XMODE) {
          bf[j] = *reinterpret_cast<const float4*>(&Bs[p][swz(wn * 32 + j * 16 + lr, kq)]);
        } else {
          const float* bp = &Bs[p][(4 * kq) * YLD + wn * 32 + j * 16 + lr];
          bf[j] = make_float4(bp[0], bp[YLD], bp[2 * YLD], bp[3 * YLD]);
        }
      }
      if (PREC == 0) {
        // k step outermost: four independent accumulators between two updates of one
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++) acc[p][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i].x, bf[j].x, acc[p][i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++) acc[p][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i].y, bf[j].y, acc[p][i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++) acc[p][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i].z, bf[j].z, acc[p][i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++) acc[p][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i].w, bf[j].w, acc[p][i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }

  // ---------------- epilogue ----------------
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        int m = m0 + wm * 32 + i * 16 + 4 * kq + e;
        int n = n0 + wn * 32 + j * 16 + lr;
        float v0 = acc[0][i][j][e], v1 = acc[1][i][j][e];
        if (MODE == DST_XF) {
          if (m < a.rows && n < Nh) {
            float* dst = a.out + (size_t)m * N;
            dst[n] = v0;
            dst[Nh + n] = v1;
          }
        } else if (MODE == DST_XI) {
          if (m < a.rows && n < Nh) {
            float* dst = a.out + (size_t)m * N;
            dst[n] = v0 + v1;
            dst[N - 1 - n] = v0 - v1;
          }
        } else if (MODE == DST_YF) {
          if (m < Nh && n < a.inner) {
            float* pl = a.out + (size_t)plane * planeSz;
            pl[(size_t)m * a.inner + n] = v0;
            pl[(size_t)(Nh + m) * a.inner + n] = v1;
          }
        } else {
          if (m < Nh && n < a.inner) {
            float* pl = a.out + (size_t)plane * planeSz;
            pl[(size_t)m * a.inner + n] = v0 + v1;
            pl[(size_t)(N - 1 - m) * a.inner + n] = v0 - v1;
          }
        }
      }
}

// ---------------- split hi/lo BF16 form ----------------
// Each FP32 operand x is split into hi = RNE_bf16(x) and lo = RNE_bf16(x - hi); the
// product is hi*hi + hi*lo + lo*hi accumulated in FP32 (the lo*lo term, ~2^-18 relative,
// is dropped). v_mfma_f32_16x16x32_bf16, 128 x 128 block tiles, 64 x 64 per wave.
// The sine matrices are split once on the host and staged as BF16 hi/lo tiles. The data tile is
// staged in LDS as FP32 (even / odd folded), and each wave splits the fragments it reads with the
// hardware's packed conversion (v_cvt_pk_bf16_f32), so the split overlaps its MFMAs instead of
// sitting between the global loads and the barrier.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
constexpr int BBM = 128, BBN = 128, BBK = 32;
constexpr int BLD = BBK + 8;   // matrix LDS row: 40 bf16 = 80 B
constexpr int DLD = BBK + 4;   // data LDS row: 36 floats = 144 B

// hi / lo BF16 halves of 8 consecutive floats (RNE, as the host split of the matrices)
__device__ __forceinline__ void split8_hw(float4 v0, float4 v1, bf16x8& hi, bf16x8& lo) {
  const float x[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
  uint32_t h[4], l[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    bf16x2 hb = __builtin_convertvector(f32x2{x[2 * i], x[2 * i + 1]}, bf16x2);
    uint32_t hu = __builtin_bit_cast(uint32_t, hb);
    float h0 = __uint_as_float(hu << 16), h1 = __uint_as_float(hu & 0xffff0000u);
    bf16x2 lb = __builtin_convertvector(f32x2{x[2 * i] - h0, x[2 * i + 1] - h1}, bf16x2);
    h[i] = hu;
    l[i] = __builtin_bit_cast(uint32_t, lb);
  }
  hi = __builtin_bit_cast(bf16x8, make_uint4(h[0], h[1], h[2], h[3]));
  lo = __builtin_bit_cast(bf16x8, make_uint4(l[0], l[1], l[2], l[3]));
}

template <int MODE>
__global__ void __launch_bounds__(256, 2) k_dst_bf16(DstArgs a) {
  // One LDS stage. The global loads of stage s + 1 are issued into registers before the MFMAs of
  // stage s and written to LDS after them, so their latency overlaps this stage's MFMAs.
  constexpr bool XMODE = (MODE == DST_XF || MODE == DST_XI);
  // data: [parity][row (X: data row; Y: output column n)][k] FP32; matrix: [parity][hi/lo][row][k]
  __shared__ __attribute__((aligned(16))) float Ds[2][XMODE ? BBM : BBN][DLD];
  __shared__ __attribute__((aligned(16))) uint16_t Ms[2][2][XMODE ? BBN : BBM][BLD];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, g = lane >> 4;
  int mt, nt, plane;
  dst_tile<XMODE>(mt, nt, plane);
  const int m0 = mt * BBM, n0 = nt * BBN;
  const int N = a.N, Nh = a.Nh;
  const size_t planeSz = (size_t)N * (size_t)a.inner;

  v4f acc[2][4][4];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) acc[p][i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  // staging registers: matrix [array][half of the 32 k] and data (X: 4 float4 of row `row` at
  // k, then 4 at the mirrored / odd position; Y: 4 k rows of 4 columns, then their partners)
  uint4 w0a, w0b, w1a, w1b, w2a, w2b, w3a, w3b;
  float4 pd[8];
  const int row = tid >> 1, half = tid & 1;
  const int ng = tid & 31, kq = tid >> 5;
  const float* pl = a.in + (XMODE ? 0 : (size_t)plane * planeSz);
  auto fetch = [&](int k0) {
    const size_t wo = (size_t)((XMODE ? n0 : m0) + row) * a.KpB + k0 + 16 * half;
    const uint4* q0 = reinterpret_cast<const uint4*>(a.W0h + wo);
    const uint4* q1 = reinterpret_cast<const uint4*>(a.W0l + wo);
    const uint4* q2 = reinterpret_cast<const uint4*>(a.W1h + wo);
    const uint4* q3 = reinterpret_cast<const uint4*>(a.W1l + wo);
    w0a = q0[0]; w0b = q0[1];
    w1a = q1[0]; w1b = q1[1];
    w2a = q2[0]; w2b = q2[1];
    w3a = q3[0]; w3b = q3[1];
#pragma unroll
    for (int c = 0; c < 8; c++) pd[c] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (XMODE) {
      const int r = m0 + row;
      const float* sr = a.in + (size_t)r * N;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int kk = k0 + 16 * half + 4 * c;
        if (r < a.rows && kk < Nh) {
          pd[c] = *reinterpret_cast<const float4*>(sr + kk);
          pd[4 + c] = *reinterpret_cast<const float4*>(sr + (MODE == DST_XF ? N - 4 - kk : Nh + kk));
        }
      }
    } else {
      const int gn = n0 + 4 * ng;
#pragma unroll
      for (int kk = 0; kk < 4; kk++) {
        const int k = k0 + 4 * kq + kk;
        if (k < Nh && gn < a.inner) {
          pd[kk] = *reinterpret_cast<const float4*>(pl + (size_t)k * a.inner + gn);
          pd[4 + kk] = *reinterpret_cast<const float4*>(pl + (size_t)(MODE == DST_YF ? N - 1 - k : Nh + k) * a.inner + gn);
        }
      }
    }
  };
  auto stash = [&]() {
    *reinterpret_cast<uint4*>(&Ms[0][0][row][16 * half]) = w0a;
    *reinterpret_cast<uint4*>(&Ms[0][0][row][16 * half + 8]) = w0b;
    *reinterpret_cast<uint4*>(&Ms[0][1][row][16 * half]) = w1a;
    *reinterpret_cast<uint4*>(&Ms[0][1][row][16 * half + 8]) = w1b;
    *reinterpret_cast<uint4*>(&Ms[1][0][row][16 * half]) = w2a;
    *reinterpret_cast<uint4*>(&Ms[1][0][row][16 * half + 8]) = w2b;
    *reinterpret_cast<uint4*>(&Ms[1][1][row][16 * half]) = w3a;
    *reinterpret_cast<uint4*>(&Ms[1][1][row][16 * half + 8]) = w3b;
    if (XMODE) {
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const float4 f = pd[c], h = pd[4 + c];
        float4 s0 = f, s1 = h;
        if (MODE == DST_XF) {
          s0 = make_float4(f.x + h.w, f.y + h.z, f.z + h.y, f.w + h.x);
          s1 = make_float4(f.x - h.w, f.y - h.z, f.z - h.y, f.w - h.x);
        }
        *reinterpret_cast<float4*>(&Ds[0][row][16 * half + 4 * c]) = s0;
        *reinterpret_cast<float4*>(&Ds[1][row][16 * half + 4 * c]) = s1;
      }
    } else {
      // 4 k x 4 n per thread, transposed to [n][k] (float4 along k)
      float s0[4][4], s1[4][4];
#pragma unroll
      for (int kk = 0; kk < 4; kk++) {
        const float4 f = pd[kk], h = pd[4 + kk];
        float4 f0 = f, f1 = h;
        if (MODE == DST_YF) {
          f0 = make_float4(f.x + h.x, f.y + h.y, f.z + h.z, f.w + h.w);
          f1 = make_float4(f.x - h.x, f.y - h.y, f.z - h.z, f.w - h.w);
        }
        s0[kk][0] = f0.x; s0[kk][1] = f0.y; s0[kk][2] = f0.z; s0[kk][3] = f0.w;
        s1[kk][0] = f1.x; s1[kk][1] = f1.y; s1[kk][2] = f1.z; s1[kk][3] = f1.w;
      }
#pragma unroll
      for (int nn = 0; nn < 4; nn++) {
        *reinterpret_cast<float4*>(&Ds[0][4 * ng + nn][4 * kq]) = make_float4(s0[0][nn], s0[1][nn], s0[2][nn], s0[3][nn]);
        *reinterpret_cast<float4*>(&Ds[1][4 * ng + nn][4 * kq]) = make_float4(s1[0][nn], s1[1][nn], s1[2][nn], s1[3][nn]);
      }
    }
  };

  fetch(0);
  for (int k0 = 0; k0 < Nh; k0 += BBK) {
    stash();
    __syncthreads();
    if (k0 + BBK < Nh) fetch(k0 + BBK);
    // ---- MFMA: lane group g feeds k = 8 g .. 8 g + 7 of both operands ----
    // A = rows m (X: data, Y: matrix), B = columns n (X: matrix, Y: data); the four A fragments
    // are held for the parity, one B fragment at a time.
    const int dbase = XMODE ? wm * 64 : wn * 64, mbase = XMODE ? wn * 64 : wm * 64;
#pragma unroll
    for (int p = 0; p < 2; p++) {
      bf16x8 ah[4], al[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        if (XMODE) {
          const float* dp = &Ds[p][dbase + i * 16 + lr][8 * g];
          split8_hw(*reinterpret_cast<const float4*>(dp), *reinterpret_cast<const float4*>(dp + 4), ah[i], al[i]);
        } else {
          ah[i] = *reinterpret_cast<const bf16x8*>(&Ms[p][0][mbase + i * 16 + lr][8 * g]);
          al[i] = *reinterpret_cast<const bf16x8*>(&Ms[p][1][mbase + i * 16 + lr][8 * g]);
        }
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        bf16x8 bh, bl;
        if (XMODE) {
          bh = *reinterpret_cast<const bf16x8*>(&Ms[p][0][mbase + j * 16 + lr][8 * g]);
          bl = *reinterpret_cast<const bf16x8*>(&Ms[p][1][mbase + j * 16 + lr][8 * g]);
        } else {
          const float* dp = &Ds[p][dbase + j * 16 + lr][8 * g];
          split8_hw(*reinterpret_cast<const float4*>(dp), *reinterpret_cast<const float4*>(dp + 4), bh, bl);
        }
        // small terms first; four independent accumulators between two updates of one
#pragma unroll
        for (int i = 0; i < 4; i++) acc[p][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, acc[p][i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) acc[p][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, acc[p][i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) acc[p][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, acc[p][i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // ---- epilogue (as the FP32 form) ----
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        int m = m0 + wm * 64 + i * 16 + 4 * g + e;
        int n = n0 + wn * 64 + j * 16 + lr;
        float v0 = acc[0][i][j][e], v1 = acc[1][i][j][e];
        if (MODE == DST_XF) {
          if (m < a.rows && n < Nh) {
            float* dst = a.out + (size_t)m * N;
            dst[n] = v0;
            dst[Nh + n] = v1;
          }
        } else if (MODE == DST_XI) {
          if (m < a.rows && n < Nh) {
            float* dst = a.out + (size_t)m * N;
            dst[n] = v0 + v1;
            dst[N - 1 - n] = v0 - v1;
          }
        } else if (MODE == DST_YF) {
          if (m < Nh && n < a.inner) {
            float* pl = a.out + (size_t)plane * planeSz;
            pl[(size_t)m * a.inner + n] = v0;
            pl[(size_t)(Nh + m) * a.inner + n] = v1;
          }
        } else {
          if (m < Nh && n < a.inner) {
            float* pl = a.out + (size_t)plane * planeSz;
            pl[(size_t)m * a.inner + n] = v0 + v1;
            pl[(size_t)(N - 1 - m) * a.inner + n] = v0 - v1;
          }
        }
      }
}

void launch_dst(hipStream_t s, int mode, int prec, const DstArgs& a) {
  dim3 blk(256);
  if (prec == GEMM_BF16X3) {
    if (mode == DST_XF || mode == DST_XI) {
      dim3 grid((a.Nh + BBN - 1) / BBN, (a.rows + BBM - 1) / BBM, 1);
      if (mode == DST_XF) hipLaunchKernelGGL((k_dst_bf16<DST_XF>), grid, blk, 0, s, a);
      else hipLaunchKernelGGL((k_dst_bf16<DST_XI>), grid, blk, 0, s, a);
    } else {
      dim3 grid((a.inner + BBN - 1) / BBN, (a.Nh + BBM - 1) / BBM, a.rows);
      if (mode == DST_YF) hipLaunchKernelGGL((k_dst_bf16<DST_YF>), grid, blk, 0, s, a);
      else hipLaunchKernelGGL((k_dst_bf16<DST_YI>), grid, blk, 0, s, a);
    }
    return;
  }
  if (mode == DST_XF || mode == DST_XI) {
    dim3 grid((a.Nh + BN - 1) / BN, (a.rows + BM - 1) / BM, 1);
    if (mode == DST_XF) hipLaunchKernelGGL((k_dst<DST_XF, 0>), grid, blk, 0, s, a);
    else hipLaunchKernelGGL((k_dst<DST_XI, 0>), grid, blk, 0, s, a);
  } else {
    dim3 grid((a.inner + BN - 1) / BN, (a.Nh + BM - 1) / BM, a.rows);
    if (mode == DST_YF) hipLaunchKernelGGL((k_dst<DST_YF, 0>), grid, blk, 0, s, a);
    else hipLaunchKernelGGL((k_dst<DST_YI, 0>), grid, blk, 0, s, a);
  }
}

// ---------------- Thomas ----------------
// FP64 sweep of one mode (the near-singular ones, or every mode in the FP64 form): c' and d' go
// through global scratch, as there are few such modes.
__device__ __forceinline__ void thomas64(const ThomasArgs& a, int sx, int sy, size_t base, size_t sxy) {
  double shift = a.lam64[sx] + a.lam64[a.nx + sy];
  double cprev = 0.0, dprev = 0.0;
  for (int k = 0; k < a.nz; k++) {
    const double* o = a.zop64 + 4 * k;
    size_t idx = base + sxy * k;
    double den = o[2] + shift * o[3] - o[0] * cprev;
    cprev = o[1] / den;
    dprev = ((double)a.hat[idx] - o[0] * dprev) / den;
    a.cp64[idx] = cprev;
    a.dp64[idx] = dprev;
  }
  double above = a.dp64[base + sxy * (a.nz - 1)];
  a.hat[base + sxy * (a.nz - 1)] = (float)above;
  for (int k = a.nz - 2; k >= 0; k--) {
    size_t idx = base + sxy * k;
    above = a.dp64[idx] - a.cp64[idx] * above;
    a.hat[idx] = (float)above;
  }
}

// FP32 sweeps with c' and d' held in registers (nz <= NZMAX): the column of the mode is read once
// (all loads issued before the recurrence needs them) and written once.
template <int NZMAX>
__global__ void __launch_bounds__(256) k_thomas(ThomasArgs a) {
  int sx = blockIdx.x * blockDim.x + threadIdx.x, sy = blockIdx.y * blockDim.y + threadIdx.y;
  if (sx >= a.nx || sy >= a.ny) return;
  const size_t sxy = (size_t)a.nx * a.ny, base = (size_t)sx + (size_t)a.nx * sy;
  bool use64 = a.mode == 1;
  if (a.mode == 2) use64 = fabs(a.lam64[sx] + a.lam64[a.nx + sy]) < a.shift_lo;
  if (use64) {
    thomas64(a, sx, sy, base, sxy);
    return;
  }
  const int nz = a.nz;
  float d[NZMAX], c[NZMAX];
#pragma unroll
  for (int k = 0; k < NZMAX; k++)
    if (k < nz) d[k] = a.hat[base + sxy * k];
  const float shift = a.lam32[sx] + a.lam32[a.nx + sy];
  float cprev = 0.0f, dprev = 0.0f;
#pragma unroll
  for (int k = 0; k < NZMAX; k++) {
    if (k < nz) {
      float4 o = a.zop32[k];
      float den = o.z + shift * o.w - o.x * cprev;
      cprev = o.y / den;
      dprev = (d[k] - o.x * dprev) / den;
      c[k] = cprev;
      d[k] = dprev;
    }
  }
  float above = 0.0f;
#pragma unroll
  for (int k = NZMAX - 1; k >= 0; k--) {
    if (k < nz) {
      above = (k == nz - 1) ? d[k] : d[k] - c[k] * above;
      a.hat[base + sxy * k] = above;
    }
  }
}

// Any nz: c' goes through global scratch.
__global__ void __launch_bounds__(256) k_thomas_any(ThomasArgs a) {
  int sx = blockIdx.x * blockDim.x + threadIdx.x, sy = blockIdx.y * blockDim.y + threadIdx.y;
  if (sx >= a.nx || sy >= a.ny) return;
  const size_t sxy = (size_t)a.nx * a.ny, base = (size_t)sx + (size_t)a.nx * sy;
  bool use64 = a.mode == 1;
  if (a.mode == 2) use64 = fabs(a.lam64[sx] + a.lam64[a.nx + sy]) < a.shift_lo;
  if (use64) {
    thomas64(a, sx, sy, base, sxy);
    return;
  }
  float shift = a.lam32[sx] + a.lam32[a.nx + sy];
  float cprev = 0.0f, dprev = 0.0f;
  for (int k = 0; k < a.nz; k++) {
    float4 o = a.zop32[k];
    size_t idx = base + sxy * k;
    float den = o.z + shift * o.w - o.x * cprev;
    cprev = o.y / den;
    dprev = (a.hat[idx] - o.x * dprev) / den;
    a.cp32[idx] = cprev;
    a.hat[idx] = dprev;
  }
  float above = a.hat[base + sxy * (a.nz - 1)];
  for (int k = a.nz - 2; k >= 0; k--) {
    size_t idx = base + sxy * k;
    above = a.hat[idx] - a.cp32[idx] * above;
    a.hat[idx] = above;
  }
}

void launch_thomas(hipStream_t s, const ThomasArgs& a) {
  dim3 blk(64, 4, 1), grid((a.nx + 63) / 64, (a.ny + 3) / 4, 1);
  if (a.nz <= 64) hipLaunchKernelGGL((k_thomas<64>), grid, blk, 0, s, a);
  else if (a.nz <= 80) hipLaunchKernelGGL((k_thomas<80>), grid, blk, 0, s, a);
  else if (a.nz <= 104) hipLaunchKernelGGL((k_thomas<104>), grid, blk, 0, s, a);
  else hipLaunchKernelGGL(k_thomas_any, grid, blk, 0, s, a);
}

}  // namespace obl
