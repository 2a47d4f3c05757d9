// Exact anelastic pressure solve for gfx950 (replaces solver.js:12-103).
//
// The reference applies dense DST-II matrices (one GEMM per transform pass).
// Here every pass uses the even/odd symmetry of the DST-II basis,
// v_l(N-1-i) = (-1)^l v_l(i): the even modes see only S(i) = F(i) + F(N-1-i),
// the odd modes only D(i) = F(i) - F(N-1-i), i < N/2. Each pass is therefore
// two half-size GEMMs, computed together by one kernel on the MFMA units, and
// the flops halve. Spectral arrays keep the even modes in the first half of each
// axis and the odd modes in the second half; the Thomas kernel uses eigenvalues
// permuted the same way.
//
// Two GEMM forms:
//  - FP32 (k_dst, v_mfma_f32_16x16x4f32) on FP32 arrays. Passes (rows of the
//    padded grid F[z][y][x], x fastest):
//      XF  G[r][s]      = sum_i S/D[r][i] W[s][i]            r = y + nyp z
//      YF  G[z][s][x]   = sum_j W[s][j] S/D[z][j][x]
//      XI  F[r][i], F[r][N-1-i] = E +- O, E = sum_s G[r][s] WT[i][s]
//      YI  F[z][j][x], F[z][N-1-j][x] = E +- O
//  - multi-piece BF16 (k_dst_mx, v_mfma_f32_32x32x16_bf16) on operands stored
//    already split, each pass's epilogue writing the next pass's operand.
// Thomas: one (sx, sy) mode per work-item, z swept in registers; FP32 with the
// reference's operation order, FP64, or FP64 only for the modes whose vertical
// system is near-singular.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "obl_kernels.h"
#include "obl_mx.h"

namespace obl {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16;

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

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// ---------------- multi-piece BF16 form (k_dst_mx) ----------------
// Every FP32 operand x is carried as P BF16 pieces, x ~ x1 + x2 (+ x3), each piece the round-to-nearest-even BF16
// of what the pieces before it leave. P = 2 keeps the products x1 y1 + x1 y2 + x2 y1 (split hi/lo BF16, about 16
// bits); P = 3 adds x1 y3 + x3 y1 + x2 y2, the terms of order 2^-16, and carries the transform at FP32 accuracy.
// v_mfma_f32_32x32x16_bf16 accumulates in FP32, the smallest products first.
//
// The operands are stored already split, in a tiled layout: an [R][K] operand (K the reduction index) is cut
// into 32-row x 16-k tiles, 1 KB per piece; tile (rb, kb) of piece q and parity p is the 1 KB block number
// ((p P + q) RB + rb) (KB + 1) + kb, and inside a tile the 16-byte chunk c holds row c / 2, k-half (c & 1) ^ s(row),
// s(r) = bit 2 of r xor bit 3 of r (a swizzle that keeps LDS reads of the layout free of bank conflicts). Each
// operand's producer writes it in this layout: the fold-and-split pass for the x forward pass, the epilogue of
// the pass before, or the Thomas sweep.
//
// Block: 4 waves, 32 MF rows (MF = MX_MF) x 64 NF columns of each parity; waves 0-1 compute the even half, waves
// 2-3 the odd half, each 32 MF x 32 NF (MF x NF fragments of 32 x 32), K advancing 16 per step. With MF = 4 a
// wave holds 192 accumulators and a CU runs one block (measured: 64-row blocks two per CU are 1.5 times slower,
// the sine-matrix traffic doubling).
typedef float v16f __attribute__((ext_vector_type(16)));

// Fold-and-split of the right-hand side into the x forward pass's operand. Its rows come in tiles of 32 MF rows
// holding 16 MF (z, y) pairs (MF = MX_MF): row 16 MF h + q of tile t is y' (h = 0) or ny - 1 - y' (h = 1) of pair
// 16 MF t + q = z (ny / 2) + y', so the pass's epilogue finds both rows of a pair in one lane and can fold y too. k runs over x' < nx / 2:
// parity 0 holds F(x') + F(nx - 1 - x'), parity 1 F(x') - F(nx - 1 - x'). One wave writes one tile of every piece
// and parity (lane l: chunk l).
template <int P>
__global__ void __launch_bounds__(256) k_mx_fold_x(const float* __restrict__ rhs, uint16_t* __restrict__ out, int Nx,
                                                   int Ny, int nz, int RB, int KB) {
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (wave >= RB * KB) return;
  const int rb = wave / KB, kb = wave % KB;
  const int ri = lane >> 1, kc = (lane & 1) ^ (((ri >> 2) ^ (ri >> 3)) & 1);
  const int r = rb * 32 + ri, t = r / (32 * MX_MF), rho = r % (32 * MX_MF);
  const int Hx = Nx >> 1, Hy = Ny >> 1;
  const int pair = 16 * MX_MF * t + rho % (16 * MX_MF);
  const int k0 = kb * 16 + kc * 8;
  float s[8], d[8];
#pragma unroll
  for (int e = 0; e < 8; e++) s[e] = d[e] = 0.0f;
  if (pair < nz * Hy) {
    const int z = pair / Hy, yp = pair % Hy, y = rho >= 16 * MX_MF ? Ny - 1 - yp : yp;
    const float* row = rhs + ((size_t)z * Ny + y) * Nx;
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int k = k0 + 4 * c;
      if (k < Hx) {  // Hx is a multiple of 4: a float4 is all in or all out
        float4 f = *reinterpret_cast<const float4*>(row + k);
        float4 g = *reinterpret_cast<const float4*>(row + (Nx - 4 - k));
        s[4 * c + 0] = f.x + g.w; d[4 * c + 0] = f.x - g.w;
        s[4 * c + 1] = f.y + g.z; d[4 * c + 1] = f.y - g.z;
        s[4 * c + 2] = f.z + g.y; d[4 * c + 2] = f.z - g.y;
        s[4 * c + 3] = f.w + g.x; d[4 * c + 3] = f.w - g.x;
      }
    }
  }
  uint2 ps[2][P], pd[2][P];
  mx_split4<P>(s[0], s[1], s[2], s[3], ps[0]);
  mx_split4<P>(s[4], s[5], s[6], s[7], ps[1]);
  mx_split4<P>(d[0], d[1], d[2], d[3], pd[0]);
  mx_split4<P>(d[4], d[5], d[6], d[7], pd[1]);
  const size_t plane = (size_t)RB * (KB + 1) * 512, tile = ((size_t)rb * (KB + 1) + kb) * 512 + lane * 8;
#pragma unroll
  for (int q = 0; q < P; q++) {
    *reinterpret_cast<uint4*>(out + (size_t)q * plane + tile) = make_uint4(ps[0][q].x, ps[0][q].y, ps[1][q].x, ps[1][q].y);
    *reinterpret_cast<uint4*>(out + (size_t)(P + q) * plane + tile) =
        make_uint4(pd[0][q].x, pd[0][q].y, pd[1][q].x, pd[1][q].y);
  }
}

template <int V>
struct IntC {
  static constexpr int value = V;
};

template <int MODE, int P, int NF>
__global__ void __launch_bounds__(256, MX_MF == 2 ? 2 : 1) k_dst_mx(MxArgs a) {
  constexpr int MF = MX_MF;
  constexpr int NPROD = P == 2 ? 3 : 6;
  // products (piece of A, piece of B), smallest first
  constexpr int PA[6] = {P == 2 ? 1 : 2, P == 2 ? 0 : 1, 0, 1, 0, 0};
  constexpr int PB[6] = {0, P == 2 ? 1 : 1, P == 2 ? 0 : 2, 0, 1, 0};
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wp = wave >> 1, wh = wave & 1;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int nt = t % a.tilesN, mt = t / a.tilesN;

  // Each lane's MFMA fragment of a tile is one 16-byte chunk (row l % 32, k-half l / 32), and the 64 lanes of a
  // wave read the 64 chunks of the tile: every fragment load is one coalesced 1 KB wave load. So each wave loads
  // its own fragments from global memory into registers (the two waves of a parity read the same A fragments,
  // the second from L1), double-buffered one k-block ahead; no LDS and no barrier in the k loop.
  const int loff = mx_chunk(lane & 31, lane >> 5) * 8;
  const int KBP = a.KB + 1;  // row-block pitch in tiles
  const size_t planeA = (size_t)a.aRB * KBP * 512, planeB = (size_t)a.bRB * KBP * 512;
  const uint16_t* pA = a.A + (size_t)(wp * P) * planeA + (size_t)(mt * MF) * KBP * 512 + loff;
  const uint16_t* pB = a.B + (size_t)(wp * P) * planeB + (size_t)(nt * 2 * NF + wh * NF) * KBP * 512 + loff;
  bf16x8 fa[2][MF][P], fb[2][NF][P];
  auto load = [&](auto S, int kb) __attribute__((always_inline)) {
    constexpr int s = decltype(S)::value;
#pragma unroll
    for (int q = 0; q < P; q++)
#pragma unroll
      for (int i = 0; i < MF; i++)
        fa[s][i][q] = *reinterpret_cast<const bf16x8*>(pA + q * planeA + (size_t)(i * KBP + kb) * 512);
#pragma unroll
    for (int q = 0; q < P; q++)
#pragma unroll
      for (int j = 0; j < NF; j++)
        fb[s][j][q] = *reinterpret_cast<const bf16x8*>(pB + q * planeB + (size_t)(j * KBP + kb) * 512);
  };

  v16f acc[MF][NF];
#pragma unroll
  for (int i = 0; i < MF; i++)
#pragma unroll
    for (int j = 0; j < NF; j++)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;

  // product-major order: MF NF independent accumulators between two updates of one
  auto compute = [&](auto S) __attribute__((always_inline)) {
    constexpr int s = decltype(S)::value;
#pragma unroll
    for (int c = 0; c < NPROD; c++)
#pragma unroll
      for (int j = 0; j < NF; j++)
#pragma unroll
        for (int i = 0; i < MF; i++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[s][i][PA[c]], fb[s][j][PB[c]], acc[i][j], 0, 0, 0);
  };

  // KB is even (the operands carry a zero k-block when N / 2 needs an odd number). The loads run unconditionally
  // (the last stage reloads the last k-block, unused): with a branch around them the wait before each compute
  // would have to cover the path without them. The scheduling barriers keep the loads ahead of the MFMAs.
  constexpr IntC<0> S0;
  constexpr IntC<1> S1;
  load(S0, 0);
  for (int kb = 0; kb < a.KB; kb += 2) {
    __builtin_amdgcn_sched_barrier(0);
    load(S1, kb + 1);
    __builtin_amdgcn_sched_barrier(0);
    compute(S0);
    __builtin_amdgcn_sched_barrier(0);
    load(S0, min(kb + 2, a.KB - 1));
    __builtin_amdgcn_sched_barrier(0);
    compute(S1);
  }
  __shared__ __attribute__((aligned(16))) float X[4 * 16 * NF * 64];  // the inverse passes' E / O exchange

  // ---- epilogue: lane l, fragment (i, j), element e sits at row 32 i + 8 (e / 4) + 4 (l / 32) + e % 4 of the
  // block tile and column 32 j + l % 32 of the wave's columns ----
  const int Hx = a.Nx >> 1, Hy = a.Ny >> 1;
  const int ncol0 = nt * 64 * NF + wh * 32 * NF + (lane & 31);
  const int rsub = 4 * (lane >> 5);
  if (MODE == DST_XF) {
    // fold y: rows 32 i + r and 16 MF + 32 i + r of the tile are y' and ny - 1 - y' of one pair; write the next
    // pass's operand: row z nx + sx, k = y', parity 0 the sum, 1 the difference
    const size_t plane = (size_t)a.oRB * (a.oKB + 1) * 512;
#pragma unroll
    for (int i = 0; i < MF / 2; i++)
#pragma unroll
      for (int j = 0; j < NF; j++) {
        const int n = ncol0 + 32 * j;
        if (n >= Hx) continue;
#pragma unroll
        for (int g = 0; g < 4; g++) {
          const int pair = mt * 16 * MF + 32 * i + 8 * g + rsub;
          if (pair >= a.Mvalid) continue;
          const int z = pair / Hy, yp = pair % Hy;
          const int R = z * a.Nx + wp * Hx + n;
          float s[4], d[4];
#pragma unroll
          for (int e = 0; e < 4; e++) {
            const float lo = acc[i][j][4 * g + e], hi = acc[i + MF / 2][j][4 * g + e];
            s[e] = lo + hi;
            d[e] = lo - hi;
          }
          uint2 ps[P], pd[P];
          mx_split4<P>(s[0], s[1], s[2], s[3], ps);
          mx_split4<P>(d[0], d[1], d[2], d[3], pd);
          const size_t o = mx_off(R, yp, a.oKB);
#pragma unroll
          for (int q = 0; q < P; q++) {
            *reinterpret_cast<uint2*>(a.outT + (size_t)q * plane + o) = ps[q];
            *reinterpret_cast<uint2*>(a.outT + (size_t)(P + q) * plane + o) = pd[q];
          }
        }
      }
  } else if (MODE == DST_YF) {
    // spectral array [z][sx][sy], sy = parity * ny / 2 + n
#pragma unroll
    for (int i = 0; i < MF; i++)
#pragma unroll
      for (int j = 0; j < NF; j++) {
        const int n = ncol0 + 32 * j;
        if (n >= Hy) continue;
#pragma unroll
        for (int e = 0; e < 16; e++) {
          const int r = mt * 32 * MF + 32 * i + 8 * (e >> 2) + rsub + (e & 3);
          if (r < a.Mvalid) a.outF[(size_t)r * a.Ny + wp * Hy + n] = acc[i][j][e];
        }
      }
  } else {
    // inverse passes: the two waves of a column half hold E (even modes) and O (odd modes) of the same outputs;
    // they swap them through LDS, one row fragment at a time, and the even wave writes F(j) = E + O, the odd
    // wave F(N - 1 - j) = E - O
    float* mine = X + wave * 16 * NF * 64;  // [16 NF][64 lanes] per wave
    const float* other = X + (wave ^ 2) * 16 * NF * 64;
    const float sg = wp ? -1.0f : 1.0f;  // F = E + O = mine + other (even wave); E - O = other - mine (odd)
    auto exch = [&](const int i) {
#pragma unroll
      for (int j = 0; j < NF; j++)
#pragma unroll
        for (int e = 0; e < 16; e++) mine[(j * 16 + e) * 64 + lane] = acc[i][j][e];
      __syncthreads();
#pragma unroll
      for (int j = 0; j < NF; j++) {
        const int n = ncol0 + 32 * j;
        if (MODE == DST_YI) {
          if (n < Hy) {
            const size_t plane = (size_t)a.oRB * (a.oKB + 1) * 512;
            const int y = wp ? a.Ny - 1 - n : n;
#pragma unroll
            for (int g = 0; g < 4; g++) {
              const int r = mt * 32 * MF + 32 * i + 8 * g + rsub;  // (z, sx) row of this pass
              if (r >= a.Mvalid) continue;
              const int z = r / a.Nx, sx = r % a.Nx, px = sx >= Hx, k = sx - px * Hx;
              float f[4];
#pragma unroll
              for (int e = 0; e < 4; e++) f[e] = other[(j * 16 + 4 * g + e) * 64 + lane] + sg * acc[i][j][4 * g + e];
              uint2 pc[P];
              mx_split4<P>(f[0], f[1], f[2], f[3], pc);
              const size_t o = mx_off(z * a.Ny + y, k, a.oKB);
#pragma unroll
              for (int q = 0; q < P; q++) *reinterpret_cast<uint2*>(a.outT + (size_t)(px * P + q) * plane + o) = pc[q];
            }
          }
        } else {  // DST_XI: phi rows (z, y), x = n (even wave) or nx - 1 - n (odd wave)
          if (n < Hx) {
            const int x = wp ? a.Nx - 1 - n : n;
#pragma unroll
            for (int e = 0; e < 16; e++) {
              const int r = mt * 32 * MF + 32 * i + 8 * (e >> 2) + rsub + (e & 3);
              if (r < a.Mvalid) a.outF[(size_t)r * a.Nx + x] = other[(j * 16 + e) * 64 + lane] + sg * acc[i][j][e];
            }
          }
        }
      }
      __syncthreads();
    };
    exch(0);
    exch(1);
    if (MF > 2) {
      exch(2);
      exch(3);
    }
  }
}

template <int P, int NF>
static void launch_mx_nf(hipStream_t s, int mode, const MxArgs& a, int blocks) {
  dim3 blk(256), grid(blocks);
  switch (mode) {
    case DST_XF: hipLaunchKernelGGL((k_dst_mx<DST_XF, P, NF>), grid, blk, 0, s, a); break;
    case DST_YF: hipLaunchKernelGGL((k_dst_mx<DST_YF, P, NF>), grid, blk, 0, s, a); break;
    case DST_YI: hipLaunchKernelGGL((k_dst_mx<DST_YI, P, NF>), grid, blk, 0, s, a); break;
    default: hipLaunchKernelGGL((k_dst_mx<DST_XI, P, NF>), grid, blk, 0, s, a); break;
  }
}

void launch_dst_mx(hipStream_t s, int mode, int P, int NF, const MxArgs& a) {
  const int blocks = a.tilesN * (a.aRB / MX_MF);
  if (P == 2) {
    if (NF == 2) launch_mx_nf<2, 2>(s, mode, a, blocks);
    else if (NF == 3) launch_mx_nf<2, 3>(s, mode, a, blocks);
    else launch_mx_nf<2, 4>(s, mode, a, blocks);
  } else {
    if (NF == 2) launch_mx_nf<3, 2>(s, mode, a, blocks);
    else if (NF == 3) launch_mx_nf<3, 3>(s, mode, a, blocks);
    else launch_mx_nf<3, 4>(s, mode, a, blocks);
  }
}

void launch_mx_fold_x(hipStream_t s, int P, const float* rhs, uint16_t* out, int Nx, int Ny, int nz, int RB, int KB) {
  dim3 blk(256), grid((RB * KB + 3) / 4);
  if (P == 2) hipLaunchKernelGGL(k_mx_fold_x<2>, grid, blk, 0, s, rhs, out, Nx, Ny, nz, RB, KB);
  else hipLaunchKernelGGL(k_mx_fold_x<3>, grid, blk, 0, s, rhs, out, Nx, Ny, nz, RB, KB);
}

void launch_dst(hipStream_t s, int mode, const DstArgs& a) {
  dim3 blk(256);
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
// Mode (sx, sy) of the spectral array: its column starts at `base` (layout 0: sx + nx sy, x fastest; layout 1:
// sy + ny sx, y fastest) and steps nx ny per level. The solution goes back into the array, or into the y
// inverse pass's tiled split operand (row z nx + sx, parity sy >= ny / 2, k = sy mod ny / 2).
__device__ __forceinline__ void thomas_put(const ThomasArgs& a, int sx, int sy, size_t idx, int k, float v) {
  if (!a.outT) {
    a.hat[idx] = v;
    return;
  }
  const int Hy = a.ny >> 1, par = sy >= Hy;
  const size_t plane = (size_t)((a.nz * a.nx + 32 * MX_MF - 1) / (32 * MX_MF) * MX_MF) * (a.oKB + 1) * 512;
  uint16_t* o = a.outT + mx_off(k * a.nx + sx, sy - par * Hy, a.oKB);
  uint16_t pc[3];
  if (a.P == 3) {
    mx_split1<3>(v, pc);
    for (int q = 0; q < 3; q++) o[(size_t)(3 * par + q) * plane] = pc[q];
  } else {
    mx_split1<2>(v, pc);
    for (int q = 0; q < 2; q++) o[(size_t)(2 * par + q) * plane] = pc[q];
  }
}

__device__ __forceinline__ void thomas_mode(const ThomasArgs& a, int& sx, int& sy, size_t& base) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y * blockDim.y + threadIdx.y;
  if (a.layout) {
    sy = f;
    sx = g;
    base = (size_t)sy + (size_t)a.ny * sx;
  } else {
    sx = f;
    sy = g;
    base = (size_t)sx + (size_t)a.nx * sy;
  }
}

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
  thomas_put(a, sx, sy, base + sxy * (a.nz - 1), a.nz - 1, (float)above);
  for (int k = a.nz - 2; k >= 0; k--) {
    size_t idx = base + sxy * k;
    above = a.dp64[idx] - a.cp64[idx] * above;
    thomas_put(a, sx, sy, idx, k, (float)above);
  }
}

// The modes of a staged chunk whose bits are set in om, value by value (a chunk that also holds modes this
// kernel does not write)
__device__ __noinline__ static void thomas_put_sparse(uint16_t* o, size_t plane, int P, uint32_t om, float4 v0,
                                                      float4 v1) {
  const float v[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
  for (int e = 0; e < 8; e++) {
    if ((om >> e) & 1u) {
      uint16_t pc[3];
      mx_split1<3>(v[e], pc);
#pragma unroll
      for (int q = 0; q < 3; q++)
        if (q < P) o[(size_t)q * plane + e] = pc[q];
    }
  }
}

// Thomas sweeps with the column in registers (nz <= NZMAX), one launch for both kinds of mode: the first nb64
// blocks take the listed FP64 modes (one per work-item), the others every remaining mode in FP32 (64 x 4 modes
// per block), so the FP64 modes' short serial chains run beside the FP32 blocks instead of after them. FAST (the
// mixed form): the FP32 modes, all of condition below the mixed
// bound, divide through one reciprocal per level (v_rcp_f32, within WGSL's 2.5 ULP for x / y); otherwise
// correctly rounded division in the reference's order (solver.js:86-101).
template <int NZMAX, bool FAST>
__global__ void __launch_bounds__(256) k_thomas(ThomasArgs a, int nb64, int gx) {
  // every level's coefficients in LDS, read once per block: a read per level from global memory would put one
  // memory latency on each step of the recurrence
  __shared__ float4 zs[NZMAX];
  __shared__ double zd[4 * NZMAX];
  const int t = threadIdx.x;
  const int nz = a.nz;
  if (t < nz) zs[t] = a.zop32[t];
  if (nb64 > 0)
    for (int q = t; q < 4 * nz; q += 256) zd[q] = a.zop64[q];
  __syncthreads();
  const size_t sxy = (size_t)a.nx * a.ny;
  if ((int)blockIdx.x < nb64) {
    // FP64 sweep of listed mode i: c' and d' to global scratch ([level][list index], coalesced), read back 8
    // levels at a time, each chunk's loads issued while the chunk above is used (registers stay within the FP32
    // modes' budget). One reciprocal per level, refined by a Newton step (relative error ~2^-52), serves both
    // quotients.
    const int i = blockIdx.x * 256 + t;
    if (i >= a.n64) return;
    const int2 m = a.modes64[i];
    const int sx = m.x, sy = m.y;
    const size_t base = a.layout ? (size_t)sy + (size_t)a.ny * sx : (size_t)sx + (size_t)a.nx * sy;
    float h[NZMAX];
    // unconditional loads (levels past nz repeat the top one): a load under a per-level branch is waited for at
    // the branch's end, one memory latency per level
#pragma unroll
    for (int k = 0; k < NZMAX; k++) h[k] = a.hat[base + sxy * min(k, nz - 1)];
    const double shift = a.lam64[sx] + a.lam64[a.nx + sy];
    double cprev = 0.0, dprev = 0.0;
#pragma unroll
    for (int k = 0; k < NZMAX; k++) {
      if (k < nz) {
        const double* o = zd + 4 * k;
        const double den = o[2] + shift * o[3] - o[0] * cprev;
        double r = __builtin_amdgcn_rcp(den);
        r = fma(r, fma(-den, r, 1.0), r);
        cprev = o[1] * r;
        dprev = ((double)h[k] - o[0] * dprev) * r;
        a.cp64[(size_t)k * a.n64 + i] = cprev;
        a.dp64[(size_t)k * a.n64 + i] = dprev;
      }
    }
    constexpr int CH = 8, NCH = (NZMAX + CH - 1) / CH;
    auto fetch = [&](int c, double* cc, double* dc) {
#pragma unroll
      for (int j = 0; j < CH; j++) {
        const size_t q = (size_t)min(c * CH + j, nz - 1) * a.n64 + i;
        cc[j] = a.cp64[q];
        dc[j] = a.dp64[q];
      }
    };
    double cbuf[2][CH], dbuf[2][CH];
    fetch(NCH - 1, cbuf[(NCH - 1) & 1], dbuf[(NCH - 1) & 1]);
    double above = 0.0;
#pragma unroll
    for (int c = NCH - 1; c >= 0; c--) {
      if (c > 0) fetch(c - 1, cbuf[(c - 1) & 1], dbuf[(c - 1) & 1]);
      const double *cc = cbuf[c & 1], *dc = dbuf[c & 1];
#pragma unroll
      for (int j = CH - 1; j >= 0; j--) {
        const int k = c * CH + j;
        if (k < nz) {
          above = (k == nz - 1) ? dc[j] : dc[j] - cc[j] * above;
          thomas_put(a, sx, sy, base + sxy * k, k, (float)above);
        }
      }
    }
    return;
  }
  // FP32 modes
  const int b = (int)blockIdx.x - nb64;
  const int f = (b % gx) * 64 + (t & 63), g = (b / gx) * 4 + (t >> 6);
  int sx, sy;
  size_t base;
  if (a.layout) {
    sy = f; sx = g;
    base = (size_t)sy + (size_t)a.ny * sx;
  } else {
    sx = f; sy = g;
    base = (size_t)sx + (size_t)a.nx * sy;
  }
  // Into the tiled operand (layout 1, ny a multiple of 16) the block's solutions are staged in LDS eight levels at
  // a time ([level][sx][sy]: 4 x 64 modes) and stored as whole 16-byte chunks of each piece (8 consecutive sy of
  // one row): at each flush a thread splits one chunk's 8 values and stores its P chunks, instead of P 2-byte
  // stores per mode and level, and a wave's stores fill whole 128-byte lines (4 rows x 2 k-halves). A chunk that
  // holds a listed FP64 mode, or a mode past the array, is stored value by value, leaving those modes' bytes alone.
  const bool packed = a.outT != nullptr && a.layout == 1 && (a.ny & 15) == 0;
  if (!packed && (sx >= a.nx || sy >= a.ny)) return;
  // the listed FP64 modes (mixed form) are the first blocks'
  bool own = sx < a.nx && sy < a.ny;
  if (own && a.mode == THOMAS_MIXED && fabs(a.lam64[sx] + a.lam64[a.nx + sy]) < a.shift_lo) {
    if (!packed) return;
    own = false;
  }
  __shared__ __align__(16) float sol[2][8][4][64];
  __shared__ uint64_t ownw[4];
  if (packed) {
    // every thread reaches the staging barriers: a lane past the array sweeps a clamped mode and stores nothing
    const uint64_t m = __ballot(own);
    if ((t & 63) == 0) ownw[t >> 6] = m;
    sx = min(sx, a.nx - 1);
    sy = min(sy, a.ny - 1);
    base = (size_t)sy + (size_t)a.ny * sx;
  }
  float d[NZMAX], c[NZMAX];
#pragma unroll
  for (int k = 0; k < NZMAX; k++) d[k] = a.hat[base + sxy * min(k, nz - 1)];
  const float shift = a.lam32[sx] + a.lam32[a.nx + sy];
  float cprev = 0.0f, dprev = 0.0f;
#pragma unroll
  for (int k = 0; k < NZMAX; k++) {
    if (k < nz) {
      float4 o = zs[k];
      float den = o.z + shift * o.w - o.x * cprev;
      if (FAST) {
        const float r = __builtin_amdgcn_rcpf(den);
        cprev = o.y * r;
        dprev = (d[k] - o.x * dprev) * r;
      } else {
        cprev = o.y / den;
        dprev = (d[k] - o.x * dprev) / den;
      }
      c[k] = cprev;
      d[k] = dprev;
    }
  }
  float above = 0.0f;
  if (!packed) {
#pragma unroll
    for (int k = NZMAX - 1; k >= 0; k--) {
      if (k < nz) {
        above = (k == nz - 1) ? d[k] : d[k] - c[k] * above;
        thomas_put(a, sx, sy, base + sxy * k, k, above);
      }
    }
    return;
  }
  // this thread's chunk at a flush: level k0 + fl, row sx0 + fw, the 8 modes fsy .. fsy + 7 (k-half kc of the
  // wave's 16-mode group kb); lanes 8 apart in a wave write adjacent 128-byte lines
  const int w = t >> 6, lane = t & 63;
  const int fw = t & 3, kc = (t >> 2) & 1, kb = (t >> 3) & 3, fl = t >> 5, fg = 2 * kb + kc;
  const int fsx = (b / gx) * 4 + fw, fsy = (b % gx) * 64 + 8 * fg;
  const int Hy = a.ny >> 1, par = fsy >= Hy, P = a.P;
  const bool fin = fsx < a.nx && fsy < a.ny;
  const size_t plane = (size_t)((a.nz * a.nx + 32 * MX_MF - 1) / (32 * MX_MF) * MX_MF) * (a.oKB + 1) * 512;
  uint16_t* const ob = a.outT + (size_t)(P * par) * plane;
  auto flush = [&](int k0, int buf) {
    __syncthreads();
    const int k = k0 + fl;
    if (k < nz && fin) {
      const float4 v0 = *reinterpret_cast<const float4*>(&sol[buf][fl][fw][8 * fg]);
      const float4 v1 = *reinterpret_cast<const float4*>(&sol[buf][fl][fw][8 * fg + 4]);
      const uint32_t om = (uint32_t)(ownw[fw] >> (8 * fg)) & 0xffu;
      uint16_t* const o = ob + mx_off(k * a.nx + fsx, fsy - par * Hy, a.oKB);
      // (the first two of three pieces are the two-piece split)
      if (om == 0xffu) {
        uint2 lo[3], hi[3];
        mx_split4<3>(v0.x, v0.y, v0.z, v0.w, lo);
        mx_split4<3>(v1.x, v1.y, v1.z, v1.w, hi);
#pragma unroll
        for (int q = 0; q < 3; q++)
          if (q < P) *reinterpret_cast<uint4*>(o + (size_t)q * plane) = make_uint4(lo[q].x, lo[q].y, hi[q].x, hi[q].y);
      } else if (om) {
        thomas_put_sparse(o, plane, P, om, v0, v1);
      }
    }
  };
  constexpr int NB = (NZMAX + 7) / 8;
#pragma unroll
  for (int kb8 = NB - 1; kb8 >= 0; kb8--) {
    if (kb8 * 8 < nz) {
#pragma unroll
      for (int j = 7; j >= 0; j--) {
        const int k = kb8 * 8 + j;
        if (k < NZMAX && k < nz) {
          above = (k == nz - 1) ? d[k] : d[k] - c[k] * above;
          sol[kb8 & 1][j][w][lane] = above;
        }
      }
      flush(kb8 * 8, kb8 & 1);
    }
  }
}

// Any nz: c' goes through global scratch.
__global__ void __launch_bounds__(256) k_thomas_any(ThomasArgs a) {
  int sx, sy;
  size_t base;
  thomas_mode(a, sx, sy, base);
  if (sx >= a.nx || sy >= a.ny) return;
  const size_t sxy = (size_t)a.nx * a.ny;
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
  thomas_put(a, sx, sy, base + sxy * (a.nz - 1), a.nz - 1, above);
  for (int k = a.nz - 2; k >= 0; k--) {
    size_t idx = base + sxy * k;
    above = a.hat[idx] - a.cp32[idx] * above;
    thomas_put(a, sx, sy, idx, k, above);
  }
}

template <int NZ>
static void launch_thomas_nz(hipStream_t s, const ThomasArgs& a, dim3 grid, int nb64, int gx, bool fast) {
  if (fast) hipLaunchKernelGGL((k_thomas<NZ, true>), grid, dim3(256), 0, s, a, nb64, gx);
  else hipLaunchKernelGGL((k_thomas<NZ, false>), grid, dim3(256), 0, s, a, nb64, gx);
}

void launch_thomas(hipStream_t s, const ThomasArgs& a) {
  const int nf = a.layout ? a.ny : a.nx, ns = a.layout ? a.nx : a.ny;
  if (a.nz > 104) {
    dim3 blk(64, 4, 1), grid((nf + 63) / 64, (ns + 3) / 4, 1);
    hipLaunchKernelGGL(k_thomas_any, grid, blk, 0, s, a);
    return;
  }
  // FP64 modes (modes 1 and 2) from the list; mode 1 has no FP32 blocks
  const int nb64 = a.mode != THOMAS_FP32 ? (a.n64 + 255) / 256 : 0;
  const int gx = (nf + 63) / 64, nb32 = a.mode != THOMAS_FP64 ? gx * ((ns + 3) / 4) : 0;
  const dim3 grid(nb64 + nb32), blk(256);
  if (grid.x == 0) return;
  const bool fast = a.mode == THOMAS_MIXED;
  if (a.nz <= 64) launch_thomas_nz<64>(s, a, grid, nb64, gx, fast);
  else if (a.nz <= 80) launch_thomas_nz<80>(s, a, grid, nb64, gx, fast);
  else launch_thomas_nz<104>(s, a, grid, nb64, gx, fast);
}

}  // namespace obl
