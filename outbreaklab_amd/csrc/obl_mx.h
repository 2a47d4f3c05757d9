// The multi-piece BF16 operands of the pressure solve's transform passes: the tiled layout and the split of FP32
// values into BF16 pieces, shared by the passes (k_solve.hip) and by the divergence kernel that writes the first
// pass's operand (k_step.hip).
//
// An [R][K] operand (K the reduction index) is cut into 32-row x 16-k tiles, 1 KB per piece; tile (rb, kb) of piece
// q and parity p is the 1 KB block ((p P + q) RB + rb) (KB + 1) + kb, and inside a tile the 16-byte chunk c holds
// row c / 2, k-half (c & 1) ^ s(row), s(r) = bit 2 of r xor bit 3 of r.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace obl {

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int mx_chunk(int ri, int kc) { return 2 * ri + (kc ^ (((ri >> 2) ^ (ri >> 3)) & 1)); }
// bf16 offset of element (row, k) inside one piece plane of a tiled operand with KB k-blocks; a row-block spans
// KB + 1 tiles (an odd pitch: with power-of-two strides, as at N = 1024, the tiles of a k-block alias in the
// memory channels)
__device__ __forceinline__ size_t mx_off(int row, int k, int KB) {
  return ((size_t)(row >> 5) * (KB + 1) + (k >> 4)) * 512 + mx_chunk(row & 31, (k >> 3) & 1) * 8 + (k & 7);
}

// P pieces of 4 floats, 4 bf16 (8 bytes) per piece; the hardware conversion rounds to nearest even
template <int P>
__device__ __forceinline__ void mx_split4(float r0, float r1, float r2, float r3, uint2* out) {
#pragma unroll
  for (int q = 0; q < P; q++) {
    uint32_t u01 = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{r0, r1}, bf16x2));
    uint32_t u23 = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{r2, r3}, bf16x2));
    out[q] = make_uint2(u01, u23);
    if (q + 1 < P) {
      r0 -= __uint_as_float(u01 << 16);
      r1 -= __uint_as_float(u01 & 0xffff0000u);
      r2 -= __uint_as_float(u23 << 16);
      r3 -= __uint_as_float(u23 & 0xffff0000u);
    }
  }
}
template <int P>
__device__ __forceinline__ void mx_split1(float r, uint16_t* out) {
#pragma unroll
  for (int q = 0; q < P; q++) {
    uint32_t u = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{r, 0.0f}, bf16x2));
    out[q] = (uint16_t)u;
    if (q + 1 < P) r -= __uint_as_float(u << 16);
  }
}

}  // namespace obl
