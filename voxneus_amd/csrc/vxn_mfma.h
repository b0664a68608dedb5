// BF16 matrix-core helpers for the radiance MLP (gfx950, wave64).
//
// v_mfma_f32_32x32x16_bf16 computes C[32x32] += A[32x16] . B[16x32] in FP32.  Lane l
// (j = l % 32, g = l / 32) holds
//   A: row j, columns 8g .. 8g+7          (8 bf16 = 4 VGPRs)
//   B: rows 8g .. 8g+7, column j          (8 bf16 = 4 VGPRs)
//   C: element v (0..15) = row 8(v/4) + 4g + v%4, column j.
// Sums over the K index do not depend on its order, so the C registers of one product
// are used unchanged as the B operand of the next one: the 16 rows a lane holds in C
// become the K positions 8g+e of two K-steps q = 0, 1 (e = 0..7, v = 8q + e), provided
// the other operand enumerates the same rows in the same order (kperm below).
#pragma once
#include <stdint.h>

#include "vxn_device.h"

namespace vxn {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));

// Two floats rounded to nearest-even bf16, packed low | high << 16.
VXN_DEV uint32_t pk_bf16(float lo, float hi) {
  const bf16x2_t r = {(__bf16)lo, (__bf16)hi};
  return __builtin_bit_cast(uint32_t, r);
}

VXN_DEV uint16_t bf16_bits(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }

VXN_DEV float bf16_round(float x) { return (float)(__bf16)x; }

VXN_DEV f32x16_t mfma_bf16(u32x4_t a, u32x4_t b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c,
                                                 0, 0, 0);
}

// Exchange across the wave halves: returns {lo, hi} where
//   lo = lanes 0-31: a of the same lane,      lanes 32-63: b of lane l-32;
//   hi = lanes 0-31: a of lane l+32,          lanes 32-63: b of the same lane.
VXN_DEV u32x2_t half_swap(uint32_t a, uint32_t b) {
  const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
  return u32x2_t{r[0], r[1]};
}

// Row (0..31 of a 32-row tile) that C element v of wave half g holds.
VXN_DEV int crow(int v, int g) { return 8 * (v >> 2) + 4 * g + (v & 3); }

// Hidden unit at K position 8g + e of K-step (mt, q) when a [64 hidden x 32] C tile pair
// (mt = 0, 1) is reused as a B operand (see the file comment).
VXN_DEV int kperm(int mt, int q, int g, int e) { return 32 * mt + 16 * q + 8 * (e >> 2) + 4 * g + (e & 3); }

}  // namespace vxn
