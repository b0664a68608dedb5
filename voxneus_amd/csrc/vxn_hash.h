// Per-level hash-grid lookups shared by the standalone and fused kernels
// (radiance_hash.py:71-166: hashes every level, 8 corners with no top clamp,
// feature = w0*e0 + w1*e1 + ... + w7*e7 summed left to right, times the lambda weight).
#pragma once

#include "vxn_device.h"

namespace vxn {

struct HashDesc {
  const void* table;  // [L, T, 2] float32 or bfloat16
  int dtype;          // 0 = float32, 1 = bfloat16
  int L;
  int64_t T;
  uint32_t mask;
  float res[32];
  float lw[32];
};

VXN_DEV float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

VXN_DEV void load_entry(const HashDesc& h, int l, uint32_t idx, float& e0, float& e1) {
  const int64_t off = (int64_t)l * h.T + idx;
  if (h.dtype == 0) {
    const float2 v = reinterpret_cast<const float2*>(h.table)[off];
    e0 = v.x;
    e1 = v.y;
  } else {
    const uint32_t v = reinterpret_cast<const uint32_t*>(h.table)[off];
    e0 = __uint_as_float(v << 16);
    e1 = __uint_as_float(v & 0xffff0000u);
  }
}

// PAIR: with a bfloat16 table, read the two entries of corners k and k + 4 with one
// 16-byte load when they share an aligned group of four entries.  The corners differ only
// in x, so their entries differ by x ^ (x + 1) in the low bits: 1 for even x, 3 when x
// ends in 01, which covers three quarters of the pairs.  (With a float32 table wide pair
// loads are slower in the fused render, so they are not used.)
template <bool PAIR>
VXN_DEV void hash_level_fwd(const HashDesc& h, int l, float x, float y, float z, float& f0, float& f1) {
  const float lw = h.lw[l];
  if (lw == 0.0f) {
    f0 = 0.0f;
    f1 = 0.0f;
    return;
  }
  const HashCell c = hash_cell(x, y, z, h.res[l], h.mask);
  float e[8][2];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t i0 = c.idx[k], i1 = c.idx[k + 4];
    if (PAIR && h.dtype == 1 && (i0 ^ i1) <= 3u && h.T >= 4) {
      // T is a power of two, so a group of four never crosses a level
      const int64_t off = ((int64_t)l * h.T + (i0 & ~3u)) >> 2;
      const uint4 v = reinterpret_cast<const uint4*>(h.table)[off];
      const uint32_t a0 = i0 & 3u, a1 = i1 & 3u;
      const uint32_t w0 = a0 == 0 ? v.x : a0 == 1 ? v.y : a0 == 2 ? v.z : v.w;
      const uint32_t w1 = a1 == 0 ? v.x : a1 == 1 ? v.y : a1 == 2 ? v.z : v.w;
      e[k][0] = __uint_as_float(w0 << 16);
      e[k][1] = __uint_as_float(w0 & 0xffff0000u);
      e[k + 4][0] = __uint_as_float(w1 << 16);
      e[k + 4][1] = __uint_as_float(w1 & 0xffff0000u);
    } else {
      load_entry(h, l, i0, e[k][0], e[k][1]);
      load_entry(h, l, i1, e[k + 4][0], e[k + 4][1]);
    }
  }
  float a0 = c.w[0] * e[0][0], a1 = c.w[0] * e[0][1];
#pragma unroll
  for (int k = 1; k < 8; ++k) {
    a0 = a0 + c.w[k] * e[k][0];
    a1 = a1 + c.w[k] * e[k][1];
  }
  if (lw < 1.0f) {
    a0 = a0 * lw;
    a1 = a1 * lw;
  }
  f0 = a0;
  f1 = a1;
}

VXN_DEV void hash_level_bwd(const HashDesc& h, int l, float x, float y, float z, float d0, float d1,
                            float* __restrict__ grad_table) {
  const float lw = h.lw[l];
  if (lw == 0.0f || (d0 == 0.0f && d1 == 0.0f)) return;
  if (lw < 1.0f) {
    d0 = d0 * lw;
    d1 = d1 * lw;
  }
  const HashCell c = hash_cell(x, y, z, h.res[l], h.mask);
  float* base = grad_table + (int64_t)l * h.T * 2;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    float* p = base + (int64_t)c.idx[k] * 2;
    atomicAdd(p, d0 * c.w[k]);
    atomicAdd(p + 1, d1 * c.w[k]);
  }
}

// Validates the hash configuration and fills a HashDesc (defined in hash.hip).
int make_hash_desc(HashDesc& h, const void* table, int dtype, int L, int log2T, int F, const int32_t* level_res,
                   const float* level_w);

}  // namespace vxn
