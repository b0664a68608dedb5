// Active vertex set V_R (bitmask -> ascending compaction), fused closed-form eikonal +
// curvature gradients with FP64 loss reduction, reference-shaped single regularisers,
// and the device-side s-gradient amplification.
#include <mutex>

#include "vxn_host.h"

namespace vxn {

constexpr int kCompactThreads = 256;  // bitmask words per block in the compaction

__global__ void __launch_bounds__(256) vr_mark_kernel(const int64_t* __restrict__ corners, int64_t n, int X, int Y,
                                                      int Z, uint32_t* __restrict__ bits) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t x = corners[3 * i], y = corners[3 * i + 1], z = corners[3 * i + 2];
    x = x < 0 ? 0 : (x > X - 1 ? X - 1 : x);  // active_vertices.py:33-35
    y = y < 0 ? 0 : (y > Y - 1 ? Y - 1 : y);
    z = z < 0 ? 0 : (z > Z - 1 ? Z - 1 : z);
    const int64_t lin = (x * Y + y) * Z + z;
    atomicOr(bits + (lin >> 5), 1u << (lin & 31));
  }
}

// Pass 1: per-block popcount totals.
__global__ void __launch_bounds__(kCompactThreads) vr_count_kernel(const uint32_t* __restrict__ bits, int64_t nwords,
                                                                   int32_t* __restrict__ block_tot,
                                                                   int32_t* __restrict__ counts) {
  __shared__ int32_t red[kCompactThreads / 64];
  const int64_t w = blockIdx.x * (int64_t)kCompactThreads + threadIdx.x;
  int c = w < nwords ? __popc(bits[w]) : 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int i = 0; i < kCompactThreads / 64; ++i) s += red[i];
    block_tot[blockIdx.x] = s;
    if (blockIdx.x == 0) counts[1] = 0;
  }
}

// Pass 2: exclusive scan of block totals (single work-group, sequential chunks).
__global__ void __launch_bounds__(256) vr_scan_kernel(int32_t* __restrict__ block_tot, int nblocks,
                                                      int32_t* __restrict__ counts) {
  __shared__ int32_t carry_s;
  __shared__ int32_t wsum[4];
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int base = 0; base < nblocks; base += 256) {
    const int i = base + threadIdx.x;
    const int v = i < nblocks ? block_tot[i] : 0;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int before = carry_s;
    for (int k = 0; k < wid; ++k) before += wsum[k];
    if (i < nblocks) block_tot[i] = before + incl - v;
    __syncthreads();
    if (threadIdx.x == 255) carry_s = before + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[0] = carry_s;
}

// Pass 3: write ascending linear indices; bit 31 flags "outside the relaxed sphere".
__global__ void __launch_bounds__(kCompactThreads) vr_write_kernel(const uint32_t* __restrict__ bits, int64_t nwords,
                                                                   const int32_t* __restrict__ block_off, int X,
                                                                   int Y, int Z, float mnx, float mny, float mnz,
                                                                   float ex, float ey, float ez, float relax,
                                                                   int32_t* __restrict__ out,
                                                                   int32_t* __restrict__ counts) {
  __shared__ int32_t wsum[kCompactThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t w = blockIdx.x * (int64_t)kCompactThreads + threadIdx.x;
  uint32_t word = w < nwords ? bits[w] : 0u;
  const int c = __popc(word);
  int incl = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  int pos = block_off[blockIdx.x] + incl - c;
  for (int k = 0; k < wid; ++k) pos += wsum[k];
  int inside_n = 0;
  const int64_t YZ = (int64_t)Y * Z;
  while (word) {
    const int b = __ffs(word) - 1;
    word &= word - 1;
    const int64_t lin = w * 32 + b;
    const int ix = (int)(lin / YZ);
    const int64_t r = lin - (int64_t)ix * YZ;
    const int iy = (int)(r / Z);
    const int iz = (int)(r - (int64_t)iy * Z);
    // world = xyz_min + idx * voxel_size; |world| < relax  (trainer.py:263-264)
    const float wx = mnx + (float)ix * ex, wy = mny + (float)iy * ey, wz = mnz + (float)iz * ez;
    const float rr = sqrtf(wx * wx + wy * wy + wz * wz);
    const bool inside = rr < relax;
    inside_n += inside ? 1 : 0;
    out[pos++] = (int32_t)lin | (inside ? 0 : (int32_t)0x80000000);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) inside_n += __shfl_xor(inside_n, o, 64);
  if (lane == 0 && inside_n) atomicAdd(counts + 1, inside_n);
}

constexpr int kRegThreads = 256;

// Fused eikonal + curvature over V_R.  Partial FP64 sums per block: {eik, curv, n_interior}.
__global__ void __launch_bounds__(kRegThreads) reg_eik_curv_kernel(GridDesc g, const int32_t* __restrict__ list,
                                                                   const int32_t* __restrict__ counts, float w_eik,
                                                                   float w_curv, float* __restrict__ gg,
                                                                   double* __restrict__ partial) {
  const int n_all = counts[0], n_in = counts[1];
  const bool filt = n_in > 0;
  const int nact = filt ? n_in : n_all;
  const float scale = (float)(2.0 / (double)(nact > 0 ? nact : 1));
  const float dx = 2.0f * g.eps[0], dy = 2.0f * g.eps[1], dz = 2.0f * g.eps[2];
  const float ex2 = g.eps[0] * g.eps[0], ey2 = g.eps[1] * g.eps[1], ez2 = g.eps[2] * g.eps[2];
  const int64_t YZ = (int64_t)g.Y * g.Z;
  double se = 0.0, sc = 0.0, cnt = 0.0;
  for (int i = blockIdx.x * kRegThreads + threadIdx.x; i < n_all; i += gridDim.x * kRegThreads) {
    const int32_t e = list[i];
    if (filt && e < 0) continue;
    const int64_t lin = e & 0x7fffffff;
    const int ix = (int)(lin / YZ);
    const int64_t r = lin - (int64_t)ix * YZ;
    const int iy = (int)(r / g.Z);
    const int iz = (int)(r - (int64_t)iy * g.Z);
    if (!(ix > 0 && ix < g.X - 1 && iy > 0 && iy < g.Y - 1 && iz > 0 && iz < g.Z - 1)) continue;
    const float* f = g.grid;
    const float fc = f[lin];
    const float fxp = f[lin + YZ], fxm = f[lin - YZ];
    const float fyp = f[lin + g.Z], fym = f[lin - g.Z];
    const float fzp = f[lin + 1], fzm = f[lin - 1];
    cnt += 1.0;
    // eikonal (eikonal.py:181-218)
    const float nx = (fxp - fxm) / dx, ny = (fyp - fym) / dy, nz = (fzp - fzm) / dz;
    const float nn = sqrtf(nx * nx + ny * ny + nz * nz);
    const double de = (double)nn - 1.0;
    se += de * de;
    if (w_eik != 0.0f) {
      const float sv = scale * (1.0f - 1.0f / (nn + 1e-8f));
      const float vx = sv * nx, vy = sv * ny, vz = sv * nz;
      atomicAdd(gg + lin + YZ, (vx / dx) * w_eik);
      atomicAdd(gg + lin - YZ, (-vx / dx) * w_eik);
      atomicAdd(gg + lin + g.Z, (vy / dy) * w_eik);
      atomicAdd(gg + lin - g.Z, (-vy / dy) * w_eik);
      atomicAdd(gg + lin + 1, (vz / dz) * w_eik);
      atomicAdd(gg + lin - 1, (-vz / dz) * w_eik);
    }
    // curvature (curvature.py:351-384)
    const float Lx = ((fxp + fxm) - 2.0f * fc) / ex2;
    const float Ly = ((fyp + fym) - 2.0f * fc) / ey2;
    const float Lz = ((fzp + fzm) - 2.0f * fc) / ez2;
    sc += (double)Lx * Lx + (double)Ly * Ly + (double)Lz * Lz;
    if (w_curv != 0.0f) {
      const float gc = scale * ((-2.0f * Lx / ex2 - 2.0f * Ly / ey2) - 2.0f * Lz / ez2);
      const float gx = scale * (Lx / ex2), gy = scale * (Ly / ey2), gz = scale * (Lz / ez2);
      atomicAdd(gg + lin, gc * w_curv);
      atomicAdd(gg + lin + YZ, gx * w_curv);
      atomicAdd(gg + lin - YZ, gx * w_curv);
      atomicAdd(gg + lin + g.Z, gy * w_curv);
      atomicAdd(gg + lin - g.Z, gy * w_curv);
      atomicAdd(gg + lin + 1, gz * w_curv);
      atomicAdd(gg + lin - 1, gz * w_curv);
    }
  }
  se = wave_sum_d(se);
  sc = wave_sum_d(sc);
  cnt = wave_sum_d(cnt);
  __shared__ double red[3][kRegThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = se;
    red[1][wid] = sc;
    red[2][wid] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0, c = 0;
    for (int k = 0; k < kRegThreads / 64; ++k) {
      a += red[0][k];
      b += red[1][k];
      c += red[2][k];
    }
    partial[3 * blockIdx.x] = a;
    partial[3 * blockIdx.x + 1] = b;
    partial[3 * blockIdx.x + 2] = c;
  }
}

// Deterministic final reduction of the per-block partials (fixed order).
__global__ void __launch_bounds__(64) reg_finish_kernel(const double* __restrict__ partial, int nblocks,
                                                        double* __restrict__ losses) {
  double a = 0, b = 0, c = 0;
  for (int i = threadIdx.x; i < nblocks; i += 64) {
    a += partial[3 * i];
    b += partial[3 * i + 1];
    c += partial[3 * i + 2];
  }
  a = wave_sum_d(a);
  b = wave_sum_d(b);
  c = wave_sum_d(c);
  if (threadIdx.x == 0) {
    losses[0] = c > 0 ? a / c : 0.0;
    losses[1] = c > 0 ? b / c : 0.0;
  }
}

constexpr int kRegBlocks = 1024;

// Reference-shaped regularisers over int64 [V,3] vertices (cuda/eikonal_grad.cu:8-69,
// cuda/curvature_grad.cu:8-71 semantics: boundary vertices skipped, 2/V scale).
__global__ void __launch_bounds__(256) eikonal_ref_kernel(GridDesc g, const int64_t* __restrict__ act, int64_t V,
                                                          float weight, float* __restrict__ gg) {
  const float scale0 = (float)(2.0 / (double)V);
  const float dx = 2.0f * g.eps[0], dy = 2.0f * g.eps[1], dz = 2.0f * g.eps[2];
  const int64_t YZ = (int64_t)g.Y * g.Z;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < V; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ix = act[3 * i], iy = act[3 * i + 1], iz = act[3 * i + 2];
    if (ix <= 0 || ix >= g.X - 1 || iy <= 0 || iy >= g.Y - 1 || iz <= 0 || iz >= g.Z - 1) continue;
    const int64_t lin = ix * YZ + iy * g.Z + iz;
    const float* f = g.grid;
    const float nx = (f[lin + YZ] - f[lin - YZ]) / dx;
    const float ny = (f[lin + g.Z] - f[lin - g.Z]) / dy;
    const float nz = (f[lin + 1] - f[lin - 1]) / dz;
    const float nn = sqrtf(nx * nx + ny * ny + nz * nz);
    const float sv = scale0 * (1.0f - 1.0f / (nn + 1e-8f));
    const float vx = sv * nx, vy = sv * ny, vz = sv * nz;
    atomicAdd(gg + lin + YZ, (vx / dx) * weight);
    atomicAdd(gg + lin - YZ, (-vx / dx) * weight);
    atomicAdd(gg + lin + g.Z, (vy / dy) * weight);
    atomicAdd(gg + lin - g.Z, (-vy / dy) * weight);
    atomicAdd(gg + lin + 1, (vz / dz) * weight);
    atomicAdd(gg + lin - 1, (-vz / dz) * weight);
  }
}

__global__ void __launch_bounds__(256) curvature_ref_kernel(GridDesc g, const int64_t* __restrict__ act, int64_t V,
                                                            float weight, float* __restrict__ gg) {
  const float scale = (float)(2.0 / (double)V);
  const float ex2 = g.eps[0] * g.eps[0], ey2 = g.eps[1] * g.eps[1], ez2 = g.eps[2] * g.eps[2];
  const int64_t YZ = (int64_t)g.Y * g.Z;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < V; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ix = act[3 * i], iy = act[3 * i + 1], iz = act[3 * i + 2];
    if (ix <= 0 || ix >= g.X - 1 || iy <= 0 || iy >= g.Y - 1 || iz <= 0 || iz >= g.Z - 1) continue;
    const int64_t lin = ix * YZ + iy * g.Z + iz;
    const float* f = g.grid;
    const float fc = f[lin];
    const float Lx = ((f[lin + YZ] + f[lin - YZ]) - 2.0f * fc) / ex2;
    const float Ly = ((f[lin + g.Z] + f[lin - g.Z]) - 2.0f * fc) / ey2;
    const float Lz = ((f[lin + 1] + f[lin - 1]) - 2.0f * fc) / ez2;
    const float gc = scale * ((-2.0f * Lx / ex2 - 2.0f * Ly / ey2) - 2.0f * Lz / ez2);
    const float gx = scale * (Lx / ex2), gy = scale * (Ly / ey2), gz = scale * (Lz / ez2);
    atomicAdd(gg + lin, gc * weight);
    atomicAdd(gg + lin + YZ, gx * weight);
    atomicAdd(gg + lin - YZ, gx * weight);
    atomicAdd(gg + lin + g.Z, gy * weight);
    atomicAdd(gg + lin - g.Z, gy * weight);
    atomicAdd(gg + lin + 1, gz * weight);
    atomicAdd(gg + lin - 1, gz * weight);
  }
}

// ------------------------------------------------------------------------------------
// V_R straight from the bitmask: per-block counts {|V_R|, |V_R inside the relaxed
// sphere|}, then the regularisers over the marked vertices tile by tile (below), so no
// vertex list is built.
// ------------------------------------------------------------------------------------

struct VrGeom {
  float mn[3];  // xyz_min
  float e[3];   // (xyz_max - xyz_min) / (res - 1), float32 (as the compaction computes it)
  float relax;  // relaxed-sphere radius
};

VXN_DEV bool vr_inside(const VrGeom& v, int ix, int iy, int iz) {
  const float wx = v.mn[0] + (float)ix * v.e[0], wy = v.mn[1] + (float)iy * v.e[1], wz = v.mn[2] + (float)iz * v.e[2];
  return sqrtf(wx * wx + wy * wy + wz * wz) < v.relax;
}

__global__ void __launch_bounds__(kCompactThreads) vr_count2_kernel(const uint32_t* __restrict__ bits, int64_t nwords,
                                                                    int X, int Y, int Z, VrGeom v,
                                                                    int32_t* __restrict__ part) {
  __shared__ int32_t red[2][kCompactThreads / 64];
  const int64_t w = blockIdx.x * (int64_t)kCompactThreads + threadIdx.x;
  const uint32_t word = w < nwords ? bits[w] : 0u;
  int c = __popc(word), ci = 0;
  if (word) {
    // the word's first vertex in 32-bit arithmetic (X*Y*Z < 2^31), then its 32 vertices
    // in order, z fastest; every bit is tested, so the loop has no divergence of its own
    const uint32_t lin0 = (uint32_t)w * 32u, YZ = (uint32_t)Y * (uint32_t)Z;
    int ix = (int)(lin0 / YZ);
    const uint32_t r = lin0 - (uint32_t)ix * YZ;
    int iy = (int)(r / (uint32_t)Z), iz = (int)(r - (uint32_t)iy * (uint32_t)Z);
#pragma unroll 8
    for (int b = 0; b < 32; ++b) {
      ci += ((word >> b) & 1u) && vr_inside(v, ix, iy, iz) ? 1 : 0;
      if (++iz == Z) {
        iz = 0;
        if (++iy == Y) {
          iy = 0;
          ++ix;
        }
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    c += __shfl_xor(c, o, 64);
    ci += __shfl_xor(ci, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = c;
    red[1][threadIdx.x >> 6] = ci;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0, b = 0;
    for (int i = 0; i < kCompactThreads / 64; ++i) {
      a += red[0][i];
      b += red[1][i];
    }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

__global__ void __launch_bounds__(256) vr_count_finish_kernel(const int32_t* __restrict__ part, int nblocks,
                                                              int32_t* __restrict__ counts) {
  __shared__ int32_t red[2][4];
  int a = 0, b = 0;
  for (int i = threadIdx.x; i < nblocks; i += 256) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    counts[0] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    counts[1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

constexpr int kGatherThreads = 256;

// Constants of the two regularisers' gradients, with the reciprocals of the six divisors
// (the divisions by them are correctly rounded through div_rn).
struct RegK {
  float scale, dx, dy, dz, ex2, ey2, ez2, w_eik, w_curv;
  float rdx, rdy, rdz, rex2, rey2, rez2;
};

// The terms a V_R vertex u adds to its six neighbours and (curvature) to itself, with the
// values its own losses need (eikonal.py:181-218: v = scale (1 - 1/(|n| + eps)) n, +v/d to
// u + e and -v/d to u - e; curvature.py:351-384: the second differences).
struct RegTerms {
  float xp, xm, yp, ym, zp, zm, cz;
  float nn, Lx, Ly, Lz;
};

VXN_DEV RegTerms reg_terms(float fc, float fxp, float fxm, float fyp, float fym, float fzp, float fzm, const RegK& k) {
  RegTerms t;
  const float nx = div_rn(fxp - fxm, k.dx, k.rdx), ny = div_rn(fyp - fym, k.dy, k.rdy),
              nz = div_rn(fzp - fzm, k.dz, k.rdz);
  t.nn = sqrtf(nx * nx + ny * ny + nz * nz);
  t.Lx = div_rn((fxp + fxm) - 2.0f * fc, k.ex2, k.rex2);
  t.Ly = div_rn((fyp + fym) - 2.0f * fc, k.ey2, k.rey2);
  t.Lz = div_rn((fzp + fzm) - 2.0f * fc, k.ez2, k.rez2);
  const float sv = k.scale * (1.0f - 1.0f / (t.nn + 1e-8f));
  const float ex = k.w_eik != 0.0f ? div_rn(sv * nx, k.dx, k.rdx) * k.w_eik : 0.0f;
  const float ey = k.w_eik != 0.0f ? div_rn(sv * ny, k.dy, k.rdy) * k.w_eik : 0.0f;
  const float ez = k.w_eik != 0.0f ? div_rn(sv * nz, k.dz, k.rdz) * k.w_eik : 0.0f;
  const float qx = div_rn(t.Lx, k.ex2, k.rex2), qy = div_rn(t.Ly, k.ey2, k.rey2), qz = div_rn(t.Lz, k.ez2, k.rez2);
  const float gx = k.w_curv != 0.0f ? (k.scale * qx) * k.w_curv : 0.0f;
  const float gy = k.w_curv != 0.0f ? (k.scale * qy) * k.w_curv : 0.0f;
  const float gz = k.w_curv != 0.0f ? (k.scale * qz) * k.w_curv : 0.0f;
  // -2 Lx / ex2 is RN(RN(-2 Lx) / ex2) = -2 RN(Lx / ex2): doubling is exact
  t.cz = k.w_curv != 0.0f ? (k.scale * ((-2.0f * qx - 2.0f * qy) - 2.0f * qz)) * k.w_curv : 0.0f;
  t.xp = ex + gx;
  t.xm = -ex + gx;
  t.yp = ey + gy;
  t.ym = -ey + gy;
  t.zp = ez + gz;
  t.zm = -ez + gz;
  return t;
}

// Tiles of kRegTX x kRegTY (x, y) rows by 64 z: each member vertex of the tile computes its
// terms once and adds them, with LDS float atomics, into the tile's accumulator, which
// spans the tile plus one vertex on every side; the accumulator is then added to grad_grid
// line by line (64 consecutive z per instruction, zero sums skipped), so a vertex gets one
// device atomic per tile that reaches it instead of one per term.  Tiles without a marked
// vertex are skipped after one load of their bitmask words.
constexpr int kRegTX = 8, kRegTY = 8;
constexpr int kRegHY = kRegTY + 2;
constexpr int kRegRows = (kRegTX + 2) * kRegHY;
constexpr int kRegLine = 66;  // z0 - 1 .. z0 + 64
constexpr int kRegGroup = 4;  // rows whose loads are in flight together

__global__ void __launch_bounds__(kGatherThreads) reg_vr_kernel(GridDesc g, const uint32_t* __restrict__ bits,
                                                                    VrGeom v, const int32_t* __restrict__ counts,
                                                                    float w_eik, float w_curv,
                                                                    float* __restrict__ gg,
                                                                    double* __restrict__ partial) {
  __shared__ float acc[kRegRows * kRegLine];
  const int n_all = counts[0], n_in = counts[1];
  const bool filt = n_in > 0;
  const int nact = filt ? n_in : n_all;
  RegK k;
  k.scale = (float)(2.0 / (double)(nact > 0 ? nact : 1));
  k.dx = 2.0f * g.eps[0];
  k.dy = 2.0f * g.eps[1];
  k.dz = 2.0f * g.eps[2];
  k.ex2 = g.eps[0] * g.eps[0];
  k.ey2 = g.eps[1] * g.eps[1];
  k.ez2 = g.eps[2] * g.eps[2];
  k.w_eik = w_eik;
  k.w_curv = w_curv;
  k.rdx = 1.0f / k.dx;
  k.rdy = 1.0f / k.dy;
  k.rdz = 1.0f / k.dz;
  k.rex2 = 1.0f / k.ex2;
  k.rey2 = 1.0f / k.ey2;
  k.rez2 = 1.0f / k.ez2;
  const int X = g.X, Y = g.Y, Z = g.Z;
  const int64_t YZ = (int64_t)Y * Z;
  const float* f = g.grid;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // a vertex is in V_R for the losses when marked, interior and (filtering) inside the sphere
  double se = 0.0, sc = 0.0, cnt = 0.0;
  const int ntx = (X + kRegTX - 1) / kRegTX, nty = (Y + kRegTY - 1) / kRegTY, ntz = (Z + 63) / 64;
  const int64_t ntiles = (int64_t)ntx * nty * ntz;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int tz = (int)(tile % ntz);
    const int64_t rest = tile / ntz;
    const int ty = (int)(rest % nty), tx = (int)(rest / nty);
    const int x0 = tx * kRegTX, y0 = ty * kRegTY, z0 = tz * 64;
    // skip a tile without a marked vertex: 64 bits of a row span at most 3 words
    bool marked = false;
    for (int q = threadIdx.x; q < kRegTX * kRegTY * 3; q += kGatherThreads) {
      const int r = q / 3, j = q - (q / 3) * 3;
      const int rx = x0 + r / kRegTY, ry = y0 + r % kRegTY;
      if (rx < X && ry < Y) {
        const int64_t rb = ((int64_t)rx * Y + ry) * Z;
        const int64_t b0 = rb + z0, b1 = rb + min(z0 + 63, Z - 1);
        const int64_t wd = (b0 >> 5) + j;
        if (wd <= (b1 >> 5)) marked = marked || bits[wd] != 0u;
      }
    }
    if (!__syncthreads_or(marked)) continue;
    for (int e = threadIdx.x; e < kRegRows * kRegLine; e += kGatherThreads) acc[e] = 0.0f;
    __syncthreads();
    const int wz = z0 + lane;
    // Wavefront w owns the tile's x slices 2w and 2w + 1 (rows r = 16 w + j: x = 2w + j / 8,
    // y = j % 8), kRegGroup rows at a time with every load of a group (membership words and
    // stencils, the stencil centre clamped into the interior) issued before any is used.  A
    // member's terms that stay in the wavefront's own slices (and their y halo) are added at
    // once with plain LDS read-add-writes: no other wavefront touches those rows.  The one
    // term that leaves (-x from slice 2w, +x from slice 2w + 1) waits in a register until the
    // barrier, after which its target row has no other writer.
    constexpr int kRows = kRegTX * kRegTY / (kGatherThreads / 64);
    static_assert(kRows == 2 * kRegTY, "a wavefront owns two x slices");
    float leave[kRows];
    uint32_t lmask = 0u;
#pragma unroll
    for (int g0 = 0; g0 < kRows; g0 += kRegGroup) {
      float sv[kRegGroup][7];
      uint32_t word[kRegGroup];
      int64_t wl[kRegGroup];
#pragma unroll
      for (int j = 0; j < kRegGroup; ++j) {
        const int jj = g0 + j;
        const int wx = x0 + 2 * wid + jj / kRegTY, wy = y0 + jj % kRegTY;
        const bool ok = wx < X && wy < Y && wz < Z;
        wl[j] = ok ? ((int64_t)wx * Y + wy) * Z + wz : -1;
        word[j] = ok ? bits[wl[j] >> 5] : 0u;
        const int cx = min(max(wx, 1), X - 2), cy = min(max(wy, 1), Y - 2), cz = min(max(wz, 1), Z - 2);
        const int64_t u = ((int64_t)cx * Y + cy) * Z + cz;
        sv[j][0] = f[u];
        sv[j][1] = f[u + YZ];
        sv[j][2] = f[u - YZ];
        sv[j][3] = f[u + Z];
        sv[j][4] = f[u - Z];
        sv[j][5] = f[u + 1];
        sv[j][6] = f[u - 1];
      }
#pragma unroll
      for (int j = 0; j < kRegGroup; ++j) {
        const int jj = g0 + j;
        const bool right = jj >= kRegTY;
        const int wx = x0 + 2 * wid + jj / kRegTY, wy = y0 + jj % kRegTY;
        const bool mem = wl[j] >= 0 && ((word[j] >> (wl[j] & 31)) & 1u) && wx > 0 && wx < X - 1 && wy > 0 &&
                         wy < Y - 1 && wz > 0 && wz < Z - 1 && (!filt || vr_inside(v, wx, wy, wz));
        if (__ballot(mem) == 0) {
          leave[jj] = 0.0f;
          continue;
        }
        RegTerms t = reg_terms(sv[j][0], sv[j][1], sv[j][2], sv[j][3], sv[j][4], sv[j][5], sv[j][6], k);
        if (mem) {
          const double de = (double)t.nn - 1.0;
          se += de * de;
          sc += (double)t.Lx * t.Lx + (double)t.Ly * t.Ly + (double)t.Lz * t.Lz;
          cnt += 1.0;
          lmask |= 1u << jj;
        } else {
          t.cz = t.zp = t.zm = t.yp = t.ym = t.xp = t.xm = 0.0f;
        }
        // every lane adds (zero when not a member); the z terms land on the adjacent lanes'
        // entries, so each add is ordered across the wavefront before the next
        const int h = ((wx - x0 + 1) * kRegHY + (wy - y0 + 1)) * kRegLine + lane + 1;
        acc[h] = acc[h] + t.cz;
        wave_sync();
        acc[h + 1] = acc[h + 1] + t.zp;
        wave_sync();
        acc[h - 1] = acc[h - 1] + t.zm;
        wave_sync();
        acc[h + kRegLine] = acc[h + kRegLine] + t.yp;
        wave_sync();
        acc[h - kRegLine] = acc[h - kRegLine] + t.ym;
        wave_sync();
        if (right) {
          acc[h - kRegHY * kRegLine] = acc[h - kRegHY * kRegLine] + t.xm;
          leave[jj] = t.xp;
        } else {
          acc[h + kRegHY * kRegLine] = acc[h + kRegHY * kRegLine] + t.xp;
          leave[jj] = t.xm;
        }
        wave_sync();
      }
    }
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < kRows; ++jj) {
      if (lmask & (1u << jj)) {
        const bool right = jj >= kRegTY;
        const int hx = 2 * wid + jj / kRegTY + 1 + (right ? 1 : -1), hy = jj % kRegTY + 1;
        const int h = (hx * kRegHY + hy) * kRegLine + lane + 1;
        acc[h] = acc[h] + leave[jj];
      }
    }
    __syncthreads();
    for (int r = wid; r < kRegRows; r += kGatherThreads / 64) {
      const int gx = x0 + r / kRegHY - 1, gy = y0 + r % kRegHY - 1;
      if (gx < 0 || gx >= X || gy < 0 || gy >= Y) continue;
      const int64_t rb = ((int64_t)gx * Y + gy) * Z;
      const float val = acc[r * kRegLine + lane + 1];
      if (val != 0.0f && wz < Z) atomicAdd(gg + rb + wz, val);
      if (lane < 2) {
        const int hz = lane == 0 ? 0 : kRegLine - 1;
        const int gz = z0 - 1 + hz;
        const float v2 = acc[r * kRegLine + hz];
        if (v2 != 0.0f && gz >= 0 && gz < Z) atomicAdd(gg + rb + gz, v2);
      }
    }
    __syncthreads();
  }
  se = wave_sum_d(se);
  sc = wave_sum_d(sc);
  cnt = wave_sum_d(cnt);
  __shared__ double red[3][kGatherThreads / 64];
  if (lane == 0) {
    red[0][wid] = se;
    red[1][wid] = sc;
    red[2][wid] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0, c = 0;
    for (int q = 0; q < kGatherThreads / 64; ++q) {
      a += red[0][q];
      b += red[1][q];
      c += red[2][q];
    }
    partial[3 * blockIdx.x] = a;
    partial[3 * blockIdx.x + 1] = b;
    partial[3 * blockIdx.x + 2] = c;
  }
}

// Post-upsample eikonal clamp (trainer.py:309-316): max |src| per work-group, then the
// scale factor, then dst += (src * w) * c.  NaN propagates through the maximum as
// torch's max does, and a NaN maximum leaves c = 1 (the reference's comparison is false).
constexpr int kClampBlocks = 1024;
constexpr int kClampThreads = 256;

VXN_DEV float nan_max(float a, float b) { return (a > b || a != a) ? a : b; }

VXN_DEV float block_nan_max(float m) {
  __shared__ float red[kClampThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = nan_max(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  m = red[0];
#pragma unroll
  for (int k = 1; k < kClampThreads / 64; ++k) m = nan_max(m, red[k]);
  return m;
}

__global__ void __launch_bounds__(kClampThreads) absmax_partial_kernel(const float* __restrict__ src, int64_t n,
                                                                       int vec, float* __restrict__ part) {
  float m = 0.0f;
  const int64_t stride = (int64_t)gridDim.x * kClampThreads;
  const int64_t t0 = blockIdx.x * (int64_t)kClampThreads + threadIdx.x;
  int64_t head = 0;
  if (vec) {
    const int64_t nv = n >> 2;
    const float4* s4 = (const float4*)src;
    for (int64_t i = t0; i < nv; i += stride) {
      const float4 v = s4[i];
      m = nan_max(m, nan_max(nan_max(fabsf(v.x), fabsf(v.y)), nan_max(fabsf(v.z), fabsf(v.w))));
    }
    head = nv << 2;
  }
  for (int64_t i = head + t0; i < n; i += stride) m = nan_max(m, fabsf(src[i]));
  m = block_nan_max(m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// part[nparts] <- the factor c of max |src * w| = max|src| * |w| (rounding is monotone)
__global__ void __launch_bounds__(kClampThreads) clamp_factor_kernel(float* __restrict__ part, int nparts, float w,
                                                                     float clamp) {
  float m = 0.0f;
  for (int i = threadIdx.x; i < nparts; i += kClampThreads) m = nan_max(m, part[i]);
  m = block_nan_max(m);
  if (threadIdx.x == 0) {
    const float mw = m * fabsf(w);
    part[nparts] = mw > clamp ? (1.0f / mw) * clamp : 1.0f;
  }
}

__global__ void __launch_bounds__(kClampThreads) clamped_add_kernel(float* __restrict__ dst,
                                                                    const float* __restrict__ src, int64_t n, int vec,
                                                                    const float* __restrict__ factor, float w) {
  const float c = *factor;
  const bool sc = c != 1.0f;
  const int64_t stride = (int64_t)gridDim.x * kClampThreads;
  const int64_t t0 = blockIdx.x * (int64_t)kClampThreads + threadIdx.x;
  int64_t head = 0;
  if (vec) {
    const int64_t nv = n >> 2;
    const float4* s4 = (const float4*)src;
    float4* d4 = (float4*)dst;
    for (int64_t i = t0; i < nv; i += stride) {
      const float4 v = s4[i];
      float4 o = d4[i];
      float g0 = v.x * w, g1 = v.y * w, g2 = v.z * w, g3 = v.w * w;
      if (sc) {
        g0 = g0 * c;
        g1 = g1 * c;
        g2 = g2 * c;
        g3 = g3 * c;
      }
      o.x = o.x + g0;
      o.y = o.y + g1;
      o.z = o.z + g2;
      o.w = o.w + g3;
      d4[i] = o;
    }
    head = nv << 2;
  }
  for (int64_t i = head + t0; i < n; i += stride) {
    float g = src[i] * w;
    if (sc) g = g * c;
    dst[i] = dst[i] + g;
  }
}

__global__ void s_amplify_kernel(float* grad_v, float k, int32_t* amplified) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const float gv = grad_v[0];
    const bool neg = gv < 0.0f;
    if (neg) grad_v[0] = gv * k;
    if (amplified) amplified[0] = neg ? 1 : 0;
  }
}

static int64_t vr_nwords(int X, int Y, int Z) { return ((int64_t)X * Y * Z + 31) / 32; }
static int vr_nblocks(int X, int Y, int Z) { return (int)((vr_nwords(X, Y, Z) + kCompactThreads - 1) / kCompactThreads); }

}  // namespace vxn

using namespace vxn;

extern "C" int vxn_vr_mark_corners(const int64_t* corners, int64_t n, int X, int Y, int Z, uint32_t* bitmask,
                                   void* stream) {
  VXN_REQUIRE(bitmask != nullptr, VXN_ERR_NULL, "bitmask is null");
  VXN_REQUIRE(X >= 1 && Y >= 1 && Z >= 1 && n >= 0, VXN_ERR_SHAPE, "bad shape");
  if (n == 0) return 0;
  VXN_REQUIRE(corners != nullptr, VXN_ERR_NULL, "corners is null");
  hipLaunchKernelGGL(vr_mark_kernel, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, (hipStream_t)stream, corners, n,
                     X, Y, Z, bitmask);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_vr_workspace_bytes(int X, int Y, int Z, int64_t* bytes) {
  VXN_REQUIRE(bytes != nullptr, VXN_ERR_NULL, "bytes is null");
  VXN_REQUIRE(X >= 1 && Y >= 1 && Z >= 1, VXN_ERR_SHAPE, "bad shape");
  *bytes = (int64_t)vr_nblocks(X, Y, Z) * (int64_t)sizeof(int32_t);
  return 0;
}

extern "C" int vxn_vr_compact(const uint32_t* bitmask, int X, int Y, int Z, const float* box6, float relax_radius,
                              int32_t* vr_list, int32_t* counts, void* workspace, int64_t workspace_bytes,
                              void* stream) {
  VXN_REQUIRE(bitmask && box6 && vr_list && counts, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(X >= 2 && Y >= 2 && Z >= 2, VXN_ERR_SHAPE, "bad shape");
  VXN_REQUIRE((int64_t)X * Y * Z < ((int64_t)1 << 31), VXN_ERR_SHAPE, "grid too large for int32 V_R entries");
  const int nb = vr_nblocks(X, Y, Z);
  VXN_REQUIRE(workspace != nullptr && workspace_bytes >= (int64_t)nb * 4, VXN_ERR_WORKSPACE,
              "V_R workspace needs %lld bytes", (long long)nb * 4);
  const int64_t nwords = vr_nwords(X, Y, Z);
  hipStream_t s = (hipStream_t)stream;
  int32_t* bt = (int32_t*)workspace;
  hipLaunchKernelGGL(vr_count_kernel, dim3(nb), dim3(kCompactThreads), 0, s, bitmask, nwords, bt, counts);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL(vr_scan_kernel, dim3(1), dim3(256), 0, s, bt, nb, counts);
  VXN_LAUNCH_CHECK();
  const int res[3] = {X, Y, Z};
  float e[3];
  for (int a = 0; a < 3; ++a) e[a] = (box6[3 + a] - box6[a]) / ((float)res[a] - 1.0f);
  hipLaunchKernelGGL(vr_write_kernel, dim3(nb), dim3(kCompactThreads), 0, s, bitmask, nwords, bt, X, Y, Z, box6[0],
                     box6[1], box6[2], e[0], e[1], e[2], relax_radius, vr_list, counts);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_reg_workspace_bytes(int X, int Y, int Z, int64_t* bytes) {
  VXN_REQUIRE(bytes != nullptr, VXN_ERR_NULL, "bytes is null");
  (void)X;
  (void)Y;
  (void)Z;
  *bytes = (int64_t)kRegBlocks * 3 * (int64_t)sizeof(double);
  return 0;
}

extern "C" int vxn_reg_eik_curv(const float* grid, int X, int Y, int Z, const float* eps3, const int32_t* vr_list,
                                const int32_t* counts, float w_eik, float w_curv, float* grad_grid, double* losses,
                                void* workspace, int64_t workspace_bytes, void* stream) {
  VXN_REQUIRE(grid && eps3 && vr_list && counts && grad_grid && losses, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(X >= 3 && Y >= 3 && Z >= 3, VXN_ERR_SHAPE, "grid resolution must be >= 3 per axis");
  VXN_REQUIRE((int64_t)X * Y * Z < ((int64_t)1 << 31), VXN_ERR_SHAPE, "grid too large for int32 V_R entries");
  for (int a = 0; a < 3; ++a) VXN_REQUIRE(eps3[a] > 0.0f, VXN_ERR_VALUE, "voxel size must be > 0");
  VXN_REQUIRE(workspace != nullptr && workspace_bytes >= (int64_t)kRegBlocks * 3 * 8, VXN_ERR_WORKSPACE,
              "regulariser workspace needs %lld bytes", (long long)kRegBlocks * 24);
  GridDesc g = {};
  g.grid = grid;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  for (int a = 0; a < 3; ++a) g.eps[a] = eps3[a];
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(reg_eik_curv_kernel, dim3(kRegBlocks), dim3(kRegThreads), 0, s, g, vr_list, counts, w_eik, w_curv,
                     grad_grid, (double*)workspace);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL(reg_finish_kernel, dim3(1), dim3(64), 0, s, (const double*)workspace, kRegBlocks, losses);
  VXN_LAUNCH_CHECK();
  return 0;
}

constexpr int kGatherBlocks = 2048;

extern "C" int vxn_reg_vr_workspace_bytes(int X, int Y, int Z, int64_t* bytes) {
  VXN_REQUIRE(bytes != nullptr, VXN_ERR_NULL, "bytes is null");
  VXN_REQUIRE(X >= 1 && Y >= 1 && Z >= 1, VXN_ERR_SHAPE, "bad shape");
  *bytes = (int64_t)kGatherBlocks * 3 * (int64_t)sizeof(double) + (int64_t)vr_nblocks(X, Y, Z) * 2 * 4;
  return 0;
}

static int check_reg_vr(const float* grid, int X, int Y, int Z, const float* eps3, const float* box6,
                        const uint32_t* bitmask, float* grad_grid, double* losses, int32_t* counts) {
  VXN_REQUIRE(grid && eps3 && box6 && bitmask && grad_grid && losses && counts, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(X >= 3 && Y >= 3 && Z >= 3, VXN_ERR_SHAPE, "grid resolution must be >= 3 per axis");
  VXN_REQUIRE((int64_t)X * Y * Z < ((int64_t)1 << 31), VXN_ERR_SHAPE, "grid too large for int32 V_R entries");
  for (int a = 0; a < 3; ++a) VXN_REQUIRE(eps3[a] > 0.0f, VXN_ERR_VALUE, "voxel size must be > 0");
  return 0;
}

// The regulariser pass over the V_R bitmask with its partial sums at `partial` (double
// [kGatherBlocks][3]) and `cpart` (int32 [vr_nblocks][2]); every block of the count and
// gather kernels overwrites its own partials, so the buffers need no clearing.
static int launch_reg_vr(const float* grid, int X, int Y, int Z, const float* eps3, const float* box6,
                         float relax_radius, const uint32_t* bitmask, float w_eik, float w_curv, float* grad_grid,
                         double* losses, int32_t* counts, double* partial, int32_t* cpart, hipStream_t s) {
  GridDesc g = {};
  g.grid = grid;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  for (int a = 0; a < 3; ++a) g.eps[a] = eps3[a];
  VrGeom v;
  const int res[3] = {X, Y, Z};
  for (int a = 0; a < 3; ++a) {
    v.mn[a] = box6[a];
    v.e[a] = (box6[3 + a] - box6[a]) / ((float)res[a] - 1.0f);
  }
  v.relax = relax_radius;
  const int nb = vr_nblocks(X, Y, Z);
  hipLaunchKernelGGL(vr_count2_kernel, dim3(nb), dim3(kCompactThreads), 0, s, bitmask, vr_nwords(X, Y, Z), X, Y, Z, v,
                     cpart);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL(vr_count_finish_kernel, dim3(1), dim3(256), 0, s, cpart, nb, counts);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL(reg_vr_kernel, dim3(kGatherBlocks), dim3(kGatherThreads), 0, s, g, bitmask, v, counts, w_eik,
                     w_curv, grad_grid, partial);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL(reg_finish_kernel, dim3(1), dim3(64), 0, s, (const double*)partial, kGatherBlocks, losses);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_reg_eik_curv_vr(const float* grid, int X, int Y, int Z, const float* eps3, const float* box6,
                                   float relax_radius, const uint32_t* bitmask, float w_eik, float w_curv,
                                   float* grad_grid, double* losses, int32_t* counts, void* workspace,
                                   int64_t workspace_bytes, void* stream) {
  if (int e = check_reg_vr(grid, X, Y, Z, eps3, box6, bitmask, grad_grid, losses, counts)) return e;
  const int nb = vr_nblocks(X, Y, Z);
  const int64_t need = (int64_t)kGatherBlocks * 3 * 8 + (int64_t)nb * 2 * 4;
  VXN_REQUIRE(workspace != nullptr && workspace_bytes >= need, VXN_ERR_WORKSPACE,
              "regulariser workspace needs %lld bytes", (long long)need);
  double* partial = (double*)workspace;
  int32_t* cpart = (int32_t*)((char*)workspace + (int64_t)kGatherBlocks * 3 * 8);
  return launch_reg_vr(grid, X, Y, Z, eps3, box6, relax_radius, bitmask, w_eik, w_curv, grad_grid, losses, counts,
                       partial, cpart, (hipStream_t)stream);
}

// ------------------------------------------------------------------------------------
// The same pass without a workspace argument: the partials live in a file-scope device
// array (one copy per GPU, held until the process ends).  Host code must not pass the
// array's name to a kernel (it is a host-side stand-in, not a device address): its
// address on each device is taken once with hipGetSymbolAddress, with that device
// current, and kept in reg_scratch_ptr[].  Calls that overlap on one GPU share the array
// and race on it.
// ------------------------------------------------------------------------------------
constexpr int kNoAllocMaxCountBlocks = (int)(((int64_t)VXN_REG_NOALLOC_MAX_VOXELS / 32 + kCompactThreads - 1) /
                                             kCompactThreads);
constexpr int64_t kNoAllocPartialBytes = (int64_t)kGatherBlocks * 3 * 8;
constexpr int64_t kNoAllocScratchBytes = kNoAllocPartialBytes + (int64_t)kNoAllocMaxCountBlocks * 2 * 4;
static_assert(kNoAllocScratchBytes == 180224, "176 KiB of scratch per GPU");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex reg_scratch_mu;
static void* reg_scratch_ptr[kMaxDevices];

// The device address of `scratch` on the current device (looked up once per device).
static int reg_scratch(void** out) {
  int dev = -1;
  VXN_HIP(hipGetDevice(&dev));
  VXN_REQUIRE(dev >= 0 && dev < kMaxDevices, VXN_ERR_VALUE, "device %d is beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(reg_scratch_mu);
  if (reg_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    VXN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    VXN_REQUIRE(p != nullptr, VXN_ERR_HIP, "no device address for the regulariser scratch on device %d", dev);
    reg_scratch_ptr[dev] = p;
  }
  *out = reg_scratch_ptr[dev];
  return 0;
}

extern "C" int vxn_reg_eik_curv_vr_noalloc(const float* grid, int X, int Y, int Z, const float* eps3,
                                           const float* box6, float relax_radius, const uint32_t* bitmask,
                                           float w_eik, float w_curv, float* grad_grid, double* losses,
                                           int32_t* counts, void* stream) {
  if (int e = check_reg_vr(grid, X, Y, Z, eps3, box6, bitmask, grad_grid, losses, counts)) return e;
  VXN_REQUIRE((int64_t)X * Y * Z <= (int64_t)VXN_REG_NOALLOC_MAX_VOXELS, VXN_ERR_SHAPE,
              "grid has %lld vertices; the entry point without a workspace takes at most %lld "
              "(use vxn_reg_eik_curv_vr with a workspace)",
              (long long)X * Y * Z, (long long)VXN_REG_NOALLOC_MAX_VOXELS);
  void* p = nullptr;
  if (int e = reg_scratch(&p)) return e;
  double* partial = (double*)p;
  int32_t* cpart = (int32_t*)((char*)p + kNoAllocPartialBytes);
  return launch_reg_vr(grid, X, Y, Z, eps3, box6, relax_radius, bitmask, w_eik, w_curv, grad_grid, losses, counts,
                       partial, cpart, (hipStream_t)stream);
}

extern "C" int vxn_eikonal_grad(const float* grid, int X, int Y, int Z, const int64_t* active, int64_t V,
                                float* grad_grid, const float* eps3, float weight, void* stream) {
  VXN_REQUIRE(grid && grad_grid && eps3, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(X >= 2 && Y >= 2 && Z >= 2 && V >= 0, VXN_ERR_SHAPE, "bad shape");
  if (V == 0) return 0;
  VXN_REQUIRE(active != nullptr, VXN_ERR_NULL, "active is null");
  GridDesc g = {};
  g.grid = grid;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  for (int a = 0; a < 3; ++a) g.eps[a] = eps3[a];
  hipLaunchKernelGGL(eikonal_ref_kernel, dim3(blocks_for(V, 256, 65536)), dim3(256), 0, (hipStream_t)stream, g,
                     active, V, weight, grad_grid);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_curvature_grad(const float* grid, int X, int Y, int Z, const int64_t* active, int64_t V,
                                  float* grad_grid, const float* eps3, float weight, void* stream) {
  VXN_REQUIRE(grid && grad_grid && eps3, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(X >= 2 && Y >= 2 && Z >= 2 && V >= 0, VXN_ERR_SHAPE, "bad shape");
  if (V == 0) return 0;
  VXN_REQUIRE(active != nullptr, VXN_ERR_NULL, "active is null");
  GridDesc g = {};
  g.grid = grid;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  for (int a = 0; a < 3; ++a) g.eps[a] = eps3[a];
  hipLaunchKernelGGL(curvature_ref_kernel, dim3(blocks_for(V, 256, 65536)), dim3(256), 0, (hipStream_t)stream, g,
                     active, V, weight, grad_grid);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_s_amplify(float* grad_v, float k, int32_t* amplified, void* stream) {
  VXN_REQUIRE(grad_v != nullptr, VXN_ERR_NULL, "grad_v is null");
  hipLaunchKernelGGL(s_amplify_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, grad_v, k, amplified);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_eik_clamp_add_workspace_bytes(int64_t* bytes) {
  VXN_REQUIRE(bytes != nullptr, VXN_ERR_NULL, "bytes is null");
  *bytes = (int64_t)(kClampBlocks + 1) * 4;
  return 0;
}

extern "C" int vxn_eik_clamp_add(float* dst, const float* src, int64_t n, float weight, float clamp, void* workspace,
                                 int64_t workspace_bytes, void* stream) {
  VXN_REQUIRE(n >= 0, VXN_ERR_SHAPE, "n must be >= 0");
  if (n == 0) return 0;
  VXN_REQUIRE(dst && src, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(!(dst < src + n && src < dst + n), VXN_ERR_VALUE, "dst and src must not overlap");
  VXN_REQUIRE(workspace != nullptr && workspace_bytes >= (int64_t)(kClampBlocks + 1) * 4, VXN_ERR_WORKSPACE,
              "clamp workspace needs %lld bytes", (long long)(kClampBlocks + 1) * 4);
  const int vec = (((uintptr_t)dst | (uintptr_t)src) & 15) == 0 ? 1 : 0;
  float* part = (float*)workspace;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(absmax_partial_kernel, dim3(kClampBlocks), dim3(kClampThreads), 0, s, src, n, vec, part);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL(clamp_factor_kernel, dim3(1), dim3(kClampThreads), 0, s, part, kClampBlocks, weight, clamp);
  VXN_LAUNCH_CHECK();
  const int64_t units = vec ? (n >> 2) : n;
  hipLaunchKernelGGL(clamped_add_kernel, dim3(blocks_for(units, kClampThreads, 8192)), dim3(kClampThreads), 0, s,
                     dst, src, n, vec, (const float*)(part + kClampBlocks), weight);
  VXN_LAUNCH_CHECK();
  return 0;
}
