// Device-side building blocks shared by the VoxNeuS gfx950 kernels.
//
// All translation units are compiled with -ffp-contract=off so that every float
// expression rounds exactly like the PyTorch CPU operations of the reference path
// (one rounding per elementwise op).  Fused multiply-adds are written explicitly
// (fmaf) only where the reference itself has no rounding order to preserve
// (accumulations whose order is not defined by the reference, MLP dot products).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define VXN_DEV __device__ __forceinline__
#define VXN_HD __host__ __device__

#include "vxn_fastdiv.h"

namespace vxn {

// ------------------------------------------------------------------------------------
// Geometry of one dense SDF grid: [X, Y, Z] float32, z fastest.
// ------------------------------------------------------------------------------------
struct GridDesc {
  const float* grid;  // X*Y*Z
  int X, Y, Z;
  float mn[3];     // xyz_min
  float den[3];    // (xyz_max - xyz_min) + 1e-8f           (sdf_grid.py:106, voxneus.py:144)
  float rden[3];   // RN(1 / den): the divisions by den are div_rn (correctly rounded)
  float rm1[3];    // float(res) - 1.0f                      (coordinates.py:28)
  float bound[3];  // float32(res - 1.0001)                  (coordinates.py:73)
  float eps[3];    // (xyz_max - xyz_min) / (res - 1)        (sdf_grid.py:83-88)
  int analytical;  // gradient mode: 0 = Eq. 7 interpolated, 1 = Eq. 4 analytical (sdf_grid.py:142-160)
};

VXN_DEV int64_t lin3(const GridDesc& g, int x, int y, int z) {
  return ((int64_t)x * g.Y + y) * g.Z + z;
}

// Correctly rounded a / b from y = RN(1 / b) (b > 0), by Markstein's residual correction
// (P. Markstein, "IA-64 and Elementary Functions", 2000, ch. 8; Muller et al., "Handbook
// of Floating-Point Arithmetic", the fma-based division theorems): with y correctly
// rounded and q within one ulp of a / b, the residual e = a - b q is exact in one fma and
// RN(q + e y) = RN(a / b).  q = RN(a y) is within 1.5 ulp, so one correction first brings
// it within one ulp and a second rounds correctly (an exact q gets zero residuals and is
// kept).  Without branches: the quotient takes the sign of a y (the division's sign,
// signed zeros included), and an infinite a y is the quotient of an infinite a (a NaN a
// stays NaN).  Quotients in the subnormal range may differ from the division in the last
// place.
VXN_DEV float div_rn(float a, float b, float y) {
  const float q0 = a * y;
  const float e0 = fmaf(-q0, b, a);
  const float q1 = fmaf(e0, y, q0);
  const float e1 = fmaf(-q1, b, a);
  const float q2 = copysignf(fmaf(e1, y, q1), q0);
  return __builtin_isinf(q0) ? q0 : q2;
}

// (p - mn) / den of one axis, correctly rounded (div_rn with y = RN(1 / den)).
VXN_DEV float over_den(const GridDesc& g, int a, float v) { return div_rn(v, g.den[a], g.rden[a]); }

// ------------------------------------------------------------------------------------
// exp and sigmoid exactly as PyTorch's CPU float kernels evaluate them: the vectorised
// path calls SLEEF's expf_u10 (Shibata, "SLEEF: A Portable Vectorized Library of C
// Standard Mathematical Functions", xexpf: Cody-Waite reduction by ln2 split in two
// parts, degree-5 minimax polynomial, ldexp by two power-of-two multiplies), and
// sigmoid(a) = 1 / (1 + exp(0 - a)).
// ------------------------------------------------------------------------------------
VXN_DEV float pow2i(int q) { return __int_as_float((q + 127) << 23); }

VXN_DEV float sleef_expf(float d) {
  const float kLog2e = 1.442695040888963407359924681001892137426645954152985934135449406931f;
  const float kL2U = 0.693145751953125f;
  const float kL2L = 1.428606765330187045e-06f;
  const float qf = rintf(d * kLog2e);
  const int q = (int)qf;
  float s = fmaf(qf, -kL2U, d);
  s = fmaf(qf, -kL2L, s);
  float u = 0.000198527617612853646278381f;
  u = fmaf(u, s, 0.00139304355252534151077271f);
  u = fmaf(u, s, 0.00833336077630519866943359f);
  u = fmaf(u, s, 0.0416664853692054748535156f);
  u = fmaf(u, s, 0.166666671633720397949219f);
  u = fmaf(u, s, 0.5f);
  u = 1.0f + fmaf(s * s, u, s);
  const int qh = q >> 1;
  u = (u * pow2i(qh)) * pow2i(q - qh);
  if (d < -104.0f) u = 0.0f;
  if (d > 100.0f) u = __int_as_float(0x7f800000);
  return u;
}

VXN_DEV float sigmoid_ref(float a) {
  const float e = sleef_expf(0.0f - a);
  return 1.0f / (1.0f + e);
}

// Orders one wavefront's memory operations across its lanes: what a lane wrote before the
// call is visible to every lane of the wavefront after it (a lane's LDS read-add-write may
// target the entry another lane updated just before; the compiler, which sees one lane at a
// time, could otherwise reorder the two).
VXN_DEV void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// NeuS sharpness from the variance parameter, s = clamp(exp(10 v), 1e-4, s_max)
// (models/variance.py:31), with the device expf that PyTorch's GPU exp also calls.
VXN_DEV float s_of_variance(float v, float s_max) { return fminf(fmaxf(expf(v * 10.0f), 1e-4f), s_max); }

// ------------------------------------------------------------------------------------
// Trilinear SDF as F.grid_sample(mode="bilinear", padding_mode="border",
// align_corners=True) on (z,y,x)-flipped [-1,1] coordinates (sdf_grid.py:90-120),
// with PyTorch's CPU 3-D kernel operation order.  Corner k = 4*ox + 2*oy + oz.
// ------------------------------------------------------------------------------------
struct GsStencil {
  int x0, y0, z0;   // floor corner
  float w[8];       // weights in corner order
  unsigned valid;   // bit k: corner k inside the grid
};

VXN_DEV float gs_unnorm(float p, float mn, float den, float rden, int size) {
  float n = div_rn(p - mn, den, rden);
  n = n * 2.0f - 1.0f;
  float c = ((n + 1.0f) / 2.0f) * (float)(size - 1);
  return fminf((float)(size - 1), fmaxf(c, 0.0f));
}

VXN_DEV GsStencil gs_stencil(const GridDesc& g, float px, float py, float pz) {
  const float cx = gs_unnorm(px, g.mn[0], g.den[0], g.rden[0], g.X);
  const float cy = gs_unnorm(py, g.mn[1], g.den[1], g.rden[1], g.Y);
  const float cz = gs_unnorm(pz, g.mn[2], g.den[2], g.rden[2], g.Z);
  const float fx = floorf(cx), fy = floorf(cy), fz = floorf(cz);
  const float gx = fx + 1.0f, gy = fy + 1.0f, gz = fz + 1.0f;
  GsStencil s;
  s.x0 = (int)fx;
  s.y0 = (int)fy;
  s.z0 = (int)fz;
  // PyTorch multiplies the W (z), H (y), D (x) factors in that order.
  s.w[0] = (gz - cz) * (gy - cy) * (gx - cx);
  s.w[1] = (cz - fz) * (gy - cy) * (gx - cx);
  s.w[2] = (gz - cz) * (cy - fy) * (gx - cx);
  s.w[3] = (cz - fz) * (cy - fy) * (gx - cx);
  s.w[4] = (gz - cz) * (gy - cy) * (cx - fx);
  s.w[5] = (cz - fz) * (gy - cy) * (cx - fx);
  s.w[6] = (gz - cz) * (cy - fy) * (cx - fx);
  s.w[7] = (cz - fz) * (cy - fy) * (cx - fx);
  unsigned v = 0;
  for (int k = 0; k < 8; ++k) {
    const int x = s.x0 + ((k >> 2) & 1), y = s.y0 + ((k >> 1) & 1), z = s.z0 + (k & 1);
    if (x < g.X && y < g.Y && z < g.Z) v |= 1u << k;
  }
  s.valid = v;
  return s;
}

// ld(x, y, z) returns the grid value at a vertex (global memory, or a staged copy).
template <class Ld>
VXN_DEV float gs_eval(const GsStencil& s, Ld&& ld) {
  float out = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (s.valid & (1u << k)) {
      const int x = s.x0 + ((k >> 2) & 1), y = s.y0 + ((k >> 1) & 1), z = s.z0 + (k & 1);
      out = out + ld(x, y, z) * s.w[k];
    }
  }
  return out;
}

VXN_DEV float gs_eval(const GridDesc& g, const GsStencil& s) {
  if (s.valid == 0xFFu) {
    // all 8 corners in the grid: the four z-pairs with one 8-byte load each, summed in
    // the same order
    float v[4][2];
#pragma unroll
    for (int q = 0; q < 4; ++q) __builtin_memcpy(v[q], g.grid + lin3(g, s.x0 + (q >> 1), s.y0 + (q & 1), s.z0), 8);
    float out = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) out = out + v[(k >> 1) & 3][k & 1] * s.w[k];
    return out;
  }
  return gs_eval(s, [&](int x, int y, int z) { return g.grid[lin3(g, x, y, z)]; });
}

// ------------------------------------------------------------------------------------
// Corner decomposition (coordinates.py:10-29, 55-124) and the Eq. 7 interpolated
// central-difference gradient (gradients.py:59-139).
// ------------------------------------------------------------------------------------
struct Corners {
  int c0[3], c1[3];
  float w[8];
  float d[3];  // fractional offsets in the cell
};

VXN_DEV Corners corner_decomp(const GridDesc& g, float px, float py, float pz) {
  const float p[3] = {px, py, pz};
  const int res[3] = {g.X, g.Y, g.Z};
  float d[3];
  Corners c;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float coord = over_den(g, a, p[a] - g.mn[a]) * g.rm1[a];
    const float cc = fminf(fmaxf(coord, 0.0f), g.bound[a]);
    int i0 = (int)floorf(cc);
    int i1 = i0 + 1;
    i0 = min(max(i0, 0), res[a] - 1);
    i1 = min(max(i1, 0), res[a] - 1);
    c.c0[a] = i0;
    c.c1[a] = i1;
    d[a] = cc - (float)i0;
    c.d[a] = d[a];
  }
  const float ox = 1.0f - d[0], oy = 1.0f - d[1], oz = 1.0f - d[2];
  c.w[0] = ox * oy * oz;
  c.w[1] = ox * oy * d[2];
  c.w[2] = ox * d[1] * oz;
  c.w[3] = ox * d[1] * d[2];
  c.w[4] = d[0] * oy * oz;
  c.w[5] = d[0] * oy * d[2];
  c.w[6] = d[0] * d[1] * oz;
  c.w[7] = d[0] * d[1] * d[2];
  return c;
}

VXN_DEV void corner_xyz(const Corners& c, int k, int& x, int& y, int& z) {
  x = ((k >> 2) & 1) ? c.c1[0] : c.c0[0];
  y = ((k >> 1) & 1) ? c.c1[1] : c.c0[1];
  z = (k & 1) ? c.c1[2] : c.c0[2];
}

// Per-axis finite-difference stencil of one vertex: (i+1, i-1) clamped, divisor
// max((i_p - i_m) * eps, eps).
VXN_DEV void fd_axis(int i, int res, float eps, int& ip, int& im, float& dv) {
  ip = min(i + 1, res - 1);
  im = max(i - 1, 0);
  dv = fmaxf((float)(ip - im) * eps, eps);
}

// The same, with y = RN(1 / dv) from reps = RN(1 / eps): dv is 2 eps inside the grid
// (ip - im = 2) and eps at a border, and halving is exact, so y = reps / 2 or reps.
VXN_DEV void fd_axis(int i, int res, float eps, float reps, int& ip, int& im, float& dv, float& y) {
  fd_axis(i, res, eps, ip, im, dv);
  y = (ip - im == 2) ? 0.5f * reps : reps;
}


// Eq. 7: grad = sum_k w_k n_k, summed in PyTorch's CPU order for an 8-long reduction
// over a strided dimension: ((t0 + t4) + (t1 + t5)) + (t2 + t6)) + (t3 + t7).
template <class Ld>
VXN_DEV void eq7_gradient(const GridDesc& g, const Corners& c, float out[3], Ld&& ld) {
  const float r0 = 1.0f / g.eps[0], r1 = 1.0f / g.eps[1], r2 = 1.0f / g.eps[2];
  float t[8][3];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    int x, y, z;
    corner_xyz(c, k, x, y, z);
    int ip, im;
    float dv, ry;
    fd_axis(x, g.X, g.eps[0], r0, ip, im, dv, ry);
    const float nx = div_rn(ld(ip, y, z) - ld(im, y, z), dv, ry);
    fd_axis(y, g.Y, g.eps[1], r1, ip, im, dv, ry);
    const float ny = div_rn(ld(x, ip, z) - ld(x, im, z), dv, ry);
    fd_axis(z, g.Z, g.eps[2], r2, ip, im, dv, ry);
    const float nz = div_rn(ld(x, y, ip) - ld(x, y, im), dv, ry);
    t[k][0] = c.w[k] * nx;
    t[k][1] = c.w[k] * ny;
    t[k][2] = c.w[k] * nz;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
    out[a] = (((t[0][a] + t[4][a]) + (t[1][a] + t[5][a])) + (t[2][a] + t[6][a])) + (t[3][a] + t[7][a]);
}

VXN_DEV void eq7_gradient(const GridDesc& g, const Corners& c, float out[3]) {
  eq7_gradient(g, c, out, [&](int x, int y, int z) { return g.grid[lin3(g, x, y, z)]; });
}

// The 32 distinct values of an interior cell's Eq. 7 stencil, read as z-columns with
// vector loads: col[2 ox + oy][j] = f(x0 + ox, y0 + oy, z0 - 1 + j), the cell's four
// columns over four z; xf[2 side + oy][j] = f(x0 - 1 + 3 side, y0 + oy, z0 + j) and
// yf[2 side + ox][j] = f(x0 + ox, y0 - 1 + 3 side, z0 + j), the face neighbours over two.
struct Brick {
  float col[4][4];
  float xf[4][2];
  float yf[4][2];
};

// Interior: the whole stencil lies inside the grid (c1 = c0 + 1 on every axis, and one
// more vertex on each side), so every fd_axis divisor is 2 eps.
VXN_DEV bool brick_interior(const GridDesc& g, const Corners& c) {
  return c.c0[0] >= 1 && c.c0[0] + 2 <= g.X - 1 && c.c0[1] >= 1 && c.c0[1] + 2 <= g.Y - 1 && c.c0[2] >= 1 &&
         c.c0[2] + 2 <= g.Z - 1;
}

VXN_DEV void load_brick(const GridDesc& g, const Corners& c, Brick& b) {
  const int x0 = c.c0[0], y0 = c.c0[1], z0 = c.c0[2];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int ox = q >> 1, oy = q & 1;
    __builtin_memcpy(b.col[q], g.grid + lin3(g, x0 + ox, y0 + oy, z0 - 1), 16);
    __builtin_memcpy(b.xf[q], g.grid + lin3(g, x0 - 1 + 3 * (q >> 1), y0 + (q & 1), z0), 8);
    __builtin_memcpy(b.yf[q], g.grid + lin3(g, x0 + (q & 1), y0 - 1 + 3 * (q >> 1), z0), 8);
  }
}

// eq7_gradient on an interior cell from its brick: the same operations in the same order
// (divisor 2 eps, reciprocal RN(1 / eps) / 2).
VXN_DEV void eq7_gradient_brick(const GridDesc& g, const Corners& c, const Brick& b, float out[3]) {
  const float r0 = 0.5f * (1.0f / g.eps[0]), r1 = 0.5f * (1.0f / g.eps[1]), r2 = 0.5f * (1.0f / g.eps[2]);
  const float d0 = fmaxf(2.0f * g.eps[0], g.eps[0]), d1 = fmaxf(2.0f * g.eps[1], g.eps[1]),
              d2 = fmaxf(2.0f * g.eps[2], g.eps[2]);
  float t[8][3];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int ox = (k >> 2) & 1, oy = (k >> 1) & 1, oz = k & 1;
    const float xp = ox ? b.xf[2 + oy][oz] : b.col[2 + oy][oz + 1];
    const float xm = ox ? b.col[oy][oz + 1] : b.xf[oy][oz];
    const float yp = oy ? b.yf[2 + ox][oz] : b.col[2 * ox + 1][oz + 1];
    const float ym = oy ? b.col[2 * ox][oz + 1] : b.yf[ox][oz];
    const float zp = b.col[2 * ox + oy][oz + 2], zm = b.col[2 * ox + oy][oz];
    t[k][0] = c.w[k] * div_rn(xp - xm, d0, r0);
    t[k][1] = c.w[k] * div_rn(yp - ym, d1, r1);
    t[k][2] = c.w[k] * div_rn(zp - zm, d2, r2);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
    out[a] = (((t[0][a] + t[4][a]) + (t[1][a] + t[5][a])) + (t[2][a] + t[6][a])) + (t[3][a] + t[7][a]);
}

// gs_eval from the brick when grid_sample's cell is the Eq. 7 cell with all 8 corners in
// the grid (the same products and sum order); false otherwise.
VXN_DEV bool gs_eval_brick(const GsStencil& s, const Corners& c, const Brick& b, float& out) {
  if (s.x0 != c.c0[0] || s.y0 != c.c0[1] || s.z0 != c.c0[2] || s.valid != 0xFFu) return false;
  float o = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) o = o + b.col[(k >> 1) & 3][(k & 1) + 1] * s.w[k];
  out = o;
  return true;
}

// Scatter dL/dgrad (the un-normalised Eq. 7 gradient) into the grid gradient:
// d n_k,a / d f[i_p] = +1/dv, d f[i_m] = -1/dv, weighted by w_k.
VXN_DEV void eq7_scatter(const GridDesc& g, const Corners& c, const float dg[3], float* grad_grid) {
  const float r0 = 1.0f / g.eps[0], r1 = 1.0f / g.eps[1], r2 = 1.0f / g.eps[2];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    int x, y, z;
    corner_xyz(c, k, x, y, z);
    int ip, im;
    float dv, ry;
    fd_axis(x, g.X, g.eps[0], r0, ip, im, dv, ry);
    float v = div_rn(c.w[k] * dg[0], dv, ry);
    atomicAdd(grad_grid + lin3(g, ip, y, z), v);
    atomicAdd(grad_grid + lin3(g, im, y, z), -v);
    fd_axis(y, g.Y, g.eps[1], r1, ip, im, dv, ry);
    v = div_rn(c.w[k] * dg[1], dv, ry);
    atomicAdd(grad_grid + lin3(g, x, ip, z), v);
    atomicAdd(grad_grid + lin3(g, x, im, z), -v);
    fd_axis(z, g.Z, g.eps[2], r2, ip, im, dv, ry);
    v = div_rn(c.w[k] * dg[2], dv, ry);
    atomicAdd(grad_grid + lin3(g, x, y, ip), v);
    atomicAdd(grad_grid + lin3(g, x, y, im), -v);
  }
}

// Eq. 4 (gradients.py:142-220): the derivative of the trilinear interpolant itself,
// grad_a = sum_k (d w_k / d x_a) f[c_k].  For corner k = 4*ox + 2*oy + oz,
// d w_k / d x = (ox ? +1 : -1) * Y_oy * Z_oz / eps_x with X_0 = 1 - dx, X_1 = dx (and
// likewise for y, z); the 8 terms are summed left to right, as PyTorch's CPU sum folds
// an 8-long contiguous row.
VXN_DEV void eq4_weights(const GridDesc& g, const Corners& c, float dw[3][8]) {
  const float dx = c.d[0], dy = c.d[1], dz = c.d[2];
  const float X[2] = {1.0f - dx, dx}, Y[2] = {1.0f - dy, dy}, Z[2] = {1.0f - dz, dz};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int ox = (k >> 2) & 1, oy = (k >> 1) & 1, oz = k & 1;
    const float px = Y[oy] * Z[oz], py = X[ox] * Z[oz], pz = X[ox] * Y[oy];
    dw[0][k] = (ox ? px : -px) / g.eps[0];
    dw[1][k] = (oy ? py : -py) / g.eps[1];
    dw[2][k] = (oz ? pz : -pz) / g.eps[2];
  }
}

VXN_DEV void eq4_gradient(const GridDesc& g, const Corners& c, float out[3]) {
  float dw[3][8], f[8];
  eq4_weights(g, c, dw);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    int x, y, z;
    corner_xyz(c, k, x, y, z);
    f[k] = g.grid[lin3(g, x, y, z)];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float acc = dw[a][0] * f[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) acc = acc + dw[a][k] * f[k];
    out[a] = acc;
  }
}

// d f[c_k] of the Eq. 4 gradient for dL/dgrad = dg.
VXN_DEV void eq4_corner_grads(const GridDesc& g, const Corners& c, const float dg[3], float v[8]) {
  float dw[3][8];
  eq4_weights(g, c, dw);
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = (dg[2] * dw[2][k] + dg[1] * dw[1][k]) + dg[0] * dw[0][k];
}

VXN_DEV void eq4_scatter(const GridDesc& g, const Corners& c, const float dg[3], float* grad_grid) {
  float v[8];
  eq4_corner_grads(g, c, dg, v);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    int x, y, z;
    corner_xyz(c, k, x, y, z);
    atomicAdd(grad_grid + lin3(g, x, y, z), v[k]);
  }
}

// The grid's gradient in its configured mode, and its backward.
VXN_DEV void grid_gradient(const GridDesc& g, const Corners& c, float out[3]) {
  if (g.analytical) eq4_gradient(g, c, out);
  else eq7_gradient(g, c, out);
}

VXN_DEV void grid_gradient_scatter(const GridDesc& g, const Corners& c, const float dg[3], float* grad_grid) {
  if (g.analytical) eq4_scatter(g, c, dg, grad_grid);
  else eq7_scatter(g, c, dg, grad_grid);
}

// V_R marks of one sample (regularization/active_vertices.py:12-50): the 8 corners of its
// cell, in a bitmask of the grid's vertices.  The two z-corners of each (x, y) column share
// a word unless the first ends one.
VXN_DEV void vr_mark_cell(const GridDesc& g, const Corners& cn, uint32_t* vr) {
#pragma unroll
  for (int k = 0; k < 8; k += 2) {
    int cx, cy, cz0, cz1;
    corner_xyz(cn, k, cx, cy, cz0);
    corner_xyz(cn, k + 1, cx, cy, cz1);
    const int64_t l0 = lin3(g, cx, cy, cz0), l1 = lin3(g, cx, cy, cz1);
    if ((l0 >> 5) == (l1 >> 5)) {
      atomicOr(vr + (l0 >> 5), (1u << (l0 & 31)) | (1u << (l1 & 31)));
    } else {
      atomicOr(vr + (l0 >> 5), 1u << (l0 & 31));
      atomicOr(vr + (l1 >> 5), 1u << (l1 & 31));
    }
  }
}

VXN_DEV void gs_scatter(const GridDesc& g, const GsStencil& s, float dsdf, float* grad_grid) {
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (s.valid & (1u << k)) {
      const int x = s.x0 + ((k >> 2) & 1), y = s.y0 + ((k >> 1) & 1), z = s.z0 + (k & 1);
      atomicAdd(grad_grid + lin3(g, x, y, z), s.w[k] * dsdf);
    }
  }
}

// normal = grad / (|grad| + 1e-8)  (gradients.py:136-137)
VXN_DEV float norm3(const float v[3]) { return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// ------------------------------------------------------------------------------------
// Multiresolution hash (radiance_hash.py:53-69): (x*1) ^ (y*2654435761) ^ (z*805459861)
// mod T.  Coordinates are non-negative and T is a power of two, so the int64 remainder
// equals the low bits of the 32-bit wrapped products.
// ------------------------------------------------------------------------------------
VXN_DEV uint32_t hash3(uint32_t x, uint32_t y, uint32_t z, uint32_t mask) {
  return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & mask;
}

struct HashCell {
  uint32_t idx[8];
  float w[8];
  float dx, dy, dz;  // position in the cell: w[k] = (x weight * y weight) * z weight
};

VXN_DEV HashCell hash_cell(float x0, float x1, float x2, float res, uint32_t mask) {
  const float s0 = x0 * res, s1 = x1 * res, s2 = x2 * res;
  const float f0 = floorf(s0), f1 = floorf(s1), f2 = floorf(s2);
  const uint32_t c0 = (uint32_t)(int)f0, c1 = (uint32_t)(int)f1, c2 = (uint32_t)(int)f2;
  const float dx = s0 - f0, dy = s1 - f1, dz = s2 - f2;
  const float ox = 1.0f - dx, oy = 1.0f - dy, oz = 1.0f - dz;
  HashCell h;
  h.dx = dx;
  h.dy = dy;
  h.dz = dz;
  h.w[0] = ox * oy * oz;
  h.w[1] = ox * oy * dz;
  h.w[2] = ox * dy * oz;
  h.w[3] = ox * dy * dz;
  h.w[4] = dx * oy * oz;
  h.w[5] = dx * oy * dz;
  h.w[6] = dx * dy * oz;
  h.w[7] = dx * dy * dz;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    h.idx[k] = hash3(c0 + ((k >> 2) & 1), c1 + ((k >> 1) & 1), c2 + (k & 1), mask);
  return h;
}

// NeRF positional encoding of a view direction (embedder.py:12-44 with num_freqs 4,
// include_input, log-spaced bands 1, 2, 4, 8): [d, sin(d), cos(d), sin(2d), cos(2d), ...,
// sin(8d), cos(8d)], each a 3-vector: 27 values.  The band products are exact.
VXN_DEV void pe27(float x, float y, float z, float o[27]) {
  o[0] = x;
  o[1] = y;
  o[2] = z;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const float b = (float)(1 << f);
    o[3 + 6 * f] = sinf(x * b);
    o[4 + 6 * f] = sinf(y * b);
    o[5 + 6 * f] = sinf(z * b);
    o[6 + 6 * f] = cosf(x * b);
    o[7 + 6 * f] = cosf(y * b);
    o[8 + 6 * f] = cosf(z * b);
  }
}

// Real spherical harmonics, degree 0-3 (embedder.py:55-87), same operation order.
VXN_DEV void sh16(float x, float y, float z, float o[16]) {
  o[0] = 0.28209479177387814f;
  o[1] = -0.4886025119029199f * y;
  o[2] = 0.4886025119029199f * z;
  o[3] = -0.4886025119029199f * x;
  o[4] = 1.0925484305920792f * x * y;
  o[5] = -1.0925484305920792f * y * z;
  o[6] = 0.31539156525252005f * (2.0f * z * z - x * x - y * y);
  o[7] = -1.0925484305920792f * x * z;
  o[8] = 0.5462742152960396f * (x * x - y * y);
  o[9] = -0.5900435899266435f * y * (3.0f * x * x - y * y);
  o[10] = 2.890611442640554f * x * y * z;
  o[11] = -0.4570457994644658f * y * (4.0f * z * z - x * x - y * y);
  o[12] = 0.3731763325901154f * z * (2.0f * z * z - 3.0f * x * x - 3.0f * y * y);
  o[13] = -0.4570457994644658f * x * (4.0f * z * z - x * x - y * y);
  o[14] = 1.445305721320277f * z * (x * x - y * y);
  o[15] = -0.5900435899266435f * x * (x * x - 3.0f * y * y);
}

// ------------------------------------------------------------------------------------
// Wave64 helpers
// ------------------------------------------------------------------------------------
VXN_DEV float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

VXN_DEV double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace vxn
