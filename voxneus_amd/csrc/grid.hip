// Dense SDF grid queries: trilinear SDF + Eq. 7 interpolated gradient, their backward,
// the reference-shaped gradient-interp pair and the full-lattice mesh query.
#include "vxn_host.h"
#include "vxn_scene.h"

namespace vxn {

__global__ void __launch_bounds__(256) grid_query_fwd_kernel(GridDesc g, const float* __restrict__ xyz, int64_t n,
                                                             float* __restrict__ sdf, float* __restrict__ grad,
                                                             float* __restrict__ normal,
                                                             int64_t* __restrict__ corners,
                                                             float* __restrict__ cweights) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
    if (sdf) {
      const GsStencil s = gs_stencil(g, px, py, pz);
      sdf[i] = gs_eval(g, s);
    }
    const Corners c = corner_decomp(g, px, py, pz);
    if (grad || normal) {
      float gv[3];
      grid_gradient(g, c, gv);
      if (grad) {
        grad[3 * i] = gv[0];
        grad[3 * i + 1] = gv[1];
        grad[3 * i + 2] = gv[2];
      }
      if (normal) {
        const float r = norm3(gv) + 1e-8f;
        normal[3 * i] = gv[0] / r;
        normal[3 * i + 1] = gv[1] / r;
        normal[3 * i + 2] = gv[2] / r;
      }
    }
    if (corners) {
      for (int k = 0; k < 8; ++k) {
        int x, y, z;
        corner_xyz(c, k, x, y, z);
        corners[24 * i + 3 * k] = x;
        corners[24 * i + 3 * k + 1] = y;
        corners[24 * i + 3 * k + 2] = z;
      }
    }
    if (cweights)
      for (int k = 0; k < 8; ++k) cweights[8 * i + k] = c.w[k];
  }
}

__global__ void __launch_bounds__(256) grid_query_bwd_kernel(GridDesc g, const float* __restrict__ xyz, int64_t n,
                                                             const float* __restrict__ d_sdf,
                                                             const float* __restrict__ d_grad,
                                                             float* __restrict__ grad_grid) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
    if (d_sdf) {
      const GsStencil s = gs_stencil(g, px, py, pz);
      gs_scatter(g, s, d_sdf[i], grad_grid);
    }
    if (d_grad) {
      const Corners c = corner_decomp(g, px, py, pz);
      const float dg[3] = {d_grad[3 * i], d_grad[3 * i + 1], d_grad[3 * i + 2]};
      grid_gradient_scatter(g, c, dg, grad_grid);
    }
  }
}

// Continuous grid coordinates in, as the reference's compiled backend takes them.
VXN_DEV Corners corners_from_coords(const GridDesc& g, float cx, float cy, float cz) {
  const float p[3] = {cx, cy, cz};
  const int res[3] = {g.X, g.Y, g.Z};
  float d[3];
  Corners c;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float cc = fminf(fmaxf(p[a], 0.0f), g.bound[a]);
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

// Gradient (and optionally its unit normal) on continuous grid coordinates, in the mode
// the descriptor names.
__global__ void __launch_bounds__(256) coords_gradient_fwd_kernel(GridDesc g, const float* __restrict__ coords,
                                                                  int64_t n, float* __restrict__ out,
                                                                  float* __restrict__ normal) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const Corners c = corners_from_coords(g, coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]);
    float gv[3];
    grid_gradient(g, c, gv);
    if (out) {
      out[3 * i] = gv[0];
      out[3 * i + 1] = gv[1];
      out[3 * i + 2] = gv[2];
    }
    if (normal) {
      const float r = norm3(gv) + 1e-8f;
      normal[3 * i] = gv[0] / r;
      normal[3 * i + 1] = gv[1] / r;
      normal[3 * i + 2] = gv[2] / r;
    }
  }
}

__global__ void __launch_bounds__(256) coords_gradient_bwd_kernel(GridDesc g, const float* __restrict__ coords,
                                                                  int64_t n, const float* __restrict__ gout,
                                                                  float* __restrict__ grad_grid) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const Corners c = corners_from_coords(g, coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]);
    const float dg[3] = {gout[3 * i], gout[3 * i + 1], gout[3 * i + 2]};
    grid_gradient_scatter(g, c, dg, grad_grid);
  }
}

__global__ void __launch_bounds__(256) voxel_corners_kernel(GridDesc g, const float* __restrict__ coords, int64_t n,
                                                            int64_t* __restrict__ corners,
                                                            float* __restrict__ cweights) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const Corners c = corners_from_coords(g, coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]);
    if (corners) {
      for (int k = 0; k < 8; ++k) {
        int x, y, z;
        corner_xyz(c, k, x, y, z);
        corners[24 * i + 3 * k] = x;
        corners[24 * i + 3 * k + 1] = y;
        corners[24 * i + 3 * k + 2] = z;
      }
    }
    if (cweights)
      for (int k = 0; k < 8; ++k) cweights[8 * i + k] = c.w[k];
  }
}

// Lattice query: one thread per lattice point; consecutive threads walk z, so the
// 8-corner and 24-neighbour stencil loads of a wavefront fall on a few cache lines.  An
// interior cell's stencil is read as 12 z-columns with vector loads (load_brick) and
// evaluated with the generic code's operations (bit-identical results).
// One lattice row (ix, iy) per work-group, z across the threads.  XCD-aware: work-groups
// are dispatched to the 8 XCDs round-robin (work-group b on XCD b % 8), and XCD k walks
// the rows of its own slab of ix in order, so each XCD's L2 streams one eighth of the
// grid (plus a one-slab halo) instead of every XCD re-reading all of it.
__global__ void __launch_bounds__(256) lattice_query_kernel(GridDesc g, int NX, int NY, int NZ, float lx0, float ly0,
                                                            float lz0, float lsx, float lsy, float lsz,
                                                            float* __restrict__ sdf, float* __restrict__ normal) {
  const int xcd = blockIdx.x & 7, j = blockIdx.x >> 3;
  const int slab = (NX + 7) / 8;
  const int ix = xcd * slab + j / NY, iy = j - (j / NY) * NY;
  if (ix >= NX || j / NY >= slab) return;
  const float px = lx0 + (float)ix * lsx, py = ly0 + (float)iy * lsy;
  const int64_t row = ((int64_t)ix * NY + iy) * NZ;
  for (int iz = threadIdx.x; iz < NZ; iz += blockDim.x) {
    const int64_t i = row + iz;
    const float pz = lz0 + (float)iz * lsz;
    const GsStencil s = gs_stencil(g, px, py, pz);
    const Corners c = corner_decomp(g, px, py, pz);
    float sd, gv[3];
    if (!g.analytical && brick_interior(g, c)) {
      Brick bk;
      load_brick(g, c, bk);
      if (!gs_eval_brick(s, c, bk, sd)) sd = gs_eval(g, s);
      eq7_gradient_brick(g, c, bk, gv);
    } else {
      sd = gs_eval(g, s);
      grid_gradient(g, c, gv);
    }
    sdf[i] = sd;
    const float rr = norm3(gv) + 1e-8f;
    normal[3 * i] = gv[0] / rr;
    normal[3 * i + 1] = gv[1] / rr;
    normal[3 * i + 2] = gv[2] / rr;
  }
}

static GridDesc coords_desc(const float* grid, int X, int Y, int Z, const float* eps3) {
  GridDesc g;
  g.grid = grid;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  const int res[3] = {X, Y, Z};
  for (int a = 0; a < 3; ++a) {
    g.mn[a] = 0.0f;
    g.den[a] = 1.0f;
    g.rden[a] = 1.0f;
    g.rm1[a] = (float)res[a] - 1.0f;
    g.bound[a] = (float)((double)res[a] - 1.0001);
    g.eps[a] = eps3[a];
  }
  g.analytical = 0;
  return g;
}

// Progressive super-resolution (geometry/upsample.py:12-39): F.interpolate(trilinear,
// align_corners=True) of the SDF grid to a new resolution, with PyTorch's CPU kernel
// arithmetic: per axis, ratio = (in - 1) / (out - 1) (formed on the host, in float) and
// src = ratio * o in float, i0 = floor(src), l1 = clamp(src - i0, 0, 1), l0 = 1 - l1 (a
// copy when in == out); the 8 corners are combined z first, then y, then x, each pair as
// fma(t0, l0, t1 * l1).  One output row (ox, oy) per work-group, z across the threads.
struct UpAxis {
  int n_in, n_out;
  float ratio;
};

VXN_DEV void upsample_axis(int o, const UpAxis& ax, int& i0, int& i1, float& l0, float& l1) {
  if (ax.n_in == ax.n_out) {
    i0 = i1 = o;
    l0 = 1.0f;
    l1 = 0.0f;
    return;
  }
  const float src = ax.ratio * (float)o;
  i0 = min((int)floorf(src), ax.n_in - 1);
  l1 = fminf(fmaxf(src - (float)i0, 0.0f), 1.0f);
  i1 = i0 + (i0 < ax.n_in - 1 ? 1 : 0);
  l0 = 1.0f - l1;
}

__global__ void __launch_bounds__(256) upsample_grid_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                            UpAxis ux, UpAxis uy, UpAxis uz) {
  const int ox = blockIdx.x / uy.n_out, oy = blockIdx.x - ox * uy.n_out;
  int x[2], y[2];
  float a[2], b[2];
  upsample_axis(ox, ux, x[0], x[1], a[0], a[1]);
  upsample_axis(oy, uy, y[0], y[1], b[0], b[1]);
  const int Y = uy.n_in, Z = uz.n_in;
  const float* r00 = in + ((int64_t)x[0] * Y + y[0]) * Z;
  const float* r01 = in + ((int64_t)x[0] * Y + y[1]) * Z;
  const float* r10 = in + ((int64_t)x[1] * Y + y[0]) * Z;
  const float* r11 = in + ((int64_t)x[1] * Y + y[1]) * Z;
  float* orow = out + ((int64_t)ox * uy.n_out + oy) * uz.n_out;
  for (int oz = threadIdx.x; oz < uz.n_out; oz += blockDim.x) {
    int z0, z1;
    float c0, c1;
    upsample_axis(oz, uz, z0, z1, c0, c1);
    const float u00 = fmaf(r00[z0], c0, r00[z1] * c1), u01 = fmaf(r01[z0], c0, r01[z1] * c1);
    const float u10 = fmaf(r10[z0], c0, r10[z1] * c1), u11 = fmaf(r11[z0], c0, r11[z1] * c1);
    const float t0 = fmaf(u00, b[0], u01 * b[1]), t1 = fmaf(u10, b[0], u11 * b[1]);
    orow[oz] = fmaf(t0, a[0], t1 * a[1]);
  }
}

static UpAxis up_axis(int n_in, int n_out) {
  UpAxis a;
  a.n_in = n_in;
  a.n_out = n_out;
  a.ratio = n_out > 1 ? (float)(n_in - 1) / (float)(n_out - 1) : 0.0f;
  return a;
}

// SDF + unit normal over a packed batch of points of many scenes (the serving query).
__global__ void __launch_bounds__(256) serve_query_kernel(const SceneDev* __restrict__ scenes, int n_scenes,
                                                          const int32_t* __restrict__ pt_scene,
                                                          const int32_t* __restrict__ n_dev, int64_t cap,
                                                          const float* __restrict__ xyz, float* __restrict__ sdf,
                                                          float* __restrict__ normal) {
  const int64_t n = min((int64_t)max(*n_dev, 0), cap);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int si = pt_scene[i];
    if (si < 0 || si >= n_scenes) {
      sdf[i] = 0.0f;
      if (normal) normal[3 * i] = normal[3 * i + 1] = normal[3 * i + 2] = 0.0f;
      continue;
    }
    const GridDesc& g = scenes[si].g;
    const float px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
    sdf[i] = gs_eval(g, gs_stencil(g, px, py, pz));
    if (normal) {
      float gv[3];
      grid_gradient(g, corner_decomp(g, px, py, pz), gv);
      const float r = norm3(gv) + 1e-8f;
      normal[3 * i] = gv[0] / r;
      normal[3 * i + 1] = gv[1] / r;
      normal[3 * i + 2] = gv[2] / r;
    }
  }
}

// ------------------------------------------------------------------------------------
// Separable Gaussian filter with edge replication (scipy.ndimage.gaussian_filter,
// mode="nearest", as apply_gaussian_filter_to_grid calls it): one 1-D correlation per
// axis, x then y then z, each accumulated in double and rounded to float, as scipy's
// correlate1d does line by line.
// ------------------------------------------------------------------------------------
constexpr int kGaussMaxRadius = 32;

struct GaussTaps {
  double w[2 * kGaussMaxRadius + 1];  // correlation weights, centre at w[r]
  int r;
};

__global__ void __launch_bounds__(256) gauss1d_kernel(const float* __restrict__ in, float* __restrict__ out, int X,
                                                      int Y, int Z, int axis, GaussTaps k) {
  const int64_t n = (int64_t)X * Y * Z;
  const int64_t stride = axis == 0 ? (int64_t)Y * Z : (axis == 1 ? (int64_t)Z : 1);
  const int len = axis == 0 ? X : (axis == 1 ? Y : Z);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)((i / stride) % len);
    const float* c = in + i - (int64_t)p * stride;  // start of the line through i
    double acc = (double)in[i] * k.w[k.r];
    for (int j = k.r; j >= 1; --j) {
      const int lo = max(p - j, 0), hi = min(p + j, len - 1);
      acc = acc + ((double)c[(int64_t)lo * stride] + (double)c[(int64_t)hi * stride]) * k.w[k.r - j];
    }
    out[i] = (float)acc;
  }
}

}  // namespace vxn

using namespace vxn;

extern "C" int vxn_serve_query(const void* scenes, int32_t n_scenes, const int32_t* pt_scene, const int32_t* n_points,
                               int64_t capacity, const float* xyz, float* sdf, float* normal, void* stream) {
  VXN_REQUIRE(capacity >= 0, VXN_ERR_SHAPE, "capacity must be >= 0");
  if (capacity == 0) return 0;
  VXN_REQUIRE(scenes && pt_scene && n_points && xyz && sdf, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(n_scenes >= 1, VXN_ERR_VALUE, "n_scenes must be >= 1");
  const int64_t nb = min((capacity + 255) / 256, (int64_t)8192);
  hipLaunchKernelGGL(serve_query_kernel, dim3((int)nb), dim3(256), 0, (hipStream_t)stream, (const SceneDev*)scenes,
                     n_scenes, pt_scene, n_points, capacity, xyz, sdf, normal);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_grid_query_fwd(const float* grid, int X, int Y, int Z, const float* box6, const float* xyz,
                                  int64_t n, float* sdf, float* grad, float* normal, int64_t* corners,
                                  float* corner_weights, int32_t gradient_mode, void* stream) {
  if (int e = check_grid(grid, X, Y, Z, box6)) return e;
  if (int e = check_gradient_mode(gradient_mode)) return e;
  VXN_REQUIRE(n >= 0, VXN_ERR_SHAPE, "n must be >= 0");
  if (n == 0) return 0;
  VXN_REQUIRE(xyz != nullptr, VXN_ERR_NULL, "xyz is null");
  GridDesc g = make_grid_desc(grid, X, Y, Z, box6);
  g.analytical = gradient_mode;
  hipLaunchKernelGGL(grid_query_fwd_kernel, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, (hipStream_t)stream, g,
                     xyz, n, sdf, grad, normal, corners, corner_weights);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_grid_query_bwd(const float* grid, int X, int Y, int Z, const float* box6, const float* xyz,
                                  int64_t n, const float* d_sdf, const float* d_grad, float* grad_grid,
                                  int32_t gradient_mode, void* stream) {
  if (int e = check_grid(grid, X, Y, Z, box6)) return e;
  if (int e = check_gradient_mode(gradient_mode)) return e;
  VXN_REQUIRE(n >= 0, VXN_ERR_SHAPE, "n must be >= 0");
  if (n == 0 || (!d_sdf && !d_grad)) return 0;
  VXN_REQUIRE(xyz != nullptr && grad_grid != nullptr, VXN_ERR_NULL, "xyz or grad_grid is null");
  GridDesc g = make_grid_desc(grid, X, Y, Z, box6);
  g.analytical = gradient_mode;
  hipLaunchKernelGGL(grid_query_bwd_kernel, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, (hipStream_t)stream, g,
                     xyz, n, d_sdf, d_grad, grad_grid);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_gradient_interp_fwd(const float* grid, int X, int Y, int Z, const float* coords, int64_t n,
                                       const float* eps3, float* out, void* stream) {
  VXN_REQUIRE(grid && coords && eps3 && out, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(X >= 2 && Y >= 2 && Z >= 2 && n >= 0, VXN_ERR_SHAPE, "bad shape");
  if (n == 0) return 0;
  const GridDesc g = coords_desc(grid, X, Y, Z, eps3);
  hipLaunchKernelGGL(coords_gradient_fwd_kernel, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, (hipStream_t)stream,
                     g, coords, n, out, nullptr);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_gradient_interp_bwd(const float* grad_out, const float* coords, int64_t n, float* grad_grid, int X,
                                       int Y, int Z, const float* eps3, void* stream) {
  VXN_REQUIRE(grad_out && coords && eps3 && grad_grid, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(X >= 2 && Y >= 2 && Z >= 2 && n >= 0, VXN_ERR_SHAPE, "bad shape");
  if (n == 0) return 0;
  const GridDesc g = coords_desc(nullptr, X, Y, Z, eps3);
  hipLaunchKernelGGL(coords_gradient_bwd_kernel, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, (hipStream_t)stream,
                     g, coords, n, grad_out, grad_grid);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_analytical_gradient_fwd(const float* grid, int X, int Y, int Z, const float* coords, int64_t n,
                                           const float* eps3, float* grad, float* normal, void* stream) {
  VXN_REQUIRE(grid && coords && eps3, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(X >= 2 && Y >= 2 && Z >= 2 && n >= 0, VXN_ERR_SHAPE, "bad shape");
  if (n == 0 || (!grad && !normal)) return 0;
  GridDesc g = coords_desc(grid, X, Y, Z, eps3);
  g.analytical = 1;
  hipLaunchKernelGGL(coords_gradient_fwd_kernel, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, (hipStream_t)stream,
                     g, coords, n, grad, normal);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_analytical_gradient_bwd(const float* grad_out, const float* coords, int64_t n, float* grad_grid,
                                           int X, int Y, int Z, const float* eps3, void* stream) {
  VXN_REQUIRE(grad_out && coords && eps3 && grad_grid, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(X >= 2 && Y >= 2 && Z >= 2 && n >= 0, VXN_ERR_SHAPE, "bad shape");
  if (n == 0) return 0;
  GridDesc g = coords_desc(nullptr, X, Y, Z, eps3);
  g.analytical = 1;
  hipLaunchKernelGGL(coords_gradient_bwd_kernel, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, (hipStream_t)stream,
                     g, coords, n, grad_out, grad_grid);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_voxel_corners(const float* coords, int64_t n, int X, int Y, int Z, int64_t* corners,
                                 float* corner_weights, void* stream) {
  VXN_REQUIRE(X >= 1 && Y >= 1 && Z >= 1 && n >= 0, VXN_ERR_SHAPE, "bad shape");
  if (n == 0) return 0;
  VXN_REQUIRE(coords != nullptr, VXN_ERR_NULL, "coords is null");
  const float eps3[3] = {1.0f, 1.0f, 1.0f};
  const GridDesc g = coords_desc(nullptr, X, Y, Z, eps3);
  hipLaunchKernelGGL(voxel_corners_kernel, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, (hipStream_t)stream, g,
                     coords, n, corners, corner_weights);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_lattice_query(const float* grid, int X, int Y, int Z, const float* box6, int NX, int NY, int NZ,
                                 const float* lattice_box6, float* sdf, float* normal, int32_t gradient_mode,
                                 void* stream) {
  if (int e = check_grid(grid, X, Y, Z, box6)) return e;
  if (int e = check_gradient_mode(gradient_mode)) return e;
  VXN_REQUIRE(lattice_box6 && sdf && normal, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(NX >= 2 && NY >= 2 && NZ >= 2, VXN_ERR_SHAPE, "lattice must be >= 2 per axis");
  GridDesc g = make_grid_desc(grid, X, Y, Z, box6);
  g.analytical = gradient_mode;
  const float* lb = lattice_box6;
  VXN_REQUIRE((int64_t)8 * ((NX + 7) / 8) * NY < ((int64_t)1 << 31), VXN_ERR_SHAPE, "lattice too large");
  const unsigned nb = (unsigned)(8 * ((NX + 7) / 8) * NY);
  const float sx = (lb[3] - lb[0]) / (float)(NX - 1), sy = (lb[4] - lb[1]) / (float)(NY - 1),
              sz = (lb[5] - lb[2]) / (float)(NZ - 1);
  hipLaunchKernelGGL(lattice_query_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, g, NX, NY, NZ, lb[0], lb[1],
                     lb[2], sx, sy, sz, sdf, normal);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_upsample_grid(const float* grid, int X, int Y, int Z, float* out, int NX, int NY, int NZ,
                                 void* stream) {
  VXN_REQUIRE(grid != nullptr && out != nullptr, VXN_ERR_NULL, "null grid");
  VXN_REQUIRE(X >= 1 && Y >= 1 && Z >= 1 && NX >= 1 && NY >= 1 && NZ >= 1, VXN_ERR_SHAPE, "sizes must be >= 1");
  VXN_REQUIRE((int64_t)NX * NY < ((int64_t)1 << 31), VXN_ERR_SHAPE, "output too large");
  hipLaunchKernelGGL(upsample_grid_kernel, dim3((unsigned)(NX * NY)), dim3(256), 0, (hipStream_t)stream, grid, out,
                     up_axis(X, NX), up_axis(Y, NY), up_axis(Z, NZ));
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_gaussian_filter_workspace_bytes(int X, int Y, int Z, int64_t* bytes) {
  VXN_REQUIRE(bytes != nullptr, VXN_ERR_NULL, "bytes is null");
  VXN_REQUIRE(X >= 1 && Y >= 1 && Z >= 1, VXN_ERR_SHAPE, "sizes must be >= 1");
  *bytes = (int64_t)X * Y * Z * (int64_t)sizeof(float);
  return 0;
}

extern "C" int vxn_gaussian_filter(const float* grid, int X, int Y, int Z, const double* weights, int radius,
                                   float* out, void* workspace, int64_t workspace_bytes, void* stream) {
  VXN_REQUIRE(grid && weights && out, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(X >= 1 && Y >= 1 && Z >= 1, VXN_ERR_SHAPE, "sizes must be >= 1");
  VXN_REQUIRE(radius >= 0 && radius <= kGaussMaxRadius, VXN_ERR_UNSUPPORTED, "radius must be in [0, %d] (got %d)",
              kGaussMaxRadius, radius);
  const int64_t n = (int64_t)X * Y * Z;
  VXN_REQUIRE(workspace != nullptr && workspace_bytes >= n * 4, VXN_ERR_WORKSPACE,
              "Gaussian filter workspace needs %lld bytes", (long long)(n * 4));
  VXN_REQUIRE(!((const char*)out < (const char*)(grid + n) && (const char*)grid < (const char*)(out + n)),
              VXN_ERR_VALUE, "out must not overlap grid");
  GaussTaps k;
  k.r = radius;
  for (int j = 0; j <= 2 * radius; ++j) k.w[j] = weights[j];
  for (int j = 0; j < radius; ++j)
    VXN_REQUIRE(weights[j] == weights[2 * radius - j], VXN_ERR_VALUE, "the weights must be symmetric");
  float* tmp = (float*)workspace;
  hipStream_t s = (hipStream_t)stream;
  const int nb = blocks_for(n, 256, 65536);
  hipLaunchKernelGGL(gauss1d_kernel, dim3(nb), dim3(256), 0, s, grid, out, X, Y, Z, 0, k);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL(gauss1d_kernel, dim3(nb), dim3(256), 0, s, (const float*)out, tmp, X, Y, Z, 1, k);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL(gauss1d_kernel, dim3(nb), dim3(256), 0, s, (const float*)tmp, out, X, Y, Z, 2, k);
  VXN_LAUNCH_CHECK();
  return 0;
}
