// Host-side helpers shared by the C-ABI entry points: argument validation, error codes
// and the thread-local error message returned by vxn_last_error().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <string>

#include "../../include/voxneus_hip.h"
#include "vxn_device.h"

namespace vxn {

void set_error(const char* fmt, ...);

#define VXN_REQUIRE(cond, code, ...)   \
  do {                                 \
    if (!(cond)) {                     \
      ::vxn::set_error(__VA_ARGS__);   \
      return (code);                   \
    }                                  \
  } while (0)

#define VXN_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::vxn::set_error("%s failed: %s", #call, hipGetErrorString(e_));             \
      return VXN_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

#define VXN_LAUNCH_CHECK()                                                         \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess) {                                                        \
      ::vxn::set_error("kernel launch failed: %s", hipGetErrorString(e_));         \
      return VXN_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

// Build a GridDesc with the reference's float32 constants (see vxn_device.h).
inline GridDesc make_grid_desc(const float* grid, int X, int Y, int Z, const float* box6) {
  GridDesc g;
  g.grid = grid;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  const int res[3] = {X, Y, Z};
  for (int a = 0; a < 3; ++a) {
    const float mn = box6[a], mx = box6[3 + a];
    g.mn[a] = mn;
    g.den[a] = (mx - mn) + 1e-8f;
    g.rden[a] = 1.0f / g.den[a];
    g.rm1[a] = (float)res[a] - 1.0f;
    g.bound[a] = (float)((double)res[a] - 1.0001);
    g.eps[a] = (mx - mn) / ((float)res[a] - 1.0f);
  }
  g.analytical = 0;
  return g;
}

inline int check_gradient_mode(int mode) {
  VXN_REQUIRE(mode == VXN_GRADIENT_INTERPOLATED || mode == VXN_GRADIENT_ANALYTICAL, VXN_ERR_VALUE,
              "gradient_mode must be %d (interpolated) or %d (analytical), got %d", VXN_GRADIENT_INTERPOLATED,
              VXN_GRADIENT_ANALYTICAL, mode);
  return 0;
}

inline int check_grid(const float* grid, int X, int Y, int Z, const float* box6) {
  VXN_REQUIRE(grid != nullptr, VXN_ERR_NULL, "grid pointer is null");
  VXN_REQUIRE(box6 != nullptr, VXN_ERR_NULL, "box pointer is null");
  VXN_REQUIRE(X >= 2 && Y >= 2 && Z >= 2, VXN_ERR_SHAPE, "grid resolution must be >= 2 per axis (got %d,%d,%d)", X,
              Y, Z);
  VXN_REQUIRE((int64_t)X * Y * Z < (int64_t)1 << 31, VXN_ERR_SHAPE, "grid has %lld vertices; the limit is 2^31",
              (long long)X * Y * Z);
  for (int a = 0; a < 3; ++a)
    VXN_REQUIRE(box6[3 + a] > box6[a], VXN_ERR_VALUE, "xyz_max must exceed xyz_min on axis %d", a);
  return 0;
}

inline int blocks_for(int64_t n, int threads, int cap = 1 << 20) {
  int64_t b = (n + threads - 1) / threads;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace vxn
