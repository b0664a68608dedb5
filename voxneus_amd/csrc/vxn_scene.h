// One resident scene as the kernels see it: dense SDF grid, hash table, radiance MLP,
// NeuS sharpness s and the occupancy grid of the sampler.  Single-scene launches take
// these fields by value; the serving kernels read them per ray (or per point) from a
// device table of packed SceneDev records.
#pragma once

#include "vxn_hash.h"

namespace vxn {

// OccupancyGrid.is_occupied state (occupancy_grid.py:121-147).
struct OccDesc {
  const uint8_t* occ;  // [res, res, res] bool, nullable (no occupancy modulation)
  int res;
  float mn[3];
  float den[3];  // (occ_max - occ_min) + 1e-8f
  float rden[3]; // RN(1 / den), for div_rn
  float bound;   // float32(res - 1.0001)
};

struct SceneDev {
  GridDesc g;
  HashDesc h;
  OccDesc oc;
  const float* mlp;  // [VXN_MLP_PARAMS] W0 | b0 | W1 | b1
  float s;           // NeuS sharpness
  float sphere_r;    // inside-sphere mask radius of the compositing (0 = none)
};

// Serving batches: the sampler over a packed batch (sampler.hip).
int launch_serve_upsample(const SceneDev* scenes, int n_scenes, const int32_t* ray_scene, const int32_t* n_dev,
                          int64_t cap, const float* o, const float* d, int n0, int per, int steps,
                          const float* t_lin, const float* u_det, float bound_r, float* z, float* dists, float* mid,
                          hipStream_t stream);

inline OccDesc make_occ_desc(const uint8_t* occ, int res, const float* box6) {
  OccDesc o;
  o.occ = occ;
  o.res = res;
  for (int k = 0; k < 3; ++k) {
    o.mn[k] = box6[k];
    o.den[k] = (box6[3 + k] - box6[k]) + 1e-8f;
    o.rden[k] = 1.0f / o.den[k];
  }
  o.bound = (float)((double)res - 1.0001);
  return o;
}

// occ.float() * 0.9 + 0.1 at a point (sampler.py:285-289).
VXN_DEV float occ_factor(const OccDesc& o, float px, float py, float pz) {
  const float p[3] = {px, py, pz};
  bool inb = true;
  int c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float nrm = div_rn(p[k] - o.mn[k], o.den[k], o.rden[k]);
    inb = inb && (nrm >= 0.0f) && (nrm <= 1.0f);
    const float v = fminf(fmaxf(nrm * (float)o.res, 0.0f), o.bound);
    c[k] = (int)v;
  }
  const int64_t r = o.res;
  const bool occ = inb && o.occ[((int64_t)c[0] * r + c[1]) * r + c[2]] != 0;
  const float f = occ ? 1.0f : 0.0f;
  return f * 0.9f + 0.1f;
}

// Rays [r0, r1) of this work-group, visited with stride `step`.  Single-scene launches
// interleave rays over the persistent grid; serving batches give each work-group a
// contiguous range (consecutive rays mostly share a scene) of the device-side count.
template <int MULTI>
VXN_DEV void ray_range(int64_t cap, const int32_t* n_dev, int64_t& r0, int64_t& r1, int64_t& step) {
  if constexpr (MULTI) {
    const int64_t n = min((int64_t)max(*n_dev, 0), cap);
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    r0 = (int64_t)blockIdx.x * per;
    r1 = min(n, r0 + per);
    step = 1;
  } else {
    r0 = blockIdx.x;
    r1 = cap;
    step = gridDim.x;
  }
}

}  // namespace vxn
