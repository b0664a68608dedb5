// OccupancyGrid.update_from_sdf (rendering/occupancy_grid.py:62-119): the EMA update of
// the sampler's occupancy grid from the NeuS2 logistic density of the SDF grid at cell
// centres.  Every update reads all the cells it visits before writing any of them (the
// reference gathers, computes and scatters as separate tensor ops), so cells drawn
// twice in one update get the same value: an indexed update computes into a workspace
// first and scatters in a second kernel.  A full sweep (every cell once) runs in place.
#include <cmath>

#include "vxn_host.h"

namespace vxn {

struct OccUpdate {
  float* density;  // [res^3]
  uint8_t* mask;   // [res^3] bool
  int res;
  float mn[3];
  float ext[3];     // xyz_max - xyz_min
  float s;          // NeuS sharpness
  float decay;      // 0.95
  float thr;        // occupancy threshold (0.01)
  float near_val;   // threshold * 2 (the near-surface density floor)
  float near_dist;  // 2 * |voxel step| (the near-surface distance)
  int first;        // total_updates == 0: no decay term
};

VXN_DEV float max_nan(float a, float b) { return (a != a || b != b) ? __int_as_float(0x7fc00000) : fmaxf(a, b); }

// The updated density of cell (x, y, z), reading its previous value.
VXN_DEV float occ_cell_update(const GridDesc& g, const OccUpdate& u, int64_t x, int64_t y, int64_t z) {
  const float r = (float)u.res;
  const float nx = ((float)x + 0.5f) / r, ny = ((float)y + 0.5f) / r, nz = ((float)z + 0.5f) / r;
  const float px = u.mn[0] + nx * u.ext[0], py = u.mn[1] + ny * u.ext[1], pz = u.mn[2] + nz * u.ext[2];
  const float sdf = gs_eval(g, gs_stencil(g, px, py, pz));
  // logistic_density (occupancy_grid.py:15-22): s * sigmoid(s d) * (1 - sigmoid(s d))
  const float sig = sigmoid_ref(u.s * sdf);
  const float dens = (u.s * sig) * (1.0f - sig);
  const float near = fabsf(sdf) < u.near_dist ? 1.0f : 0.0f;
  const float eff = max_nan(dens, near * u.near_val);
  if (u.first) return eff;
  const int64_t R = u.res;
  const float prev = u.density[(x * R + y) * R + z];
  return max_nan(prev * u.decay, eff);
}

__global__ void __launch_bounds__(256) occ_sweep_kernel(GridDesc g, OccUpdate u) {
  const int64_t R = u.res, n = R * R * R;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = i / (R * R), y = (i / R) % R, z = i % R;
    const float v = occ_cell_update(g, u, x, y, z);
    u.density[i] = v;
    u.mask[i] = v > u.thr ? 1 : 0;
  }
}

// Cells outside [0, res) are skipped (the reference's indexing would raise).
VXN_DEV bool occ_cell(const OccUpdate& u, const int64_t* ix, const int64_t* iy, const int64_t* iz, int64_t i,
                      int64_t& x, int64_t& y, int64_t& z) {
  x = ix[i];
  y = iy[i];
  z = iz[i];
  return x >= 0 && y >= 0 && z >= 0 && x < u.res && y < u.res && z < u.res;
}

__global__ void __launch_bounds__(256) occ_gather_kernel(GridDesc g, OccUpdate u, const int64_t* __restrict__ ix,
                                                         const int64_t* __restrict__ iy,
                                                         const int64_t* __restrict__ iz, int64_t n,
                                                         float* __restrict__ upd) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t x, y, z;
    if (occ_cell(u, ix, iy, iz, i, x, y, z)) upd[i] = occ_cell_update(g, u, x, y, z);
  }
}

__global__ void __launch_bounds__(256) occ_scatter_kernel(OccUpdate u, const int64_t* __restrict__ ix,
                                                          const int64_t* __restrict__ iy,
                                                          const int64_t* __restrict__ iz, int64_t n,
                                                          const float* __restrict__ upd) {
  const int64_t R = u.res;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t x, y, z;
    if (!occ_cell(u, ix, iy, iz, i, x, y, z)) continue;
    const int64_t c = (x * R + y) * R + z;
    const float v = upd[i];
    u.density[c] = v;
    u.mask[c] = v > u.thr ? 1 : 0;
  }
}

}  // namespace vxn

using namespace vxn;

extern "C" int vxn_occupancy_update_workspace_bytes(int64_t n_cells, int64_t* bytes) {
  VXN_REQUIRE(bytes != nullptr, VXN_ERR_NULL, "bytes is null");
  VXN_REQUIRE(n_cells >= 0, VXN_ERR_SHAPE, "n_cells must be >= 0");
  *bytes = n_cells * (int64_t)sizeof(float);
  return 0;
}

extern "C" int vxn_occupancy_update(float* density, uint8_t* mask, int32_t res, const float* occ_box6,
                                    const float* grid, int32_t X, int32_t Y, int32_t Z, const float* box6, float s,
                                    float decay, float threshold, int32_t first_update, const int64_t* idx_x,
                                    const int64_t* idx_y, const int64_t* idx_z, int64_t n_cells, void* workspace,
                                    int64_t workspace_bytes, void* stream) {
  if (int e = check_grid(grid, X, Y, Z, box6)) return e;
  VXN_REQUIRE(density && mask && occ_box6, VXN_ERR_NULL, "null density, mask or box");
  VXN_REQUIRE(res >= 1 && res <= 2048, VXN_ERR_SHAPE, "res must be in [1, 2048]");
  const bool indexed = idx_x || idx_y || idx_z;
  VXN_REQUIRE(!indexed || (idx_x && idx_y && idx_z), VXN_ERR_NULL, "give all three index arrays or none");
  VXN_REQUIRE(n_cells >= 0, VXN_ERR_SHAPE, "n_cells must be >= 0");
  OccUpdate u{};
  u.density = density;
  u.mask = mask;
  u.res = res;
  float step[3];
  for (int a = 0; a < 3; ++a) {
    u.mn[a] = occ_box6[a];
    u.ext[a] = occ_box6[3 + a] - occ_box6[a];
    step[a] = u.ext[a] / (float)res;
  }
  // voxel_diag = |step| (a Python float), compared with |d| as float32 after doubling
  const double diag = (double)(float)std::sqrt((double)step[0] * step[0] + (double)step[1] * step[1] +
                                               (double)step[2] * step[2]);
  u.near_dist = (float)(2.0 * diag);
  u.near_val = (float)((double)threshold * 2.0);
  u.s = s;
  u.decay = decay;
  u.thr = threshold;
  u.first = first_update ? 1 : 0;
  const GridDesc g = make_grid_desc(grid, X, Y, Z, box6);
  hipStream_t st = (hipStream_t)stream;
  if (!indexed) {
    const int64_t n = (int64_t)res * res * res;
    hipLaunchKernelGGL(occ_sweep_kernel, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, st, g, u);
    VXN_LAUNCH_CHECK();
    return 0;
  }
  if (n_cells == 0) return 0;
  VXN_REQUIRE(workspace != nullptr && workspace_bytes >= n_cells * (int64_t)sizeof(float), VXN_ERR_WORKSPACE,
              "an indexed update needs %lld workspace bytes", (long long)(n_cells * (int64_t)sizeof(float)));
  float* upd = (float*)workspace;
  const int nb = blocks_for(n_cells, 256, 65536);
  hipLaunchKernelGGL(occ_gather_kernel, dim3(nb), dim3(256), 0, st, g, u, idx_x, idx_y, idx_z, n_cells, upd);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL(occ_scatter_kernel, dim3(nb), dim3(256), 0, st, u, idx_x, idx_y, idx_z, n_cells,
                     (const float*)upd);
  VXN_LAUNCH_CHECK();
  return 0;
}
