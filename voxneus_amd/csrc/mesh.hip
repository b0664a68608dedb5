// Marching cubes over the dense SDF grid (the extraction of geometry/mesh.py:34-77, which
// calls scikit-image's marching_cubes on the host): one vertex per grid edge whose end
// values straddle the level (value < level is inside), at the linear-interpolation
// crossing, shared by the cells around that edge; triangles per cell from a case table.
//
// The case table (mc_table.h, compiled into constant memory) follows a rule on the cube's
// faces, so that the two cells sharing a face always cut it the same way and the mesh has
// no cracks: walking
// each face's corners counter-clockwise about its outward normal, every run of inside
// corners is cut off by one segment from the crossing where the walk leaves the run to
// the crossing where it entered it (diagonal inside corners on a face stay separated).
// Each crossing edge starts one segment in one of its two faces and ends one in the
// other, so the segments close into loops; each loop, oriented so that its normal points
// from the inside corners to the outside ones, is fanned into triangles from its first
// vertex whose fan puts no triangle flat in a cube face (a loop can pass all four
// crossings of a face; a fan across that face would be covered from both sides).  At
// most five triangles per case.
//
// Numbering: corner k = 4 ox + 2 oy + oz; edge 4 a + 2 u + v runs along axis a from the
// corner whose other two coordinates (in ascending axis order) are (u, v).  A grid
// vertex owns the edges leaving it along +x, +y, +z; vertices are numbered by owner in
// grid order, then by axis.  Triangles are numbered by cell in grid order, then in table
// order.
#include <string.h>

#include "mc_table.h"
#include "vxn_host.h"

namespace vxn {

constexpr int kMcThreads = 256;
constexpr int kScanItems = 4;                          // per thread
constexpr int kScanBlock = kMcThreads * kScanItems;    // items per work-group

// The table in constant memory, initialised when the code object loads: no upload, no
// state to track per device.
__constant__ int8_t c_mc_tri[256][16] = VXN_MC_TRI_INIT;  // edge triples, -1 terminated
__constant__ uint8_t c_mc_ntri[256] = VXN_MC_NTRI_INIT;
static const int8_t kMcTri[256][16] = VXN_MC_TRI_INIT;
static const uint8_t kMcNtri[256] = VXN_MC_NTRI_INIT;

// ------------------------------------------------------------------------------------
// Device
// ------------------------------------------------------------------------------------
struct McDev {
  const float* grid;
  int X, Y, Z;
  float level;
  int32_t* vcount;  // [n] crossing edges owned by each grid vertex, then their offsets
  int32_t* tcount;  // [n] triangles of the cell based at each grid vertex, then offsets
  uint8_t* vmask;   // [n] owned crossing edges, bit a for axis a
  int32_t* bsum;    // [2 * (nb + 1)] work-group sums, then their offsets; totals at [nb]
  int32_t nb;
};

VXN_DEV bool mc_inside(float v, float level) { return v < level; }

__global__ void __launch_bounds__(kMcThreads) mc_classify_kernel(McDev m) {
  const int64_t n = (int64_t)m.X * m.Y * m.Z, YZ = (int64_t)m.Y * m.Z;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i / YZ), y = (int)((i / m.Z) % m.Y), z = (int)(i % m.Z);
    const bool in0 = mc_inside(m.grid[i], m.level);
    unsigned vm = 0;
    if (x + 1 < m.X && mc_inside(m.grid[i + YZ], m.level) != in0) vm |= 1u;
    if (y + 1 < m.Y && mc_inside(m.grid[i + m.Z], m.level) != in0) vm |= 2u;
    if (z + 1 < m.Z && mc_inside(m.grid[i + 1], m.level) != in0) vm |= 4u;
    m.vmask[i] = (uint8_t)vm;
    m.vcount[i] = __popc(vm);
    int nt = 0;
    if (x + 1 < m.X && y + 1 < m.Y && z + 1 < m.Z) {
      unsigned cs = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int64_t j = i + ((k >> 2) & 1) * YZ + ((k >> 1) & 1) * (int64_t)m.Z + (k & 1);
        if (mc_inside(m.grid[j], m.level)) cs |= 1u << k;
      }
      nt = c_mc_ntri[cs];
    }
    m.tcount[i] = nt;
  }
}

// Exclusive scan of a[n] in place, work-group by work-group (kScanBlock items each);
// each work-group's total goes to bsum[b].
VXN_DEV void scan_block(int32_t* a, int64_t n, int32_t* bsum) {
  __shared__ int32_t wsum[kMcThreads / 64];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int64_t base = (int64_t)blockIdx.x * kScanBlock + (int64_t)t * kScanItems;
  int32_t v[kScanItems];
  int32_t tot = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? a[base + k] : 0;
    tot += v[k];
  }
  int32_t inc = tot;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int32_t off = 0;
  for (int w = 0; w < wid; ++w) off += wsum[w];
  int32_t run = off + inc - tot;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) a[base + k] = run;
    run += v[k];
  }
  if (t == kMcThreads - 1) bsum[blockIdx.x] = off + inc;
}

__global__ void __launch_bounds__(kMcThreads) mc_scan1_kernel(McDev m) {
  const int64_t n = (int64_t)m.X * m.Y * m.Z;
  scan_block(m.vcount, n, m.bsum);
  __syncthreads();
  scan_block(m.tcount, n, m.bsum + m.nb + 1);
}

// One work-group: exclusive scan of the nb work-group sums of both arrays; the totals
// land at bsum[nb] and bsum[2 nb + 1].
__global__ void __launch_bounds__(1024) mc_scan2_kernel(McDev m, int32_t* totals) {
  __shared__ int32_t wsum[1024 / 64];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int per = (m.nb + 1023) / 1024;
  for (int arr = 0; arr < 2; ++arr) {
    int32_t* s = m.bsum + arr * (m.nb + 1);
    const int b0 = t * per;
    int32_t tot = 0;
    for (int k = 0; k < per; ++k)
      if (b0 + k < m.nb) tot += s[b0 + k];
    int32_t inc = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int32_t off = 0;
    for (int w = 0; w < wid; ++w) off += wsum[w];
    int32_t run = off + inc - tot;
    for (int k = 0; k < per; ++k) {
      if (b0 + k < m.nb) {
        const int32_t x = s[b0 + k];
        s[b0 + k] = run;
        run += x;
      }
    }
    if (t == 1023) {
      s[m.nb] = off + inc;
      if (totals) totals[arr] = off + inc;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kMcThreads) mc_scan3_kernel(McDev m) {
  const int64_t n = (int64_t)m.X * m.Y * m.Z;
  const int32_t ov = m.bsum[blockIdx.x], ot = m.bsum[m.nb + 1 + blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kScanBlock;
  for (int k = threadIdx.x; k < kScanBlock; k += kMcThreads) {
    const int64_t i = base + k;
    if (i < n) {
      m.vcount[i] += ov;
      m.tcount[i] += ot;
    }
  }
}

__global__ void __launch_bounds__(kMcThreads) mc_vertices_kernel(McDev m, double mn0, double mn1, double mn2,
                                                                 double sc0, double sc1, double sc2,
                                                                 float* __restrict__ verts, int64_t max_verts) {
  const int64_t n = (int64_t)m.X * m.Y * m.Z, YZ = (int64_t)m.Y * m.Z;
  const double mn[3] = {mn0, mn1, mn2}, sc[3] = {sc0, sc1, sc2};
  const int64_t step[3] = {YZ, (int64_t)m.Z, 1};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned vm = m.vmask[i];
    if (!vm) continue;
    const int64_t p[3] = {i / YZ, (i / m.Z) % m.Y, i % m.Z};
    const double v0 = m.grid[i];
    int64_t o = m.vcount[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (!((vm >> a) & 1u)) continue;
      const double v1 = m.grid[i + step[a]];
      const double t = ((double)m.level - v0) / (v1 - v0);
      if (o < max_verts) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const double q = (double)p[b] + (b == a ? t : 0.0);
          verts[3 * o + b] = (float)(q * sc[b] + mn[b]);
        }
      }
      ++o;
    }
  }
}

__global__ void __launch_bounds__(kMcThreads) mc_faces_kernel(McDev m, int32_t* __restrict__ faces,
                                                              int64_t max_faces) {
  const int64_t n = (int64_t)m.X * m.Y * m.Z, YZ = (int64_t)m.Y * m.Z;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i / YZ), y = (int)((i / m.Z) % m.Y), z = (int)(i % m.Z);
    if (x + 1 >= m.X || y + 1 >= m.Y || z + 1 >= m.Z) continue;
    unsigned cs = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int64_t j = i + ((k >> 2) & 1) * YZ + ((k >> 1) & 1) * (int64_t)m.Z + (k & 1);
      if (mc_inside(m.grid[j], m.level)) cs |= 1u << k;
    }
    const int nt = c_mc_ntri[cs];
    int64_t f = m.tcount[i];
    for (int q = 0; q < 3 * nt; ++q) {
      const int e = c_mc_tri[cs][q];
      const int a = e >> 2, u = (e >> 1) & 1, v = e & 1;
      // owner vertex: the edge's low corner; the two bits are the other axes in order
      int off[3] = {0, 0, 0}, k = 0;
      const int uv[2] = {u, v};
#pragma unroll
      for (int b = 0; b < 3; ++b)
        if (b != a) off[b] = uv[k++];
      const int64_t owner = i + off[0] * YZ + off[1] * (int64_t)m.Z + off[2];
      const unsigned vm = m.vmask[owner];
      const int32_t id = m.vcount[owner] + __popc(vm & ((1u << a) - 1u));
      const int64_t slot = 3 * f + q;
      if (f + q / 3 < max_faces) faces[slot] = id;
    }
  }
}

struct McLayout {
  int64_t n, nb, off_vc, off_tc, off_vm, off_bs, total;
};

static McLayout mc_layout(int X, int Y, int Z) {
  McLayout l{};
  l.n = (int64_t)X * Y * Z;
  l.nb = (l.n + kScanBlock - 1) / kScanBlock;
  auto al = [](int64_t b) { return (b + 255) & ~(int64_t)255; };
  l.off_vc = 0;
  l.off_tc = al(l.off_vc + 4 * l.n);
  l.off_vm = al(l.off_tc + 4 * l.n);
  l.off_bs = al(l.off_vm + l.n);
  l.total = al(l.off_bs + 4 * 2 * (l.nb + 1));
  return l;
}

static int mc_dev(const float* grid, int X, int Y, int Z, float level, void* ws, int64_t ws_bytes, McDev& m) {
  VXN_REQUIRE(grid != nullptr, VXN_ERR_NULL, "grid is null");
  VXN_REQUIRE(X >= 2 && Y >= 2 && Z >= 2, VXN_ERR_SHAPE, "grid must be >= 2 per axis");
  const McLayout l = mc_layout(X, Y, Z);
  VXN_REQUIRE(l.n < ((int64_t)1 << 31) / 3, VXN_ERR_SHAPE, "grid too large for 32-bit vertex ids");
  VXN_REQUIRE(l.nb <= 1024 * 1024, VXN_ERR_SHAPE, "grid too large");
  VXN_REQUIRE(ws != nullptr && ws_bytes >= l.total, VXN_ERR_WORKSPACE, "marching cubes needs %lld workspace bytes",
              (long long)l.total);
  char* w = (char*)ws;
  m.grid = grid;
  m.X = X;
  m.Y = Y;
  m.Z = Z;
  m.level = level;
  m.vcount = (int32_t*)(w + l.off_vc);
  m.tcount = (int32_t*)(w + l.off_tc);
  m.vmask = (uint8_t*)(w + l.off_vm);
  m.bsum = (int32_t*)(w + l.off_bs);
  m.nb = (int32_t)l.nb;
  return 0;
}

}  // namespace vxn

using namespace vxn;

extern "C" int vxn_mc_case_table(int8_t* tri, uint8_t* ntri) {
  VXN_REQUIRE(tri != nullptr && ntri != nullptr, VXN_ERR_NULL, "null output");
  memcpy(tri, kMcTri, sizeof(kMcTri));
  memcpy(ntri, kMcNtri, sizeof(kMcNtri));
  return 0;
}

extern "C" int vxn_mc_workspace_bytes(int X, int Y, int Z, int64_t* bytes) {
  VXN_REQUIRE(bytes != nullptr, VXN_ERR_NULL, "bytes is null");
  VXN_REQUIRE(X >= 2 && Y >= 2 && Z >= 2, VXN_ERR_SHAPE, "grid must be >= 2 per axis");
  *bytes = mc_layout(X, Y, Z).total;
  return 0;
}

extern "C" int vxn_mc_count(const float* grid, int X, int Y, int Z, float level, void* workspace,
                            int64_t workspace_bytes, int32_t* counts, void* stream) {
  McDev m{};
  if (int e = mc_dev(grid, X, Y, Z, level, workspace, workspace_bytes, m)) return e;
  VXN_REQUIRE(counts != nullptr, VXN_ERR_NULL, "counts is null");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = (int64_t)X * Y * Z;
  hipLaunchKernelGGL(mc_classify_kernel, dim3(blocks_for(n, kMcThreads, 65536)), dim3(kMcThreads), 0, s, m);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL(mc_scan1_kernel, dim3(m.nb), dim3(kMcThreads), 0, s, m);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL(mc_scan2_kernel, dim3(1), dim3(1024), 0, s, m, counts);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL(mc_scan3_kernel, dim3(m.nb), dim3(kMcThreads), 0, s, m);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_mc_emit(const float* grid, int X, int Y, int Z, const float* box6, float level, void* workspace,
                           int64_t workspace_bytes, float* verts, int64_t max_verts, int32_t* faces,
                           int64_t max_faces, void* stream) {
  McDev m{};
  if (int e = mc_dev(grid, X, Y, Z, level, workspace, workspace_bytes, m)) return e;
  VXN_REQUIRE(box6 != nullptr, VXN_ERR_NULL, "box is null");
  VXN_REQUIRE(max_verts >= 0 && max_faces >= 0, VXN_ERR_SHAPE, "negative capacity");
  VXN_REQUIRE((verts || max_verts == 0) && (faces || max_faces == 0), VXN_ERR_NULL, "null output");
  // world = index * (xyz_max - xyz_min) / (res - 1) + xyz_min, the extent in float32
  // and the rest in float64 (mesh.py:71-73)
  const int res[3] = {X, Y, Z};
  double mn[3], sc[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = (double)box6[a];
    sc[a] = (double)(box6[3 + a] - box6[a]) / ((double)res[a] - 1.0);
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = (int64_t)X * Y * Z;
  const int nb = blocks_for(n, kMcThreads, 65536);
  if (max_verts > 0) {
    hipLaunchKernelGGL(mc_vertices_kernel, dim3(nb), dim3(kMcThreads), 0, s, m, mn[0], mn[1], mn[2], sc[0], sc[1],
                       sc[2], verts, max_verts);
    VXN_LAUNCH_CHECK();
  }
  if (max_faces > 0) {
    hipLaunchKernelGGL(mc_faces_kernel, dim3(nb), dim3(kMcThreads), 0, s, m, faces, max_faces);
    VXN_LAUNCH_CHECK();
  }
  return 0;
}
