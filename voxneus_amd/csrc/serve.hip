// Serving batches: moving request rows between their ranges of a device arena and the
// fixed-capacity batch buffers of one step.  A step is described by a segment table
// (arena row offset, row count, scene) in device memory, so the same launch sequence
// serves every step and can be captured once in a HIP graph.
#include "vxn_host.h"

namespace vxn {

constexpr int kMoveThreads = 256;
constexpr int kMaxSegs = 1024;

// Exclusive prefix sum of the segment row counts into pre[0..ns], by one work-group.
VXN_DEV void segment_prefix(const int32_t* __restrict__ segs, int ns, int64_t* pre) {
  __shared__ int64_t wsum[kMoveThreads / 64];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  int64_t c[4];
  int64_t tot = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int s = 4 * t + k;
    c[k] = s < ns ? (int64_t)max(segs[3 * s + 1], 0) : 0;
    tot += c[k];
  }
  int64_t inc = tot;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int64_t base = 0;
  for (int w = 0; w < wid; ++w) base += wsum[w];
  int64_t run = base + inc - tot;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int s = 4 * t + k;
    if (s <= ns) pre[s] = run;
    run += c[k];
  }
  if (t == kMoveThreads - 1) pre[kMaxSegs] = run;  // the total, for ns == kMaxSegs
  __syncthreads();
}

__global__ void __launch_bounds__(kMoveThreads) serve_move_kernel(const int32_t* __restrict__ segs,
                                                                   const int32_t* __restrict__ n_segs_dev,
                                                                   int max_segs, int64_t cap, int64_t arena_rows,
                                                                   int to_batch, float* arena0, float* batch0, int w0,
                                                                   float* arena1, float* batch1, int w1,
                                                                   int32_t* batch_scene, int32_t* batch_n) {
  __shared__ int64_t pre[kMaxSegs + 1];
  const int ns = min(max(*n_segs_dev, 0), max_segs);
  segment_prefix(segs, ns, pre);
  const int64_t total = min(pre[ns], cap);
  if (to_batch && batch_n && blockIdx.x == 0 && threadIdx.x == 0) *batch_n = (int32_t)total;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = ns - 1;  // last segment with pre[k] <= i
    while (lo < hi) {
      const int m = (lo + hi + 1) >> 1;
      if (pre[m] <= i) lo = m;
      else hi = m - 1;
    }
    const int64_t row = (int64_t)segs[3 * lo] + (i - pre[lo]);
    const bool ok = row >= 0 && row < arena_rows;
    if (to_batch) {
      for (int c = 0; c < w0; ++c) batch0[i * w0 + c] = ok ? arena0[row * w0 + c] : 0.0f;
      for (int c = 0; c < w1; ++c) batch1[i * w1 + c] = ok ? arena1[row * w1 + c] : 0.0f;
      if (batch_scene) batch_scene[i] = ok ? segs[3 * lo + 2] : -1;
    } else if (ok) {
      for (int c = 0; c < w0; ++c) arena0[row * w0 + c] = batch0[i * w0 + c];
      for (int c = 0; c < w1; ++c) arena1[row * w1 + c] = batch1[i * w1 + c];
    }
  }
}

}  // namespace vxn

using namespace vxn;

extern "C" int vxn_serve_move(const int32_t* segs, const int32_t* n_segs, int32_t max_segs, int64_t capacity,
                              int64_t arena_rows, int32_t to_batch, float* arena0, float* batch0, int32_t w0,
                              float* arena1, float* batch1, int32_t w1, int32_t* batch_scene, int32_t* batch_n,
                              void* stream) {
  VXN_REQUIRE(max_segs >= 1 && max_segs <= kMaxSegs, VXN_ERR_VALUE, "max_segs must be in [1, %d]", kMaxSegs);
  VXN_REQUIRE(capacity >= 0 && arena_rows >= 0, VXN_ERR_SHAPE, "negative capacity or arena size");
  VXN_REQUIRE(w0 >= 1 && w1 >= 0 && w0 <= 16 && w1 <= 16, VXN_ERR_VALUE, "row widths must be in [1, 16] / [0, 16]");
  if (capacity == 0) return 0;
  VXN_REQUIRE(segs && n_segs && arena0 && batch0, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(w1 == 0 || (arena1 && batch1), VXN_ERR_NULL, "null second array");
  const int64_t nb = min((capacity + kMoveThreads - 1) / kMoveThreads, (int64_t)1024);
  hipLaunchKernelGGL(serve_move_kernel, dim3((int)nb), dim3(kMoveThreads), 0, (hipStream_t)stream, segs, n_segs,
                     max_segs, capacity, arena_rows, to_batch, arena0, batch0, w0, arena1, batch1, w1, batch_scene,
                     batch_n);
  VXN_LAUNCH_CHECK();
  return 0;
}
