// Work-group reductions over the 128 samples of one ray (two wave64s), shared by the fused
// render kernel and the standalone compositing kernel.
#pragma once

#include "vxn_device.h"

namespace vxn {

struct Red {
  double d[2];
  float f[2][12];
};

// Exclusive prefix product over the work-group's 128 samples, in double.
VXN_DEV double block_excl_prod(double v, Red& red, int lane, int wid) {
  double inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double u = __shfl_up(inc, o, 64);
    if (lane >= o) inc = u * inc;
  }
  if (lane == 63) red.d[wid] = inc;
  __syncthreads();
  const double pre = wid == 0 ? 1.0 : red.d[0];
  const double incl = pre * inc;
  double ex = __shfl_up(incl, 1, 64);
  if (lane == 0) ex = pre;
  __syncthreads();
  return ex;
}

// Inclusive prefix sum over the 128 samples and the total, in double.
VXN_DEV double block_incl_sum(double v, double& total, Red& red, int lane, int wid) {
  double inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double u = __shfl_up(inc, o, 64);
    if (lane >= o) inc = u + inc;
  }
  if (lane == 63) red.d[wid] = inc;
  __syncthreads();
  const double pre = wid == 0 ? 0.0 : red.d[0];
  total = red.d[0] + red.d[1];
  __syncthreads();
  return pre + inc;
}

// Sums of N floats over the 128 samples (returned to every thread).
template <int N>
VXN_DEV void block_sums(float (&v)[N], Red& red, int lane, int wid) {
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) red.f[wid][k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = red.f[0][k] + red.f[1][k];
  __syncthreads();
}

}  // namespace vxn
