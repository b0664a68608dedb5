// Optimizer updates, on the device so that a training iteration needs no host sync:
//  * the SDF grid's sparse Adam (SparseGridAdam with the native backend: the dense and
//    masked Adam kernels of cuda/adam_upd_kernel.cu:6-82, and the masked PyTorch
//    formula of training/optimizer.py:92-113 that it takes with weight decay);
//  * the network's Adam (torch.optim.Adam, betas (0.9, 0.999), eps 1e-8, one parameter
//    group each for the hash table, the MLP and the variance, training/optimizer.py:143-150)
//    as one multi-tensor launch, with the variance group's learning rate chosen on the
//    device from the s-amplification flag (trainer.py:329-338).
// Every kernel streams float4 vectors: the updates are HBM-bound elementwise work.
#include <cmath>

#include "vxn_host.h"
#include "vxn_mfma.h"

namespace vxn {

constexpr int kAdamThreads = 256;

VXN_DEV void load4(const float* p, int64_t i, int64_t n, float o[4]) {
  if (i + 4 <= n) {
    const float4 v = *reinterpret_cast<const float4*>(p + i);
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) o[e] = i + e < n ? p[i + e] : 0.0f;
  }
}

VXN_DEV void store4(float* p, int64_t i, int64_t n, const float o[4]) {
  if (i + 4 <= n) {
    *reinterpret_cast<float4*>(p + i) = make_float4(o[0], o[1], o[2], o[3]);
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (i + e < n) p[i + e] = o[e];
  }
}

VXN_DEV void count_add(int32_t* counter, int c) {
  // one atomic per wavefront
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0 && c != 0) atomicAdd(counter, c);
}

// Float constants of one update, rounded as the reference's code rounds them.
struct AdamConsts {
  float b1, omb1;  // beta1 and 1 - beta1
  float b2, omb2;
  float step_size, eps, wd, bc2_sqrt;
};

// FORM 0: the native kernel (adam_upd_kernel.cu:17-19, masked :32-35):
//   m = b1 m + (1 - b1) g;  v = b2 v + (1 - b2) g g;  p -= step_size m / (sqrt(v) + eps)
//   with step_size = lr sqrt(1 - b2^t) / (1 - b1^t) from the host.
// FORM 1: the masked PyTorch formula with weight decay (optimizer.py:95-113):
//   g' = g + wd p;  m = m b1 + (1 - b1) g';  v = v b2 + (1 - b2) (g' g');
//   p -= step_size (m / (sqrt(v) / sqrt(1 - b2^t) + eps))  with step_size = lr / (1 - b1^t).
// MASKED: only elements with g != 0 change (and are counted).
template <int FORM, bool MASKED>
VXN_DEV void grid_adam_update(float pv[4], float mv[4], float vv[4], const float gv[4], int64_t i, int64_t n,
                              const AdamConsts& c, int& cnt) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (MASKED && gv[e] == 0.0f) continue;
    if (i + e >= n) continue;
    ++cnt;
    if constexpr (FORM == 0) {
      const float gr = gv[e];
      mv[e] = c.b1 * mv[e] + c.omb1 * gr;
      vv[e] = c.b2 * vv[e] + c.omb2 * gr * gr;
      pv[e] = pv[e] - c.step_size * mv[e] / (sqrtf(vv[e]) + c.eps);
    } else {
      const float gr = c.wd != 0.0f ? gv[e] + c.wd * pv[e] : gv[e];
      mv[e] = mv[e] * c.b1 + c.omb1 * gr;
      vv[e] = vv[e] * c.b2 + c.omb2 * (gr * gr);
      const float denom = sqrtf(vv[e]) / c.bc2_sqrt + c.eps;
      pv[e] = pv[e] - c.step_size * (mv[e] / denom);
    }
  }
}

// Each thread takes kGroups float4 groups a work-group stride apart: their gradient loads
// are all issued before any moment is read, so the dependent loads of the masked form
// overlap across groups instead of serialising.
constexpr int kGroups = 4;

template <int FORM, bool MASKED>
__global__ void __launch_bounds__(kAdamThreads) grid_adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                                                  float* __restrict__ m, float* __restrict__ v,
                                                                  int64_t n, AdamConsts c,
                                                                  int32_t* __restrict__ n_updated) {
  int cnt = 0;
  const int64_t tile = (int64_t)kAdamThreads * 4 * kGroups;
  for (int64_t base = (int64_t)blockIdx.x * tile; base < n; base += (int64_t)gridDim.x * tile) {
    float gv[kGroups][4];
    bool act[kGroups];
#pragma unroll
    for (int q = 0; q < kGroups; ++q) {
      const int64_t i = base + ((int64_t)q * kAdamThreads + threadIdx.x) * 4;
      act[q] = i < n;
      if (act[q]) load4(g, i, n, gv[q]);
      if constexpr (MASKED) {
        bool any = false;
#pragma unroll
        for (int e = 0; e < 4; ++e) any = any || (act[q] && gv[q][e] != 0.0f);
        act[q] = any;
      }
    }
    float pv[kGroups][4], mv[kGroups][4], vv[kGroups][4];
#pragma unroll
    for (int q = 0; q < kGroups; ++q) {
      if (!act[q]) continue;
      const int64_t i = base + ((int64_t)q * kAdamThreads + threadIdx.x) * 4;
      load4(p, i, n, pv[q]);
      load4(m, i, n, mv[q]);
      load4(v, i, n, vv[q]);
    }
#pragma unroll
    for (int q = 0; q < kGroups; ++q) {
      if (!act[q]) continue;
      const int64_t i = base + ((int64_t)q * kAdamThreads + threadIdx.x) * 4;
      grid_adam_update<FORM, MASKED>(pv[q], mv[q], vv[q], gv[q], i, n, c, cnt);
      store4(p, i, n, pv[q]);
      store4(m, i, n, mv[q]);
      store4(v, i, n, vv[q]);
    }
  }
  if (n_updated) count_add(n_updated, cnt);
}

// torch.optim.Adam over several parameters in one launch (its multi-tensor CUDA order):
//   m = lerp(m, g, 1 - b1) = fma(1 - b1, g - m, m);  v = v b2 + (1 - b2) (g g);
//   p = p + (-lr / (1 - b1^t)) (m / (sqrt(v) / sqrt(1 - b2^t) + eps)).
struct AdamTensorDev {
  float* p;
  const float* g;
  float* m;
  float* v;
  int64_t n;
  float neg_step[2];  // -lr / bc1, without and with the flag's learning-rate scale
  const int32_t* flag;
  uint16_t* p_bf16;  // nullable bf16 copy of the updated parameter
  float p_min;       // clamp_min after the update (-inf: none)
  int64_t block0;    // first work-group of this tensor
};

constexpr int kMaxAdamTensors = 8;

struct AdamMultiDev {
  AdamTensorDev t[kMaxAdamTensors];
  int nt;
  float w1;  // 1 - b1 as the lerp weight
  float b2, omb2, bc2_sqrt, eps;
};

__global__ void __launch_bounds__(kAdamThreads) torch_adam_kernel(AdamMultiDev a) {
  int k = 0;
#pragma unroll 1
  for (int j = 1; j < a.nt; ++j)
    if ((int64_t)blockIdx.x >= a.t[j].block0) k = j;
  const AdamTensorDev& T = a.t[k];
  const int64_t i = (((int64_t)blockIdx.x - T.block0) * blockDim.x + threadIdx.x) * 4;
  if (i >= T.n) return;
  const float neg_step = T.flag ? T.neg_step[*T.flag != 0 ? 1 : 0] : T.neg_step[0];
  float pv[4], gv[4], mv[4], vv[4];
  load4(T.p, i, T.n, pv);
  load4(T.g, i, T.n, gv);
  load4(T.m, i, T.n, mv);
  load4(T.v, i, T.n, vv);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    mv[e] = fmaf(a.w1, gv[e] - mv[e], mv[e]);  // lerp, weight < 0.5: one rounding, as PyTorch
    vv[e] = vv[e] * a.b2;
    vv[e] = vv[e] + a.omb2 * (gv[e] * gv[e]);
    const float denom = sqrtf(vv[e]) / a.bc2_sqrt + a.eps;
    pv[e] = pv[e] + neg_step * (mv[e] / denom);
    pv[e] = pv[e] < T.p_min ? T.p_min : pv[e];  // clamp_min_, NaN stays NaN
  }
  store4(T.p, i, T.n, pv);
  store4(T.m, i, T.n, mv);
  store4(T.v, i, T.n, vv);
  if (T.p_bf16) {
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (i + e < T.n) T.p_bf16[i + e] = bf16_bits(pv[e]);
  }
}

// step_size of the native kernels' host code (adam_upd_kernel.cu:50, :70), in float.
static float native_step_size(int step, float b1, float b2, float lr) {
  return lr * sqrtf(1.0f - powf(b2, (float)step)) / (1.0f - powf(b1, (float)step));
}

static int grid_adam_launch(int form, bool masked, float* param, const float* grad, float* exp_avg,
                            float* exp_avg_sq, int64_t n, const AdamConsts& c, int32_t* n_updated, hipStream_t s) {
  const int nb = blocks_for((n + 4 * kGroups - 1) / (4 * kGroups), kAdamThreads, 4096);
  if (form == 0 && masked)
    hipLaunchKernelGGL((grid_adam_kernel<0, true>), dim3(nb), dim3(kAdamThreads), 0, s, param, grad, exp_avg,
                       exp_avg_sq, n, c, n_updated);
  else if (form == 0)
    hipLaunchKernelGGL((grid_adam_kernel<0, false>), dim3(nb), dim3(kAdamThreads), 0, s, param, grad, exp_avg,
                       exp_avg_sq, n, c, n_updated);
  else
    hipLaunchKernelGGL((grid_adam_kernel<1, true>), dim3(nb), dim3(kAdamThreads), 0, s, param, grad, exp_avg,
                       exp_avg_sq, n, c, n_updated);
  VXN_LAUNCH_CHECK();
  return 0;
}

// The native kernels take float betas, lr and eps (pybind11 converts the Python floats)
// and form 1 - beta in float.
static AdamConsts native_consts(int step, float b1, float b2, float lr, float eps) {
  AdamConsts c{};
  c.b1 = b1;
  c.omb1 = 1.0f - b1;
  c.b2 = b2;
  c.omb2 = 1.0f - b2;
  c.step_size = native_step_size(step, b1, b2, lr);
  c.eps = eps;
  c.wd = 0.0f;
  c.bc2_sqrt = 1.0f;
  return c;
}

static int check_adam(const float* param, const float* grad, const float* exp_avg, const float* exp_avg_sq,
                      int64_t n, int64_t step, double b1, double b2) {
  VXN_REQUIRE(n >= 0, VXN_ERR_SHAPE, "n must be >= 0");
  VXN_REQUIRE(n == 0 || (param && grad && exp_avg && exp_avg_sq), VXN_ERR_NULL, "null tensor");
  VXN_REQUIRE(step >= 1, VXN_ERR_VALUE, "step must be >= 1 (the count after this update)");
  VXN_REQUIRE(b1 >= 0.0 && b1 < 1.0 && b2 >= 0.0 && b2 < 1.0, VXN_ERR_VALUE, "betas must be in [0, 1)");
  const uintptr_t al = (uintptr_t)param | (uintptr_t)grad | (uintptr_t)exp_avg | (uintptr_t)exp_avg_sq;
  VXN_REQUIRE((al & 15u) == 0, VXN_ERR_VALUE, "tensors must be 16-byte aligned");
  return 0;
}

}  // namespace vxn

using namespace vxn;

extern "C" int vxn_adam_upd(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                            int32_t step, float beta1, float beta2, float lr, float eps, void* stream) {
  if (int e = check_adam(param, grad, exp_avg, exp_avg_sq, n, step, beta1, beta2)) return e;
  if (n == 0) return 0;
  return grid_adam_launch(0, false, param, grad, exp_avg, exp_avg_sq, n, native_consts(step, beta1, beta2, lr, eps),
                          nullptr, (hipStream_t)stream);
}

extern "C" int vxn_masked_adam_upd(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                                   int32_t step, float beta1, float beta2, float lr, float eps, int32_t* n_updated,
                                   void* stream) {
  if (int e = check_adam(param, grad, exp_avg, exp_avg_sq, n, step, beta1, beta2)) return e;
  if (n == 0) return 0;
  return grid_adam_launch(0, true, param, grad, exp_avg, exp_avg_sq, n, native_consts(step, beta1, beta2, lr, eps),
                          n_updated, (hipStream_t)stream);
}

extern "C" int vxn_sparse_adam_decay(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                                     int64_t step, double beta1, double beta2, double lr, double eps,
                                     double weight_decay, int32_t* n_updated, void* stream) {
  if (int e = check_adam(param, grad, exp_avg, exp_avg_sq, n, step, beta1, beta2)) return e;
  if (n == 0) return 0;
  // Python floats: scalars and bias corrections in double, each rounded to float where it
  // meets a float32 tensor
  const double bc1 = 1.0 - std::pow(beta1, (double)step);
  const double bc2 = 1.0 - std::pow(beta2, (double)step);
  AdamConsts c{};
  c.b1 = (float)beta1;
  c.omb1 = (float)(1.0 - beta1);
  c.b2 = (float)beta2;
  c.omb2 = (float)(1.0 - beta2);
  c.step_size = (float)(lr / bc1);
  c.eps = (float)eps;
  c.wd = (float)weight_decay;
  c.bc2_sqrt = (float)std::sqrt(bc2);
  return grid_adam_launch(1, true, param, grad, exp_avg, exp_avg_sq, n, c, n_updated, (hipStream_t)stream);
}

extern "C" int vxn_adam_step(const vxn_adam_tensor* tensors, int32_t n_tensors, int64_t step, double beta1,
                             double beta2, double eps, void* stream) {
  VXN_REQUIRE(tensors != nullptr, VXN_ERR_NULL, "tensors is null");
  VXN_REQUIRE(n_tensors >= 1 && n_tensors <= kMaxAdamTensors, VXN_ERR_VALUE, "n_tensors must be in [1, %d]",
              kMaxAdamTensors);
  VXN_REQUIRE(step >= 1, VXN_ERR_VALUE, "step must be >= 1 (the count after this update)");
  VXN_REQUIRE(beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0, VXN_ERR_VALUE,
              "betas must be in [0, 1)");
  // torch.optim.Adam's host scalars: Python floats (double), cast to float at the kernels
  const double b1 = beta1, b2 = beta2;
  const double bc1 = 1.0 - std::pow(b1, (double)step);
  const double bc2 = 1.0 - std::pow(b2, (double)step);
  AdamMultiDev a{};
  a.nt = n_tensors;
  a.w1 = (float)(1.0 - b1);
  a.b2 = (float)b2;
  a.omb2 = (float)(1.0 - b2);
  a.bc2_sqrt = (float)std::sqrt(bc2);
  a.eps = (float)eps;
  int64_t blocks = 0;
  for (int j = 0; j < n_tensors; ++j) {
    const vxn_adam_tensor& s = tensors[j];
    VXN_REQUIRE(s.n >= 0, VXN_ERR_SHAPE, "tensor %d: n must be >= 0", j);
    VXN_REQUIRE(s.n == 0 || (s.param && s.grad && s.exp_avg && s.exp_avg_sq), VXN_ERR_NULL, "tensor %d: null", j);
    const uintptr_t al = (uintptr_t)s.param | (uintptr_t)s.grad | (uintptr_t)s.exp_avg | (uintptr_t)s.exp_avg_sq;
    VXN_REQUIRE((al & 15u) == 0, VXN_ERR_VALUE, "tensor %d: must be 16-byte aligned", j);
    VXN_REQUIRE(((uintptr_t)s.param_bf16 & 1u) == 0, VXN_ERR_VALUE, "tensor %d: bf16 copy misaligned", j);
    AdamTensorDev& t = a.t[j];
    t.p = s.param;
    t.g = s.grad;
    t.m = s.exp_avg;
    t.v = s.exp_avg_sq;
    t.n = s.n;
    t.neg_step[0] = (float)((s.lr / bc1) * -1.0);
    t.neg_step[1] = (float)(((s.lr * s.lr_flag_scale) / bc1) * -1.0);
    t.flag = s.lr_flag;
    t.p_bf16 = (uint16_t*)s.param_bf16;
    t.p_min = s.param_min;
    t.block0 = blocks;
    blocks += (s.n + 4 * kAdamThreads - 1) / (4 * kAdamThreads);
  }
  if (blocks == 0) return 0;
  VXN_REQUIRE(blocks < ((int64_t)1 << 31), VXN_ERR_SHAPE, "tensors too large");
  hipLaunchKernelGGL(torch_adam_kernel, dim3((unsigned)blocks), dim3(kAdamThreads), 0, (hipStream_t)stream, a);
  VXN_LAUNCH_CHECK();
  return 0;
}
