// The radiance MLP of one point at a time, for VoxNeuS.forward_points (the per-point
// pipeline the mesh colouring calls, geometry/mesh.py:291-364; models/voxneus.py:115-167)
// outside the fused render: rgb = sigmoid(W1 relu(W0 [hash32, view, normal3] + b0) + b1)
// (models/radiance_mlp.py:81-99) with its hand-written backward.
//
// point_mlp_fwd_kernel: one thread per point, the parameters staged in LDS (every lane of
// a wavefront reads the same weight: LDS broadcast), the hidden units walked one at a time
// so only the 35 + NV inputs and 3 output sums stay in registers.
// point_mlp_bwd_kernel: per point the input gradients, and relu(h), dL/dh and dL/dout kept
// for the weight gradient; point_mlp_param_kernel: one thread per parameter sums its
// product over a 2048-point chunk in double; point_mlp_reduce_kernel sums the chunks in a
// fixed order (deterministic).
//
// BF = true is the BF16 form's arithmetic (the rounding points of the fused render's
// matrix-core MLP): hash features, normal, W0's hash / normal columns, relu(h) and W1
// rounded to bf16, the view part b0 + W0_view . view in FP32; in the backward dL/dout
// and dL/dh rounded to bf16, and the view columns of dW0 taken with bf16 view values.
#include "vxn_host.h"
#include "vxn_mfma.h"

namespace vxn {

constexpr int kPmThreads = 256;
constexpr int kPmHidden = 64;
constexpr int kPmChunk = 2048;  // points per chunk of the parameter gradient

struct PointMlpDev {
  const float* feats;    // [n][32]
  const float* view;     // [n][NV]
  const float* normals;  // [n][3]
  const float* mlp;      // W0 [64][35 + NV], b0 [64], W1 [3][64], b1 [3]
  int64_t n;
  float* rgb;            // [n][3]
  const float* d_rgb;    // [n][3]
  float* d_feats;        // [n][32] (nullable)
  float* d_view;         // [n][NV] (nullable)
  float* d_normals;      // [n][3]  (nullable)
  float* hid;            // [n][64] relu(h) as the weight gradient reads it
  float* dh;             // [n][64] dL/dh
  float* dout;           // [n][3]  dL/d(pre-sigmoid output)
  float* part;           // [nchunk][np]
  float* d_mlp;          // [np]
};

template <int NV>
struct PmLayout {
  static constexpr int WS = 35 + NV;  // W0 row stride
  static constexpr int B0 = 64 * WS;
  static constexpr int W1 = B0 + 64;
  static constexpr int B1 = W1 + 192;
  static constexpr int NP = B1 + 3;
};

VXN_DEV float maybe_bf16(float x, bool bf) { return bf ? bf16_round(x) : x; }

// Loads the parameters into LDS (BF16 form: W0's hash / normal columns and W1 rounded).
template <int NV, bool BF>
VXN_DEV void stage_params(const float* __restrict__ mlp, float* w) {
  using ML = PmLayout<NV>;
  for (int k = threadIdx.x; k < ML::NP; k += blockDim.x) {
    float v = mlp[k];
    if (BF) {
      if (k < ML::B0) {
        const int col = k % ML::WS;
        if (col < 32 || col >= 32 + NV) v = bf16_round(v);
      } else if (k >= ML::W1 && k < ML::B1) {
        v = bf16_round(v);
      }
    }
    w[k] = v;
  }
}

// The point's inputs in W0 column order: x[0..31] hash features, x[32..32+NV) view,
// x[32+NV..35+NV) normal.
template <int NV, bool BF>
VXN_DEV void load_inputs(const PointMlpDev& a, int64_t i, float (&x)[35 + NV]) {
#pragma unroll
  for (int k = 0; k < 32; ++k) x[k] = maybe_bf16(a.feats[i * 32 + k], BF);
#pragma unroll
  for (int k = 0; k < NV; ++k) x[32 + k] = a.view[i * NV + k];
#pragma unroll
  for (int k = 0; k < 3; ++k) x[32 + NV + k] = maybe_bf16(a.normals[i * 3 + k], BF);
}

// Pre-activation of hidden unit j: b0 + view part first, then the hash and normal inputs.
template <int NV>
VXN_DEV float hidden_pre(const float* w, int j, const float (&x)[35 + NV]) {
  using ML = PmLayout<NV>;
  const float* row = w + j * ML::WS;
  float acc = w[ML::B0 + j];
#pragma unroll
  for (int k = 32; k < 32 + NV; ++k) acc = fmaf(row[k], x[k], acc);
#pragma unroll
  for (int k = 0; k < 32; ++k) acc = fmaf(row[k], x[k], acc);
#pragma unroll
  for (int k = 32 + NV; k < 35 + NV; ++k) acc = fmaf(row[k], x[k], acc);
  return acc;
}

template <int NV, bool BF>
__global__ void __launch_bounds__(kPmThreads) point_mlp_fwd_kernel(PointMlpDev a) {
  using ML = PmLayout<NV>;
  extern __shared__ float w[];
  stage_params<NV, BF>(a.mlp, w);
  __syncthreads();
  const int64_t i = blockIdx.x * (int64_t)kPmThreads + threadIdx.x;
  if (i >= a.n) return;
  float x[35 + NV];
  load_inputs<NV, BF>(a, i, x);
  float o[3] = {w[ML::B1], w[ML::B1 + 1], w[ML::B1 + 2]};
  for (int j = 0; j < kPmHidden; ++j) {
    const float h = maybe_bf16(fmaxf(hidden_pre<NV>(w, j, x), 0.0f), BF);
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = fmaf(w[ML::W1 + c * 64 + j], h, o[c]);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) a.rgb[i * 3 + c] = sigmoid_ref(o[c]);
}

template <int NV, bool BF>
__global__ void __launch_bounds__(kPmThreads) point_mlp_bwd_kernel(PointMlpDev a) {
  using ML = PmLayout<NV>;
  extern __shared__ float w[];
  stage_params<NV, BF>(a.mlp, w);
  __syncthreads();
  const int64_t i = blockIdx.x * (int64_t)kPmThreads + threadIdx.x;
  if (i >= a.n) return;
  float x[35 + NV];
  load_inputs<NV, BF>(a, i, x);
  // forward again (the activations are not stored by the forward)
  float o[3] = {w[ML::B1], w[ML::B1 + 1], w[ML::B1 + 2]};
  for (int j = 0; j < kPmHidden; ++j) {
    const float h = maybe_bf16(fmaxf(hidden_pre<NV>(w, j, x), 0.0f), BF);
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = fmaf(w[ML::W1 + c * 64 + j], h, o[c]);
  }
  float dout[3], doutw[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float r = sigmoid_ref(o[c]);
    dout[c] = (a.d_rgb[i * 3 + c] * (1.0f - r)) * r;
    doutw[c] = maybe_bf16(dout[c], BF);
    a.dout[i * 3 + c] = dout[c];
  }
  float dx[35 + NV];
#pragma unroll
  for (int k = 0; k < 35 + NV; ++k) dx[k] = 0.0f;
  for (int j = 0; j < kPmHidden; ++j) {
    const float pre = hidden_pre<NV>(w, j, x);
    const float h = maybe_bf16(fmaxf(pre, 0.0f), BF);
    float g = (w[ML::W1 + j] * doutw[0] + w[ML::W1 + 64 + j] * doutw[1]) + w[ML::W1 + 128 + j] * doutw[2];
    g = pre > 0.0f ? maybe_bf16(g, BF) : 0.0f;
    a.hid[i * kPmHidden + j] = h;
    a.dh[i * kPmHidden + j] = g;
    const float* row = w + j * ML::WS;
#pragma unroll
    for (int k = 0; k < 35 + NV; ++k) dx[k] = fmaf(row[k], g, dx[k]);
  }
  if (a.d_feats) {
#pragma unroll
    for (int k = 0; k < 32; ++k) a.d_feats[i * 32 + k] = dx[k];
  }
  if (a.d_view) {
#pragma unroll
    for (int k = 0; k < NV; ++k) a.d_view[i * NV + k] = dx[32 + k];
  }
  if (a.d_normals) {
#pragma unroll
    for (int k = 0; k < 3; ++k) a.d_normals[i * 3 + k] = dx[32 + NV + k];
  }
}

// One thread per parameter p, one chunk of points per blockIdx.y: the sum over the chunk
// of p's product (in double): dW0[j][k] = dh_j x_k, db0[j] = dh_j, dW1[c][j] = dout_c
// relu(h_j) (bf16 dout in the BF16 form), db1[c] = dout_c.
template <int NV, bool BF>
__global__ void __launch_bounds__(kPmThreads) point_mlp_param_kernel(PointMlpDev a) {
  using ML = PmLayout<NV>;
  const int p = blockIdx.x * kPmThreads + threadIdx.x;
  if (p >= ML::NP) return;
  const int64_t i0 = (int64_t)blockIdx.y * kPmChunk;
  const int64_t i1 = min(a.n, i0 + kPmChunk);
  double acc = 0.0;
  if (p < ML::B0) {
    const int j = p / ML::WS, k = p - j * ML::WS;
    for (int64_t i = i0; i < i1; ++i) {
      float xk;
      if (k < 32) xk = maybe_bf16(a.feats[i * 32 + k], BF);
      else if (k < 32 + NV) xk = maybe_bf16(a.view[i * NV + k - 32], BF);
      else xk = maybe_bf16(a.normals[i * 3 + k - 32 - NV], BF);
      acc += (double)a.dh[i * kPmHidden + j] * (double)xk;
    }
  } else if (p < ML::W1) {
    const int j = p - ML::B0;
    for (int64_t i = i0; i < i1; ++i) acc += (double)a.dh[i * kPmHidden + j];
  } else if (p < ML::B1) {
    const int c = (p - ML::W1) / 64, j = (p - ML::W1) - c * 64;
    for (int64_t i = i0; i < i1; ++i)
      acc += (double)maybe_bf16(a.dout[i * 3 + c], BF) * (double)a.hid[i * kPmHidden + j];
  } else {
    const int c = p - ML::B1;
    for (int64_t i = i0; i < i1; ++i) acc += (double)a.dout[i * 3 + c];
  }
  a.part[(int64_t)blockIdx.y * ML::NP + p] = (float)acc;
}

template <int NV>
__global__ void __launch_bounds__(kPmThreads) point_mlp_reduce_kernel(PointMlpDev a, int nchunk) {
  using ML = PmLayout<NV>;
  const int p = blockIdx.x * kPmThreads + threadIdx.x;
  if (p >= ML::NP) return;
  double acc = 0.0;
  for (int c = 0; c < nchunk; ++c) acc += (double)a.part[(int64_t)c * ML::NP + p];
  a.d_mlp[p] = (float)acc;
}

static int pm_view_dim(int view_dim) {
  VXN_REQUIRE(view_dim == 16 || view_dim == 27, VXN_ERR_UNSUPPORTED,
              "the per-point MLP takes the 16 SH or 27 positional-encoding view values (got %d)", view_dim);
  return 0;
}

static int64_t pm_nchunk(int64_t n) { return (n + kPmChunk - 1) / kPmChunk; }

template <int NV, bool BF>
static int pm_launch_fwd(PointMlpDev& a, hipStream_t s) {
  const size_t lds = (size_t)PmLayout<NV>::NP * sizeof(float);
  hipLaunchKernelGGL((point_mlp_fwd_kernel<NV, BF>), dim3((unsigned)((a.n + kPmThreads - 1) / kPmThreads)),
                     dim3(kPmThreads), lds, s, a);
  VXN_LAUNCH_CHECK();
  return 0;
}

template <int NV, bool BF>
static int pm_launch_bwd(PointMlpDev& a, hipStream_t s) {
  using ML = PmLayout<NV>;
  const size_t lds = (size_t)ML::NP * sizeof(float);
  const int64_t nchunk = pm_nchunk(a.n);
  hipLaunchKernelGGL((point_mlp_bwd_kernel<NV, BF>), dim3((unsigned)((a.n + kPmThreads - 1) / kPmThreads)),
                     dim3(kPmThreads), lds, s, a);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL((point_mlp_param_kernel<NV, BF>), dim3((ML::NP + kPmThreads - 1) / kPmThreads, (unsigned)nchunk),
                     dim3(kPmThreads), 0, s, a);
  VXN_LAUNCH_CHECK();
  hipLaunchKernelGGL((point_mlp_reduce_kernel<NV>), dim3((ML::NP + kPmThreads - 1) / kPmThreads), dim3(kPmThreads), 0,
                     s, a, (int)nchunk);
  VXN_LAUNCH_CHECK();
  return 0;
}

}  // namespace vxn

using namespace vxn;

extern "C" int vxn_point_mlp_fwd(const float* feats, const float* view, int view_dim, const float* normals,
                                 const float* mlp, int64_t n, int mlp_mode, float* rgb, void* stream) {
  if (int e = pm_view_dim(view_dim)) return e;
  VXN_REQUIRE(n >= 0, VXN_ERR_SHAPE, "n must be >= 0");
  VXN_REQUIRE(mlp_mode == VXN_MLP_FP32 || mlp_mode == VXN_MLP_BF16_MFMA, VXN_ERR_VALUE, "unknown mlp_mode %d",
              mlp_mode);
  if (n == 0) return 0;
  VXN_REQUIRE(feats && view && normals && mlp && rgb, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(n < ((int64_t)1 << 31) / kPmThreads * kPmThreads, VXN_ERR_SHAPE, "too many points");
  PointMlpDev a = {};
  a.feats = feats;
  a.view = view;
  a.normals = normals;
  a.mlp = mlp;
  a.n = n;
  a.rgb = rgb;
  hipStream_t s = (hipStream_t)stream;
  const bool bf = mlp_mode == VXN_MLP_BF16_MFMA;
  if (view_dim == 16) return bf ? pm_launch_fwd<16, true>(a, s) : pm_launch_fwd<16, false>(a, s);
  return bf ? pm_launch_fwd<27, true>(a, s) : pm_launch_fwd<27, false>(a, s);
}

extern "C" int vxn_point_mlp_bwd_workspace_bytes(int64_t n, int view_dim, int64_t* bytes) {
  VXN_REQUIRE(bytes != nullptr, VXN_ERR_NULL, "bytes is null");
  if (int e = pm_view_dim(view_dim)) return e;
  VXN_REQUIRE(n >= 0, VXN_ERR_SHAPE, "n must be >= 0");
  const int np = view_dim == 16 ? PmLayout<16>::NP : PmLayout<27>::NP;
  *bytes = n * (int64_t)(2 * kPmHidden + 3) * 4 + pm_nchunk(n) * np * 4;
  return 0;
}

extern "C" int vxn_point_mlp_bwd(const float* feats, const float* view, int view_dim, const float* normals,
                                 const float* mlp, int64_t n, int mlp_mode, const float* d_rgb, float* d_feats,
                                 float* d_view, float* d_normals, float* d_mlp, void* workspace,
                                 int64_t workspace_bytes, void* stream) {
  if (int e = pm_view_dim(view_dim)) return e;
  VXN_REQUIRE(n >= 0, VXN_ERR_SHAPE, "n must be >= 0");
  VXN_REQUIRE(mlp_mode == VXN_MLP_FP32 || mlp_mode == VXN_MLP_BF16_MFMA, VXN_ERR_VALUE, "unknown mlp_mode %d",
              mlp_mode);
  VXN_REQUIRE(d_mlp != nullptr, VXN_ERR_NULL, "d_mlp is null");
  const int np = view_dim == 16 ? PmLayout<16>::NP : PmLayout<27>::NP;
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {  // no points: a zero gradient (the reduction over no chunks)
    PointMlpDev z = {};
    z.d_mlp = d_mlp;
    if (view_dim == 16)
      hipLaunchKernelGGL((point_mlp_reduce_kernel<16>), dim3((np + kPmThreads - 1) / kPmThreads), dim3(kPmThreads),
                         0, s, z, 0);
    else
      hipLaunchKernelGGL((point_mlp_reduce_kernel<27>), dim3((np + kPmThreads - 1) / kPmThreads), dim3(kPmThreads),
                         0, s, z, 0);
    VXN_LAUNCH_CHECK();
    return 0;
  }
  VXN_REQUIRE(feats && view && normals && mlp && d_rgb, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(n < ((int64_t)1 << 31) / kPmThreads * kPmThreads, VXN_ERR_SHAPE, "too many points");
  int64_t need = 0;
  vxn_point_mlp_bwd_workspace_bytes(n, view_dim, &need);
  VXN_REQUIRE(workspace != nullptr && workspace_bytes >= need, VXN_ERR_WORKSPACE,
              "point MLP workspace needs %lld bytes", (long long)need);
  PointMlpDev a = {};
  a.feats = feats;
  a.view = view;
  a.normals = normals;
  a.mlp = mlp;
  a.n = n;
  a.d_rgb = d_rgb;
  a.d_feats = d_feats;
  a.d_view = d_view;
  a.d_normals = d_normals;
  a.hid = (float*)workspace;
  a.dh = a.hid + n * kPmHidden;
  a.dout = a.dh + n * kPmHidden;
  a.part = a.dout + n * 3;
  a.d_mlp = d_mlp;
  const bool bf = mlp_mode == VXN_MLP_BF16_MFMA;
  if (view_dim == 16) return bf ? pm_launch_bwd<16, true>(a, s) : pm_launch_bwd<16, false>(a, s);
  return bf ? pm_launch_bwd<27, true>(a, s) : pm_launch_bwd<27, false>(a, s);
}
