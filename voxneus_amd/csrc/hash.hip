// Multiresolution hash encoding (forward / backward) as standalone kernels.  The fused
// render kernels use the same device functions (vxn_hash.h).
#include "vxn_hash.h"
#include "vxn_host.h"

namespace vxn {

__global__ void __launch_bounds__(256) hash_fwd_kernel(HashDesc h, const float* __restrict__ x01, int64_t n,
                                                       float* __restrict__ feats) {
  const int L = h.L;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float x = fminf(fmaxf(x01[3 * i], 0.0f), 1.0f);
    const float y = fminf(fmaxf(x01[3 * i + 1], 0.0f), 1.0f);
    const float z = fminf(fmaxf(x01[3 * i + 2], 0.0f), 1.0f);
    for (int l = 0; l < L; ++l) {
      float f0, f1;
      hash_level_fwd<true>(h, l, x, y, z, f0, f1);
      feats[i * 2 * L + 2 * l] = f0;
      feats[i * 2 * L + 2 * l + 1] = f1;
    }
  }
}

__global__ void __launch_bounds__(256) hash_bwd_kernel(HashDesc h, const float* __restrict__ x01, int64_t n,
                                                       const float* __restrict__ dfeats,
                                                       float* __restrict__ grad_table) {
  const int L = h.L;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float x = fminf(fmaxf(x01[3 * i], 0.0f), 1.0f);
    const float y = fminf(fmaxf(x01[3 * i + 1], 0.0f), 1.0f);
    const float z = fminf(fmaxf(x01[3 * i + 2], 0.0f), 1.0f);
    for (int l = 0; l < L; ++l)
      hash_level_bwd(h, l, x, y, z, dfeats[i * 2 * L + 2 * l], dfeats[i * 2 * L + 2 * l + 1], grad_table);
  }
}

int make_hash_desc(HashDesc& h, const void* table, int dtype, int L, int log2T, int F, const int32_t* level_res,
                   const float* level_w) {
  VXN_REQUIRE(level_res != nullptr, VXN_ERR_NULL, "level_res is null");
  VXN_REQUIRE(L >= 1 && L <= VXN_MAX_LEVELS, VXN_ERR_SHAPE, "n_levels must be in [1, %d]", VXN_MAX_LEVELS);
  VXN_REQUIRE(log2T >= 1 && log2T <= 24, VXN_ERR_SHAPE, "log2 table size must be in [1, 24]");
  VXN_REQUIRE(F == 2, VXN_ERR_UNSUPPORTED, "only 2 features per level are implemented (got %d)", F);
  VXN_REQUIRE(dtype == VXN_DTYPE_F32 || dtype == VXN_DTYPE_BF16, VXN_ERR_VALUE, "unknown table dtype %d", dtype);
  // entries are read as one 8-byte (FP32 pair) or 4-byte (BF16 pair) load, and the BF16
  // table's gather loads 16 bytes at once where four entries sit together
  VXN_REQUIRE(((uintptr_t)table & (dtype == VXN_DTYPE_BF16 ? 15u : 7u)) == 0, VXN_ERR_VALUE,
              "table must be %d-byte aligned", dtype == VXN_DTYPE_BF16 ? 16 : 8);
  h.table = table;
  h.dtype = dtype;
  h.L = L;
  h.T = (int64_t)1 << log2T;
  h.mask = (uint32_t)(h.T - 1);
  for (int l = 0; l < VXN_MAX_LEVELS; ++l) {
    h.res[l] = l < L ? (float)level_res[l] : 0.0f;
    h.lw[l] = (l < L && level_w) ? level_w[l] : 1.0f;
  }
  for (int l = 0; l < L; ++l) {
    VXN_REQUIRE(level_res[l] >= 1 && level_res[l] < (1 << 24), VXN_ERR_VALUE, "level %d resolution %d out of range", l,
                level_res[l]);
    VXN_REQUIRE(h.lw[l] >= 0.0f && h.lw[l] <= 1.0f, VXN_ERR_VALUE, "level weight must be in [0,1]");
  }
  return 0;
}

}  // namespace vxn

using namespace vxn;

extern "C" int vxn_hash_fwd(const void* table, int dtype, int L, int log2T, int F, const int32_t* level_res,
                            const float* level_w, const float* x01, int64_t n, float* feats, void* stream) {
  HashDesc h;
  if (int e = make_hash_desc(h, table, dtype, L, log2T, F, level_res, level_w)) return e;
  VXN_REQUIRE(n >= 0, VXN_ERR_SHAPE, "n must be >= 0");
  if (n == 0) return 0;
  VXN_REQUIRE(table && x01 && feats, VXN_ERR_NULL, "null argument");
  hipLaunchKernelGGL(hash_fwd_kernel, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, (hipStream_t)stream, h, x01, n,
                     feats);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_hash_bwd(int L, int log2T, int F, const int32_t* level_res, const float* level_w,
                            const float* x01, int64_t n, const float* d_feats, float* grad_table, void* stream) {
  HashDesc h;
  if (int e = make_hash_desc(h, nullptr, VXN_DTYPE_F32, L, log2T, F, level_res, level_w)) return e;
  VXN_REQUIRE(n >= 0, VXN_ERR_SHAPE, "n must be >= 0");
  if (n == 0) return 0;
  VXN_REQUIRE(x01 && d_feats && grad_table, VXN_ERR_NULL, "null argument");
  hipLaunchKernelGGL(hash_bwd_kernel, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, (hipStream_t)stream, h, x01, n,
                     d_feats, grad_table);
  VXN_LAUNCH_CHECK();
  return 0;
}
