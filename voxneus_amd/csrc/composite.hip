// NeuS alpha and front-to-back compositing of per-sample values that the caller already
// holds: volume_render_from_samples (rendering/volume_render.py:57-153), composite_weights
// (:12-40) and compute_neus_alpha (rendering/neus_alpha.py:12-81), each with its
// hand-written backward.  The fused render kernel does the same per-ray work after its
// own per-sample pipeline; this file takes the samples as inputs.
//
// composite_kernel: one ray per 128-thread work-group (thread t = sample t, two wave64s),
// in the fused render kernel's operation order: Eq. 2 alpha from the SDF, the normal's
// cosine with the ray (or the iter_cos annealing) and s, the unit-sphere mask, the
// transmittance as an exclusive FP64 prefix product of clamp(1 - alpha + 1e-7, 0, 1), the
// T < 1e-4 cutoff, and the FP32 composites.  The backward takes suffix sums (FP64) of
// dL/dw * w and dL/dT * T for the product's terms and writes dL/d(colour, sdf, normal) per
// sample and the per-ray part of dL/ds, which composite_s_finish_kernel sums in a fixed
// order.  With alpha_in the alpha is an input (composite_weights) and its gradient is
// returned instead.
// neus_alpha_kernel: the same alpha, one thread per sample, for compute_neus_alpha.
#include "vxn_block.h"
#include "vxn_host.h"

namespace vxn {

constexpr int kCt = 128;  // samples per ray (one per thread)

struct CompositeDev {
  const float* colors;
  const float* sdf;
  const float* normals;
  const float* dirs;
  int dirs_per_sample;
  const float* z;
  const float* dists;
  const float* s;
  const float* pts;
  const float* alpha_in;
  float sphere_r, cos_anneal;
  int iter_cos;
  int64_t R;
  int S;
  float *rgb, *depth, *normal, *opacity, *weights, *alpha, *trans;
  const float *d_rgb, *d_depth, *d_normal, *d_opacity, *d_weights, *d_alpha, *d_trans;
  float *d_colors, *d_sdf, *d_normals, *d_alpha_in;
  double* part_s;  // [R] per-ray dL/ds
  float* d_s;
};

// Eq. 2 alpha of one sample (neus_alpha.py:40-76) with the quantities its backward needs.
struct AlphaTerms {
  float tcos, dist, fp, fn, pp, pn, den, q, araw;
};

VXN_DEV AlphaTerms neus_alpha(float sdf, const float n[3], const float d[3], float dist, float sv, int iter_cos,
                              float cos_anneal) {
  AlphaTerms t;
  t.tcos = (n[0] * d[0] + n[1] * d[1]) + n[2] * d[2];
  float cosv = t.tcos;
  if (iter_cos) {
    const float ra = fmaxf(-t.tcos * 0.5f + 0.5f, 0.0f), rb = fmaxf(-t.tcos, 0.0f);
    cosv = -(ra * (1.0f - cos_anneal) + rb * cos_anneal);
  }
  t.dist = dist;
  const float off = 0.5f * dist * cosv;
  t.fp = sdf - off;
  t.fn = sdf + off;
  t.pp = sigmoid_ref(sv * t.fp);
  t.pn = sigmoid_ref(sv * t.fn);
  t.den = fmaxf(t.pp, 1e-5f);
  t.q = (t.pp - t.pn) / t.den;
  t.araw = fminf(fmaxf(t.q, 0.0f), 1.0f);
  return t;
}

// Backward of neus_alpha for dL/dalpha_raw = dA: dL/dsdf, dL/dcos (to the normal through
// the ray direction) and this sample's share of dL/ds.
VXN_DEV void neus_alpha_bwd(const AlphaTerms& t, float dA, float sv, int iter_cos, float cos_anneal, float& dsdf,
                            float& dcos, double& ds) {
  const float dq = (t.q >= 0.0f && t.q <= 1.0f) ? dA : 0.0f;
  float dpp = dq / t.den;
  const float dpn = -dq / t.den;
  if (t.pp >= 1e-5f) dpp -= dq * t.q / t.den;
  const float dzp = dpp * (1.0f - t.pp) * t.pp;
  const float dzn = dpn * (1.0f - t.pn) * t.pn;
  ds = (double)dzp * t.fp + (double)dzn * t.fn;
  const float dfp = dzp * sv, dfn = dzn * sv;
  dsdf = dfp + dfn;
  const float doff = dfn - dfp;
  dcos = doff * 0.5f * t.dist;
  if (iter_cos) {
    const float ra = -t.tcos * 0.5f + 0.5f, rb = -t.tcos;
    dcos = dcos * (-((ra > 0.0f ? -0.5f : 0.0f) * (1.0f - cos_anneal) + (rb > 0.0f ? -1.0f : 0.0f) * cos_anneal));
  }
}

template <bool BWD, bool AIN>
__global__ void __launch_bounds__(kCt) composite_kernel(CompositeDev a) {
  __shared__ Red red;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int64_t ray = blockIdx.x;
  const bool valid = t < a.S;
  const int64_t si = ray * a.S + t;
  const float sv = AIN ? 0.0f : a.s[0];
  float c[3] = {0.0f, 0.0f, 0.0f}, n[3] = {0.0f, 0.0f, 0.0f}, d[3] = {0.0f, 0.0f, 0.0f};
  float sdf = 0.0f, z = 0.0f, inside = 1.0f, alpha = 0.0f;
  AlphaTerms at = {};
  if (valid) {
    if (a.colors)
      for (int k = 0; k < 3; ++k) c[k] = a.colors[3 * si + k];
    if (a.z) z = a.z[si];
    if (a.normals)
      for (int k = 0; k < 3; ++k) n[k] = a.normals[3 * si + k];
    if (AIN) {
      alpha = a.alpha_in[si];
    } else {
      sdf = a.sdf[si];
      const float* dp = a.dirs + 3 * (a.dirs_per_sample ? si : ray);
      for (int k = 0; k < 3; ++k) d[k] = dp[k];
      at = neus_alpha(sdf, n, d, a.dists[si], sv, a.iter_cos, a.cos_anneal);
      if (a.pts && a.sphere_r > 0.0f) {
        const float px = a.pts[3 * si], py = a.pts[3 * si + 1], pz = a.pts[3 * si + 2];
        inside = sqrtf(px * px + py * py + pz * pz) < a.sphere_r ? 1.0f : 0.0f;
      }
      alpha = at.araw * inside;
    }
  }
  const float ov = (1.0f - alpha) + 1e-7f;
  const float oma = valid ? fminf(fmaxf(ov, 0.0f), 1.0f) : 1.0f;
  const float T = (float)block_excl_prod((double)oma, red, lane, wid);
  const bool keep = !(T < 1e-4f);
  const float w = (valid && keep) ? alpha * T : 0.0f;
  float comp[8] = {w * c[0], w * c[1], w * c[2], w * z, w, w * n[0], w * n[1], w * n[2]};
  block_sums<8>(comp, red, lane, wid);
  const float Nr = sqrtf(comp[5] * comp[5] + comp[6] * comp[6] + comp[7] * comp[7]);
  if (!BWD) {
    if (t == 0) {
      if (a.rgb)
        for (int k = 0; k < 3; ++k) a.rgb[3 * ray + k] = comp[k];
      if (a.depth) a.depth[ray] = comp[3];
      if (a.opacity) a.opacity[ray] = comp[4];
      if (a.normal)
        for (int k = 0; k < 3; ++k) a.normal[3 * ray + k] = comp[5 + k] / (Nr + 1e-8f);
    }
    if (valid) {
      if (a.weights) a.weights[si] = w;
      if (a.alpha) a.alpha[si] = alpha;
      if (a.trans) a.trans[si] = T;
    }
    return;
  }
  // ------------------------------------------------------------------ backward
  float dr[3] = {0.0f, 0.0f, 0.0f}, dd = 0.0f, dop = 0.0f, dN[3] = {0.0f, 0.0f, 0.0f};
  if (a.d_rgb)
    for (int k = 0; k < 3; ++k) dr[k] = a.d_rgb[3 * ray + k];
  if (a.d_depth) dd = a.d_depth[ray];
  if (a.d_opacity) dop = a.d_opacity[ray];
  if (a.d_normal) {
    // normal = N / (|N| + 1e-8)
    const float gn[3] = {a.d_normal[3 * ray], a.d_normal[3 * ray + 1], a.d_normal[3 * ray + 2]};
    const float rr = Nr + 1e-8f;
    const float dotv = comp[5] * gn[0] + comp[6] * gn[1] + comp[7] * gn[2];
    for (int k = 0; k < 3; ++k) dN[k] = gn[k] / rr - (Nr > 0.0f ? comp[5 + k] * dotv / (Nr * rr * rr) : 0.0f);
  }
  float gw = (dr[0] * c[0] + dr[1] * c[1]) + dr[2] * c[2] + dd * z + dop + ((dN[0] * n[0] + dN[1] * n[1]) + dN[2] * n[2]);
  if (a.d_weights && valid) gw += a.d_weights[si];
  if (!valid || !keep) gw = 0.0f;
  const float gT = (a.d_trans && valid) ? a.d_trans[si] : 0.0f;
  // d/d(oma_i) of sum_k (g_k w_k + gT_k T_k) = suffix_i / oma_i over k > i
  double tot;
  const double qi = (double)gw * (double)alpha * (double)T + (double)gT * (double)T;
  const double pre = block_incl_sum(qi, tot, red, lane, wid);
  const double suffix = tot - pre;
  float dA = gw * T;
  if (a.d_alpha && valid) dA += a.d_alpha[si];
  if (ov <= 1.0f && ov >= 0.0f && valid) dA -= (float)(suffix / (double)oma);
  double ds = 0.0;
  if (valid) {
    for (int k = 0; k < 3; ++k)
      if (a.d_colors) a.d_colors[3 * si + k] = w * dr[k];
    if (AIN) {
      if (a.d_alpha_in) a.d_alpha_in[si] = dA;
      if (a.d_normals)
        for (int k = 0; k < 3; ++k) a.d_normals[3 * si + k] = w * dN[k];
    } else {
      float dsdf, dcos;
      neus_alpha_bwd(at, dA * inside, sv, a.iter_cos, a.cos_anneal, dsdf, dcos, ds);
      if (a.d_sdf) a.d_sdf[si] = dsdf;
      if (a.d_normals)
        for (int k = 0; k < 3; ++k) a.d_normals[3 * si + k] = dcos * d[k] + w * dN[k];
    }
  }
  if (!AIN && a.part_s) {
    const double dsw = wave_sum_d(ds);
    if (lane == 0) red.d[wid] = dsw;
    __syncthreads();
    if (t == 0) a.part_s[ray] = red.d[0] + red.d[1];
  }
}

// Fixed-order sum of the per-ray (or per-block) dL/ds parts: one wavefront, strided
// partial sums, then a wave reduction.
__global__ void __launch_bounds__(64) composite_s_finish_kernel(const double* __restrict__ part, int64_t n,
                                                                float* __restrict__ d_s) {
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 64) acc += part[i];
  acc = wave_sum_d(acc);
  if (threadIdx.x == 0) d_s[0] = (float)acc;
}

constexpr int kAt = 256;

// compute_neus_alpha: one thread per sample; normals / dirs [n][3] (dirs per sample),
// sdf / dists [n].  The backward writes dL/dsdf, dL/dnormal and the block's dL/ds part.
template <bool BWD>
__global__ void __launch_bounds__(kAt) neus_alpha_kernel(const float* __restrict__ sdf,
                                                         const float* __restrict__ normals,
                                                         const float* __restrict__ dirs,
                                                         const float* __restrict__ dists,
                                                         const float* __restrict__ s, int64_t n, int iter_cos,
                                                         float cos_anneal, float* __restrict__ alpha,
                                                         const float* __restrict__ d_alpha, float* __restrict__ d_sdf,
                                                         float* __restrict__ d_normals, double* __restrict__ part_s) {
  const int64_t i = blockIdx.x * (int64_t)kAt + threadIdx.x;
  const float sv = s[0];
  double ds = 0.0;
  if (i < n) {
    const float nv[3] = {normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]};
    const float dv[3] = {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
    const AlphaTerms at = neus_alpha(sdf[i], nv, dv, dists[i], sv, iter_cos, cos_anneal);
    if (!BWD) {
      alpha[i] = at.araw;
    } else {
      float dsdf, dcos;
      neus_alpha_bwd(at, d_alpha[i], sv, iter_cos, cos_anneal, dsdf, dcos, ds);
      if (d_sdf) d_sdf[i] = dsdf;
      if (d_normals)
        for (int k = 0; k < 3; ++k) d_normals[3 * i + k] = dcos * dv[k];
    }
  }
  if (BWD && part_s) {
    __shared__ double red[kAt / 64];
    const double w = wave_sum_d(ds);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) part_s[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
  }
}

static int build_composite(const vxn_composite_args* p, CompositeDev& a) {
  VXN_REQUIRE(p != nullptr, VXN_ERR_NULL, "args is null");
  VXN_REQUIRE(p->n_rays >= 0 && p->n_rays < ((int64_t)1 << 31), VXN_ERR_SHAPE, "n_rays out of range");
  VXN_REQUIRE(p->n_samples >= 1 && p->n_samples <= kCt, VXN_ERR_UNSUPPORTED,
              "compositing takes 1 to %d samples per ray (got %d)", kCt, p->n_samples);
  if (p->alpha_in) {
    VXN_REQUIRE(p->sdf == nullptr, VXN_ERR_VALUE, "alpha_in replaces the SDF-based alpha: pass sdf = null");
  } else {
    VXN_REQUIRE(p->sdf && p->normals && p->ray_dirs && p->dists && p->s, VXN_ERR_NULL,
                "sdf, normals, ray_dirs, dists and s are required");
  }
  a = {};
  a.colors = p->colors;
  a.sdf = p->sdf;
  a.normals = p->normals;
  a.dirs = p->ray_dirs;
  a.dirs_per_sample = p->dirs_per_sample;
  a.z = p->z_vals;
  a.dists = p->dists;
  a.s = p->s;
  a.pts = p->pts;
  a.alpha_in = p->alpha_in;
  a.sphere_r = p->sphere_radius;
  a.cos_anneal = p->cos_anneal_ratio;
  a.iter_cos = p->use_iter_cos;
  a.R = p->n_rays;
  a.S = p->n_samples;
  a.rgb = p->rgb;
  a.depth = p->depth;
  a.normal = p->normal;
  a.opacity = p->opacity;
  a.weights = p->weights;
  a.alpha = p->alpha;
  a.trans = p->transmittance;
  a.d_rgb = p->d_rgb;
  a.d_depth = p->d_depth;
  a.d_normal = p->d_normal;
  a.d_opacity = p->d_opacity;
  a.d_weights = p->d_weights;
  a.d_alpha = p->d_alpha;
  a.d_trans = p->d_transmittance;
  a.d_colors = p->d_colors;
  a.d_sdf = p->d_sdf;
  a.d_normals = p->d_normals;
  a.d_alpha_in = p->d_alpha_in;
  a.d_s = p->d_s;
  return 0;
}

}  // namespace vxn

using namespace vxn;

extern "C" int vxn_composite_fwd(const vxn_composite_args* p, void* stream) {
  CompositeDev a;
  if (int e = build_composite(p, a)) return e;
  if (a.R == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (a.alpha_in)
    hipLaunchKernelGGL((composite_kernel<false, true>), dim3((unsigned)a.R), dim3(kCt), 0, s, a);
  else
    hipLaunchKernelGGL((composite_kernel<false, false>), dim3((unsigned)a.R), dim3(kCt), 0, s, a);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_composite_bwd_workspace_bytes(int64_t n_rays, int64_t* bytes) {
  VXN_REQUIRE(bytes != nullptr, VXN_ERR_NULL, "bytes is null");
  VXN_REQUIRE(n_rays >= 0, VXN_ERR_SHAPE, "n_rays must be >= 0");
  *bytes = n_rays * (int64_t)sizeof(double);
  return 0;
}

extern "C" int vxn_composite_bwd(const vxn_composite_args* p, void* workspace, int64_t workspace_bytes,
                                 void* stream) {
  CompositeDev a;
  if (int e = build_composite(p, a)) return e;
  hipStream_t s = (hipStream_t)stream;
  const bool want_s = !a.alpha_in && a.d_s != nullptr;
  if (want_s) {
    VXN_REQUIRE(workspace != nullptr && workspace_bytes >= a.R * (int64_t)sizeof(double), VXN_ERR_WORKSPACE,
                "compositing workspace needs %lld bytes", (long long)(a.R * (int64_t)sizeof(double)));
    a.part_s = (double*)workspace;
  }
  if (a.R > 0) {
    if (a.alpha_in)
      hipLaunchKernelGGL((composite_kernel<true, true>), dim3((unsigned)a.R), dim3(kCt), 0, s, a);
    else
      hipLaunchKernelGGL((composite_kernel<true, false>), dim3((unsigned)a.R), dim3(kCt), 0, s, a);
    VXN_LAUNCH_CHECK();
  }
  if (want_s) {
    hipLaunchKernelGGL(composite_s_finish_kernel, dim3(1), dim3(64), 0, s, (const double*)a.part_s, a.R, a.d_s);
    VXN_LAUNCH_CHECK();
  }
  return 0;
}

extern "C" int vxn_neus_alpha_workspace_bytes(int64_t n, int64_t* bytes) {
  VXN_REQUIRE(bytes != nullptr, VXN_ERR_NULL, "bytes is null");
  VXN_REQUIRE(n >= 0, VXN_ERR_SHAPE, "n must be >= 0");
  *bytes = ((n + kAt - 1) / kAt) * (int64_t)sizeof(double);
  return 0;
}

extern "C" int vxn_neus_alpha_fwd(const float* sdf, const float* normals, const float* dirs, const float* dists,
                                  const float* s, int64_t n, float cos_anneal_ratio, int32_t use_iter_cos,
                                  float* alpha, void* stream) {
  VXN_REQUIRE(n >= 0 && n < ((int64_t)1 << 40), VXN_ERR_SHAPE, "n out of range");
  if (n == 0) return 0;
  VXN_REQUIRE(sdf && normals && dirs && dists && s && alpha, VXN_ERR_NULL, "null argument");
  hipLaunchKernelGGL(neus_alpha_kernel<false>, dim3((unsigned)((n + kAt - 1) / kAt)), dim3(kAt), 0,
                     (hipStream_t)stream, sdf, normals, dirs, dists, s, n, use_iter_cos, cos_anneal_ratio, alpha,
                     nullptr, nullptr, nullptr, nullptr);
  VXN_LAUNCH_CHECK();
  return 0;
}

extern "C" int vxn_neus_alpha_bwd(const float* sdf, const float* normals, const float* dirs, const float* dists,
                                  const float* s, int64_t n, float cos_anneal_ratio, int32_t use_iter_cos,
                                  const float* d_alpha, float* d_sdf, float* d_normals, float* d_s, void* workspace,
                                  int64_t workspace_bytes, void* stream) {
  VXN_REQUIRE(n >= 0 && n < ((int64_t)1 << 40), VXN_ERR_SHAPE, "n out of range");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nb = (n + kAt - 1) / kAt;
  double* part = nullptr;
  if (d_s) {
    VXN_REQUIRE(workspace != nullptr && workspace_bytes >= nb * (int64_t)sizeof(double), VXN_ERR_WORKSPACE,
                "alpha workspace needs %lld bytes", (long long)(nb * (int64_t)sizeof(double)));
    part = (double*)workspace;
  }
  if (n > 0) {
    VXN_REQUIRE(sdf && normals && dirs && dists && s && d_alpha, VXN_ERR_NULL, "null argument");
    hipLaunchKernelGGL(neus_alpha_kernel<true>, dim3((unsigned)nb), dim3(kAt), 0, st, sdf, normals, dirs, dists, s,
                       n, use_iter_cos, cos_anneal_ratio, nullptr, d_alpha, d_sdf, d_normals, part);
    VXN_LAUNCH_CHECK();
  }
  if (d_s) {
    hipLaunchKernelGGL(composite_s_finish_kernel, dim3(1), dim3(64), 0, st, (const double*)part, nb, d_s);
    VXN_LAUNCH_CHECK();
  }
  return 0;
}
