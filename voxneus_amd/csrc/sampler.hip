// 4-step progressive importance upsampler (NeuS up_sample + sample_pdf + cat_z_vals),
// one ray per one-wavefront work-group (each lane holds two samples), all per-ray state in
// LDS.
//
// Rounding follows the reference's PyTorch CPU evaluation so that the searchsorted
// indices, and therefore the z values, come out bit-identical:
//   * sigmoid = SLEEF expf_u10 based (vxn_device.h);
//   * torch.cumprod / torch.cumsum accumulate float32 inputs in double, sequentially;
//   * torch.sum over a row of m floats uses an 8-lane vector cascade: 4 vector
//     accumulators over rows of 4 vectors, leftovers into accumulator 0, accumulators
//     folded 0+1+2+3, a sequential scalar tail, then the 8 lanes folded in order.
#include "vxn_host.h"
#include "vxn_scene.h"

namespace vxn {

constexpr int kSampThreads = 64;
constexpr int kMaxS = 128;
constexpr int kMaxPer = 64;

struct SamplerDev {
  const float *o, *d, *near, *far;
  int64_t R;
  GridDesc g;
  OccDesc oc;
  int n0, per, steps, perturb;
  const float *t_lin, *u_det, *t_rand, *u_rand;
  float *z_out, *dists_out, *mid_out;
  int32_t* inds;
  uint32_t* vr;  // nullable V_R bitmask: the corners of every final sample's cell
  uint32_t* vr_cells;  // nullable (then vr is null): one bit per final sample's cell (low corner)
  float bound_r;
  // serving batches: per-ray scene of a packed batch; R is then the capacity and the
  // batch size is read from n_rays_dev
  const SceneDev* scenes;
  const int32_t* ray_scene;
  const int32_t* n_rays_dev;
  int n_scenes;
};

// One sequential double-precision scan over the n <= 64 values v (lane j holds element j),
// continuing from acc: the exclusive product (PROD, written before each multiply) or the
// inclusive sum, rounded to float into lane j of out.  The order of the operations is the
// CPU's (torch.cumprod / torch.cumsum accumulate a float tensor in double, element by
// element).  Lane indices are constants after unrolling, so an element costs a lane read,
// two conversions, the double operation and one select.
template <bool PROD>
VXN_DEV void seq_scan(float v, int n, int t, double& acc, float& out) {
#pragma unroll
  for (int j = 0; j < 64; ++j) {
    if (j >= n) break;
    const double x = (double)__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
    if constexpr (PROD) {
      out = t == j ? (float)acc : out;
      acc = acc * x;
    } else {
      acc = acc + x;
      out = t == j ? (float)acc : out;
    }
  }
}

template <int MULTI>
__global__ void __launch_bounds__(kSampThreads, 8) upsample_kernel(SamplerDev a) {
  // One wavefront per ray: lane t holds samples t and t + 64 (kMaxS = 128).
  __shared__ float zs[kMaxS], sd[kMaxS], wv[kMaxS], cdf[kMaxS + 1], cosv[kMaxS];
  __shared__ float nz[kMaxPer], nzs[kMaxPer];
  __shared__ unsigned char ins[kMaxS];
  __shared__ float acc8[8];
  const int t = threadIdx.x;
  int64_t r0, r1, rstep;
  ray_range<MULTI>(a.R, a.n_rays_dev, r0, r1, rstep);
  for (int64_t ray = r0; ray < r1; ray += rstep) {
    const SceneDev* sc = nullptr;
    if constexpr (MULTI) {
      const int si = a.ray_scene[ray];
      if (si < 0 || si >= a.n_scenes) continue;  // the render writes zeros for such rays
      sc = a.scenes + si;
    }
    const GridDesc& G = MULTI ? sc->g : a.g;
    const OccDesc& OC = MULTI ? sc->oc : a.oc;
    const float ox = a.o[3 * ray], oy = a.o[3 * ray + 1], oz = a.o[3 * ray + 2];
    const float dx = a.d[3 * ray], dy = a.d[3 * ray + 1], dz = a.d[3 * ray + 2];
    float nr, fr;
    if (a.near) {
      nr = a.near[ray];
      fr = a.far[ray];
    } else {  // near_far_from_sphere with the reference's float operation order
      const float qa = (dx * dx + dy * dy) + dz * dz;
      const float qb = 2.0f * ((ox * dx + oy * dy) + oz * dz);
      const float mid = (0.5f * (-qb)) / qa;
      nr = fmaxf(mid - a.bound_r, 0.05f);
      fr = fmaxf(mid + a.bound_r, nr + 0.1f);
    }
    int n = a.n0;
    // 1. stratified samples (sampler.py:85-125)
    float z[2] = {0.0f, 0.0f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = t + 64 * h;
      if (i < n) {
        const float tl = a.t_lin[i];
        z[h] = nr * (1.0f - tl) + fr * tl;
        zs[i] = z[h];
      }
    }
    __syncthreads();
    if (a.perturb) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = t + 64 * h;
        if (i < n) {
          const float lower = i == 0 ? zs[0] : 0.5f * (zs[i] + zs[i - 1]);
          const float upper = i == n - 1 ? zs[n - 1] : 0.5f * (zs[i + 1] + zs[i]);
          z[h] = lower + (upper - lower) * a.t_rand[ray * n + i];
        }
      }
      __syncthreads();
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (t + 64 * h < n) zs[t + 64 * h] = z[h];
      __syncthreads();
    }
    // 2. SDF at the stratified points
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = t + 64 * h;
      if (i < n) {
        const float px = ox + dx * z[h], py = oy + dy * z[h], pz = oz + dz * z[h];
        sd[i] = gs_eval(G, gs_stencil(G, px, py, pz));
      }
    }
    __syncthreads();
    for (int step = 0; step < a.steps; ++step) {
      const float inv_s = 64.0f * (float)(1 << step);
      const int m = n - 1;
      const bool last = step + 1 == a.steps;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = t + 64 * h;
        if (i < n) {
          const float zz = zs[i];
          const float px = ox + dx * zz, py = oy + dy * zz, pz = oz + dz * zz;
          ins[i] = sqrtf(px * px + py * py + pz * pz) < 1.0f;
        }
        if (i < m) cosv[i] = (sd[i + 1] - sd[i]) / ((zs[i + 1] - zs[i]) + 1e-5f);
      }
      __syncthreads();
      float alpha[2] = {0.0f, 0.0f};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = t + 64 * h;
        if (i < m) {
          const float pc = i == 0 ? 0.0f : cosv[i - 1];
          float c = fminf(pc, cosv[i]);
          c = fminf(fmaxf(c, -1e3f), 0.0f) * ((ins[i] | ins[i + 1]) ? 1.0f : 0.0f);
          const float dist = zs[i + 1] - zs[i];
          const float mid = (sd[i] + sd[i + 1]) * 0.5f;
          const float pe = mid - c * dist * 0.5f;
          const float ne = mid + c * dist * 0.5f;
          const float pcdf = sigmoid_ref(pe * inv_s);
          const float ncdf = sigmoid_ref(ne * inv_s);
          alpha[h] = ((pcdf - ncdf) + 1e-5f) / (pcdf + 1e-5f);
        }
      }
      {
        // exclusive cumprod of (1 - alpha + 1e-7) in double, sequential (as torch.cumprod
        // on the CPU), walked lane by lane through the registers
        const float w0 = t < m ? (1.0f - alpha[0]) + 1e-7f : 0.0f;
        const float w1 = t + 64 < m ? (1.0f - alpha[1]) + 1e-7f : 0.0f;
        double T = 1.0;
        float c0 = 0.0f, c1 = 0.0f;
        seq_scan<true>(w0, m, t, T, c0);
        seq_scan<true>(w1, m - 64, t, T, c1);
        const float cc[2] = {c0, c1};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int i = t + 64 * h;
          if (i < m) {
            float w = alpha[h] * cc[h];
            if (OC.occ) {
              const float mz = (zs[i] + zs[i + 1]) * 0.5f;
              w = w * occ_factor(OC, ox + dx * mz, oy + dy * mz, oz + dz * mz);
            }
            wv[i] = w + 1e-5f;
          }
        }
      }
      __syncthreads();
      // 3. sample_pdf: row sum in the 8-lane cascade order
      const int nvec = m >> 3, ilp = nvec >> 2;
      if (t < 8) {
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int i = 0; i < ilp; ++i)
#pragma unroll
          for (int k = 0; k < 4; ++k) acc[k] = acc[k] + wv[8 * (4 * i + k) + t];
        for (int i = 4 * ilp; i < nvec; ++i) acc[0] = acc[0] + wv[8 * i + t];
        acc8[t] = ((acc[0] + acc[1]) + acc[2]) + acc[3];
      }
      __syncthreads();
      {
        float tot = 0.0f;
        if (t == 0) {
          if (nvec == 0) {  // rows shorter than one vector: scalar 4-accumulator cascade
            float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            const int q = m >> 2;
            for (int i = 0; i < q; ++i)
              for (int k = 0; k < 4; ++k) acc[k] = acc[k] + wv[4 * i + k];
            for (int i = 4 * q; i < m; ++i) acc[0] = acc[0] + wv[i];
            tot = ((acc[0] + acc[1]) + acc[2]) + acc[3];
          } else {
            for (int k = 8 * nvec; k < m; ++k) tot = tot + wv[k];
            for (int l = 0; l < 8; ++l) tot = tot + acc8[l];
          }
        }
        tot = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(tot)));
        // pdf = w / sum (independent divisions, in parallel), then the cdf as a
        // sequential double cumsum (torch.cumsum on the CPU), walked lane by lane
        const float q0 = t < m ? wv[t] / tot : 0.0f, q1 = t + 64 < m ? wv[t + 64] / tot : 0.0f;
        double c = 0.0;
        float r0 = 0.0f, r1 = 0.0f;
        seq_scan<false>(q0, m, t, c, r0);
        seq_scan<false>(q1, m - 64, t, c, r1);
        if (t == 0) cdf[0] = 0.0f;
        if (t < m) cdf[t + 1] = r0;
        if (t + 64 < m) cdf[t + 65] = r1;
      }
      __syncthreads();
      if (t < a.per) {
        const float u = a.perturb ? a.u_rand[((int64_t)step * a.R + ray) * a.per + t] : a.u_det[t];
        int lo = 0, hi = n;  // searchsorted(cdf[0..n-1], u, right=True)
        while (lo < hi) {
          const int md = (lo + hi) >> 1;
          if (cdf[md] <= u)
            lo = md + 1;
          else
            hi = md;
        }
        const int ind = lo;
        const int below = max(ind - 1, 0), above = min(ind, n - 1);
        const float cg0 = cdf[below], cg1 = cdf[above];
        float den = cg1 - cg0;
        if (den < 1e-5f) den = 1.0f;
        const float tt = (u - cg0) / den;
        const float b0 = zs[below], b1 = zs[above];
        nz[t] = b0 + tt * (b1 - b0);
        if (a.inds) a.inds[((int64_t)step * a.R + ray) * a.per + t] = ind;
      }
      __syncthreads();
      // 4. cat_z_vals: stable rank of the new samples, then merge
      if (t < a.per) {
        const float v = nz[t];
        int r = 0;
        for (int j = 0; j < a.per; ++j) {
          const float u2 = nz[j];
          r += (u2 < v) || (u2 == v && j < t);
        }
        nzs[r] = v;
      }
      __syncthreads();
      float zo[2] = {0.0f, 0.0f}, so[2] = {0.0f, 0.0f}, zn = 0.0f, sn = 0.0f;
      int po[2] = {-1, -1}, pn = -1;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = t + 64 * h;
        if (i < n) {
          zo[h] = zs[i];
          so[h] = sd[i];
          int lo = 0, hi = a.per;  // count of new < zo
          while (lo < hi) {
            const int md = (lo + hi) >> 1;
            if (nzs[md] < zo[h])
              lo = md + 1;
            else
              hi = md;
          }
          po[h] = i + lo;
        }
      }
      if (t < a.per) {
        zn = nzs[t];
        int lo = 0, hi = n;  // count of old <= zn
        while (lo < hi) {
          const int md = (lo + hi) >> 1;
          if (zs[md] <= zn)
            lo = md + 1;
          else
            hi = md;
        }
        pn = t + lo;
        if (!last) {
          const float px = ox + dx * zn, py = oy + dy * zn, pz = oz + dz * zn;
          sn = gs_eval(G, gs_stencil(G, px, py, pz));
        }
      }
      __syncthreads();
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (po[h] >= 0) {
          zs[po[h]] = zo[h];
          sd[po[h]] = so[h];
        }
      }
      if (pn >= 0) {
        zs[pn] = zn;
        sd[pn] = sn;
      }
      __syncthreads();
      n += a.per;
    }
    // 5. final intervals and midpoints (sampler.py:197-202)
    float midv[2] = {0.0f, 0.0f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = t + 64 * h;
      if (i < n) {
        const float zz = zs[i];
        const float dist = i < n - 1 ? zs[i + 1] - zz : zs[n - 1] - zs[n - 2];
        midv[h] = zz + dist * 0.5f;
        a.z_out[ray * n + i] = zz;
        a.dists_out[ray * n + i] = dist;
        a.mid_out[ray * n + i] = midv[h];
      }
    }
    if (!MULTI && (a.vr || a.vr_cells)) {
      // V_R of the render's samples (their positions are o + d * mid, as the render forms
      // them), marked here so that the regularisers need not wait for the render: a
      // sample in the same cell as the sample before it marks nothing new.  With vr_cells
      // a sample sets one bit (its cell's) in place of its 8 corners' (4-8 atomics), and
      // vr_dilate_kernel turns the cells into their corners afterwards
      Corners cn[2];
      int key[2] = {-1, -1};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = t + 64 * h;
        if (i < n) {
          cn[h] = corner_decomp(G, ox + dx * midv[h], oy + dy * midv[h], oz + dz * midv[h]);
          key[h] = (int)lin3(G, cn[h].c0[0], cn[h].c0[1], cn[h].c0[2]);
        }
      }
      const int kp0 = __shfl_up(key[0], 1, 64);
      const int k63 = __shfl(key[0], 63, 64);
      const int kp1 = t == 0 ? k63 : __shfl_up(key[1], 1, 64);
      if (a.vr_cells) {
        if (t < n && (t == 0 || kp0 != key[0])) atomicOr(a.vr_cells + (key[0] >> 5), 1u << (key[0] & 31));
        if (t + 64 < n && kp1 != key[1]) atomicOr(a.vr_cells + (key[1] >> 5), 1u << (key[1] & 31));
      } else {
        if (t < n && (t == 0 || kp0 != key[0])) vr_mark_cell(G, cn[0], a.vr);
        if (t + 64 < n && kp1 != key[1]) vr_mark_cell(G, cn[1], a.vr);
      }
    }
    __syncthreads();
  }
}

// V_R from the cell marks: vertex (x, y, z) is a corner of cell c exactly when
// c = (x - i, y - j, z - k) for some i, j, k in {0, 1} (corner_decomp's corners are c0 and
// min(c0 + 1, res - 1) per axis), so each bitmask word is the OR of the cell words of its
// own and the three lower (x, y) columns, each with itself shifted up by one along z (the
// bit below a word's first comes from the column's previous word).  Every column holds
// whole words (Z a multiple of 32).  Overwrites vr.
__global__ void __launch_bounds__(256) vr_dilate_kernel(const uint32_t* __restrict__ cells, uint32_t* __restrict__ vr,
                                                        int Y, int wpc, int64_t nwords) {
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w >= nwords) return;
  const int64_t col = w / wpc;
  const int zw = (int)(w - col * wpc);
  const int64_t x = col / Y;
  const int y = (int)(col - x * Y);
  uint32_t acc = 0u;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (x < i || y < j) continue;
      const int64_t cw = w - ((int64_t)i * Y + j) * wpc;
      const uint32_t c = cells[cw];
      const uint32_t below = zw > 0 ? cells[cw - 1] : 0u;
      acc |= c | (c << 1) | (below >> 31);
    }
  vr[w] = acc;
}

int launch_serve_upsample(const SceneDev* scenes, int n_scenes, const int32_t* ray_scene, const int32_t* n_dev,
                          int64_t cap, const float* o, const float* d, int n0, int per, int steps,
                          const float* t_lin, const float* u_det, float bound_r, float* z, float* dists, float* mid,
                          hipStream_t stream) {
  SamplerDev a{};
  a.o = o;
  a.d = d;
  a.near = nullptr;
  a.far = nullptr;
  a.R = cap;
  a.n0 = n0;
  a.per = per;
  a.steps = steps;
  a.perturb = 0;
  a.t_lin = t_lin;
  a.u_det = u_det;
  a.z_out = z;
  a.dists_out = dists;
  a.mid_out = mid;
  a.inds = nullptr;
  a.bound_r = bound_r;
  a.scenes = scenes;
  a.ray_scene = ray_scene;
  a.n_rays_dev = n_dev;
  a.n_scenes = n_scenes;
  a.vr = nullptr;
  a.vr_cells = nullptr;
  const int nb = cap < 8192 ? (int)cap : 8192;
  hipLaunchKernelGGL(upsample_kernel<1>, dim3(nb), dim3(kSampThreads), 0, stream, a);
  VXN_LAUNCH_CHECK();
  return 0;
}

}  // namespace vxn

using namespace vxn;

extern "C" int vxn_upsample(const vxn_sampler_args* p, void* stream) {
  VXN_REQUIRE(p != nullptr, VXN_ERR_NULL, "args is null");
  const vxn_sampler_args& A = *p;
  if (int e = check_grid(A.grid, A.X, A.Y, A.Z, A.box)) return e;
  VXN_REQUIRE(A.n_rays >= 0, VXN_ERR_SHAPE, "n_rays must be >= 0");
  VXN_REQUIRE(A.n_samples >= 2, VXN_ERR_SHAPE, "n_samples must be >= 2");
  VXN_REQUIRE(A.up_sample_steps >= 0 && A.n_importance >= 0, VXN_ERR_VALUE, "negative importance settings");
  const int steps = (A.n_importance > 0 && A.up_sample_steps > 0) ? A.up_sample_steps : 0;
  const int per = steps ? (A.n_importance / steps > 1 ? A.n_importance / steps : 1) : 0;
  const int S = A.n_samples + steps * per;
  VXN_REQUIRE(S <= kMaxS, VXN_ERR_UNSUPPORTED, "at most %d samples per ray are supported (got %d)", kMaxS, S);
  VXN_REQUIRE(per <= kMaxPer, VXN_ERR_UNSUPPORTED, "at most %d importance samples per step", kMaxPer);
  VXN_REQUIRE(steps <= 20, VXN_ERR_UNSUPPORTED, "at most 20 upsampling steps");
  if (A.n_rays == 0) return 0;
  VXN_REQUIRE(A.rays_o && A.rays_d && A.t_lin && A.z_vals && A.dists && A.mid_z, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE((A.near == nullptr) == (A.far == nullptr), VXN_ERR_NULL, "near and far must both be given or both null");
  if (!A.near) VXN_REQUIRE(A.bound_radius > 0.0f, VXN_ERR_VALUE, "bound_radius must be > 0 when near/far are null");
  VXN_REQUIRE(!steps || A.perturb || A.u_det, VXN_ERR_NULL, "u_det is required in deterministic mode");
  VXN_REQUIRE(!A.perturb || (A.t_rand && (!steps || A.u_rand)), VXN_ERR_NULL,
              "t_rand and u_rand are required when perturb is set");
  if (A.occ_mask) VXN_REQUIRE(A.occ_res >= 1, VXN_ERR_SHAPE, "occ_res must be >= 1");
  SamplerDev a{};
  a.o = A.rays_o;
  a.d = A.rays_d;
  a.near = A.near;
  a.far = A.far;
  a.R = A.n_rays;
  a.g = make_grid_desc(A.grid, A.X, A.Y, A.Z, A.box);
  a.oc = make_occ_desc(A.occ_mask, A.occ_res, A.occ_box);
  a.n0 = A.n_samples;
  a.per = per;
  a.steps = steps;
  a.perturb = A.perturb;
  a.t_lin = A.t_lin;
  a.u_det = A.u_det;
  a.t_rand = A.t_rand;
  a.u_rand = A.u_rand;
  a.z_out = A.z_vals;
  a.dists_out = A.dists;
  a.mid_out = A.mid_z;
  a.inds = A.inds;
  // cell marks and the dilation pass where every grid column holds whole bitmask words
  const bool cells = A.vr_bitmask && A.vr_cells && A.Z % 32 == 0;
  a.vr = cells ? nullptr : A.vr_bitmask;
  a.vr_cells = cells ? A.vr_cells : nullptr;
  a.bound_r = A.bound_radius;
  a.scenes = nullptr;
  a.ray_scene = nullptr;
  a.n_rays_dev = nullptr;
  a.n_scenes = 0;
  const int nb = A.n_rays < 8192 ? (int)A.n_rays : 8192;
  hipLaunchKernelGGL(upsample_kernel<0>, dim3(nb), dim3(kSampThreads), 0, (hipStream_t)stream, a);
  VXN_LAUNCH_CHECK();
  if (cells) {
    const int64_t nwords = (int64_t)A.X * A.Y * (A.Z / 32);
    hipLaunchKernelGGL(vr_dilate_kernel, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const uint32_t*)A.vr_cells, A.vr_bitmask, A.Y, A.Z / 32, nwords);
    VXN_LAUNCH_CHECK();
  }
  return 0;
}
