// Fused per-ray render: forward, backward and the fused training step.
//
// render_kernel: one ray per 128-thread work-group (two wave64s), thread t = sample t of
// the ray.  A persistent grid walks the rays; the MLP parameter gradients are accumulated
// in registers across all rays of a work-group and written once as partials that a
// fixed-order reduction sums (deterministic).  In the backward the per-sample gradients
// w.r.t. the hash features and the SDF geometry are written as records (feature-major,
// so every store coalesces) instead of being scattered in place.
//
// hash_scatter_kernel / grid_scatter_kernel replay those records at full occupancy: a
// wavefront covers 64 consecutive samples, samples whose gradient is exactly zero are
// skipped, and runs of consecutive samples in the same cell are summed with a
// wavefront segmented scan so each run issues one set of atomics.
//
// Reference: VoxNeuS.forward_points and the render passes 2-3
// (models/voxneus.py:115-167, 226-258), compute_neus_alpha (rendering/neus_alpha.py:12-81),
// composite_weights and volume_render_from_samples (rendering/volume_render.py:12-153),
// the Huber colour loss (rendering/volume_render.py:43-54), and the scatters of the
// reference backward: index_put into the hash table (radiance_hash.py:151-160) and the
// Eq. 7 / grid_sample gradients (cuda/gradient_interp.cu:86-152).
#include <string.h>

#include "vxn_block.h"
#include "vxn_hash.h"
#include "vxn_host.h"
#include "vxn_mfma.h"
#include "vxn_scene.h"

namespace vxn {

constexpr int kRT = 128;
constexpr int kNP = VXN_MLP_PARAMS;
constexpr int kOffB0 = 64 * 51;
constexpr int kOffW1 = kOffB0 + 64;
constexpr int kOffB1 = kOffW1 + 3 * 64;

// Parameter layout of the radiance MLP with a view embedding of NV values (16 SH, 27
// positional encoding): W0 [64][32 + NV + 3] (hash features, view embedding, normal, the
// order RadianceMLP concatenates them), b0 [64], W1 [3][64], b1 [3], flattened in turn.
template <int NV>
struct MlpLayout {
  static constexpr int WS = 35 + NV;  // W0 row stride
  static constexpr int B0 = 64 * WS;
  static constexpr int W1 = B0 + 64;
  static constexpr int B1 = W1 + 3 * 64;
  static constexpr int NP = B1 + 3;
};
constexpr int kNPMax = MlpLayout<27>::NP;
static_assert(MlpLayout<16>::NP == kNP, "SH layout");

template <int NV>
VXN_DEV void view_embedding(float x, float y, float z, float (&o)[NV]) {
  if constexpr (NV == 16)
    sh16(x, y, z, o);
  else
    pe27(x, y, z, o);
}
constexpr int kXS = 37;  // LDS row stride (floats): 32 hash features + 3 normal components
constexpr int kCS = 17;  // LDS row stride (floats) for a 16-hidden-unit chunk
constexpr int kStage = kRT * kXS + kRT * 3 + 2 * kRT * kCS;
constexpr int kGeo = 4;  // geometry gradient record: dL/dsdf, dL/dgrad[3]

enum { kFwd = 0, kBwd = 1, kTrain = 2 };
enum { kMlpF32 = 0, kMlpBf16 = 1 };

// BF16 matrix-core MLP: weight fragments staged in LDS once per work-group (u32x4 units,
// one per lane of a fragment; fragments that only a few lanes use are stored compactly).
constexpr int kFragFW = 0;                 // layer 1, W0x [h][k]: 6 fragments (mt, ks)
constexpr int kFragDX = kFragFW + 6 * 64;  // dX, W0x^T [k<32][h]: 4 fragments (mt, q)
constexpr int kFragDXN = kFragDX + 4 * 64; // dX, W0x^T [32..34][h]: lanes j < 3 only
constexpr int kFragL2 = kFragDXN + 4 * 6;  // layer 2, W1 [c][h]: lanes j < 3 only
constexpr int kFragDH = kFragL2 + 4 * 6;   // dH, W1^T [h][c]: lanes g = 0, as u32x2
constexpr int kFragWords = kFragDH + 2 * 32 / 2;
constexpr int kTS = 136;  // LDS row stride (bf16) of the 128-sample transposed tiles

struct RenderDev {
  const float *o, *d, *mid, *dist;
  int64_t R;
  int S;
  vxn_fastdiv sdiv;  // sample / S (R * S < 2^31)
  GridDesc g;
  HashDesc h;
  const float* mlp;
  float s, sphere_r, cos_anneal;
  const float* s_dev;  // nullable: s read from device memory
  const float* var_dev;  // nullable: s = clamp(exp(10 v), 1e-4, s_max) from the variance v
  float s_max;
  int iter_cos;
  float *rgb, *depth, *normal, *opacity, *weights, *alpha, *trans, *sdf_s, *grad_s, *normal_s;
  int64_t* corners;
  const float *d_rgb, *d_depth, *d_normal, *d_opacity, *d_weights;
  const float* target;
  float huber_delta;
  float inv_n3;
  // training colour loss: VXN_LOSS_*; target_mask [R] (nullable) weights it by the
  // foreground and adds mask_weight * BCE(opacity, mask); loss_norm (device, in the
  // workspace): [0] the colour loss's normalisation, [1] 1 if the mask weights it
  int loss_type;
  const float* target_mask;
  float mask_weight;
  float inv_r;
  const float* loss_norm;
  float* grad_grid;
  float* grad_table;
  uint32_t* vr;
  float* part_mlp;  // [gridDim.x][kNP]
  double* part_sc;  // [gridDim.x][3]: dL/ds, colour-loss sum, mask-BCE sum
  float* gfeat;     // [32][R*S] dL/d(hash features), feature-major
  float* ggeo;      // [4][R*S]  dL/dsdf, dL/dgrad
  // BF16 backward / training: hash features gathered beforehand by hash_features_kernel,
  // [16][R*S] bf16 pairs packed in uint32
  void* hfeat;
  // backward / training with the binned hash-table gradient: live[i] = 1 where sample i has
  // a nonzero weight or alpha gradient (every other sample has zero gradients), then the
  // live samples' indices in order (live_idx, n_live of them), which the record passes walk
  uint8_t* live;
  const int32_t* live_idx;
  const int32_t* n_live;
  // BF16 backward / training: the geometry pass's SDF and gradient of every sample
  // ([4][R*S]: sdf, gx, gy, gz), read by the render kernel in place of its own stencil, and
  // need[i] = 0 where sample i's colour reaches no output or gradient (q < 0 or outside
  // the sphere), whose hash features the gather then skips (zeros)
  float* geo;
  uint8_t* need;
  // the record passes walk lists (binned hash-table gradient): ggeo (and in the FP32 form
  // gfeat) is then written only for samples that a list can hold (0 <= q, inside)
  int sparse;
  // serving batches (forward only): per-ray scene of a packed batch; R is the capacity
  // and the batch size is read from n_rays_dev
  const SceneDev* scenes;
  const int32_t* ray_scene;
  const int32_t* n_rays_dev;
  int n_scenes;
};

// W0 restricted to the 35 per-sample inputs: k < 32 hash features, 32..34 the normal
// (W0 columns 32 + NV .. 34 + NV), zero beyond.
template <int NV>
VXN_DEV float w0x(const float* W0, int h, int k) {
  constexpr int WS = MlpLayout<NV>::WS;
  return k < 32 ? W0[h * WS + k] : (k < 35 ? W0[h * WS + NV + k] : 0.0f);
}

// Builds the BF16 weight fragments of the matrix-core MLP in LDS (all threads).
template <int NV>
VXN_DEV void build_mlp_fragments(const float* W0, const float* W1, u32x4_t* frag, int t) {
  for (int idx = t; idx < kFragWords; idx += kRT) {
    float v[8];
    if (idx < kFragDX) {  // A = W0x: row h = 32mt + j, K = k = 16ks + 8g + e
      const int f = idx >> 6, l = idx & 63, mt = f / 3, ks = f % 3;
      for (int e = 0; e < 8; ++e) v[e] = w0x<NV>(W0, 32 * mt + (l & 31), 16 * ks + 8 * (l >> 5) + e);
    } else if (idx < kFragDXN) {  // A = W0x^T: row k = j, K = hidden kperm(mt, q, g, e)
      const int f = (idx - kFragDX) >> 6, l = (idx - kFragDX) & 63;
      for (int e = 0; e < 8; ++e) v[e] = w0x<NV>(W0, kperm(f >> 1, f & 1, l >> 5, e), l & 31);
    } else if (idx < kFragL2) {  // rows k = 32 + j (j < 3) of W0x^T
      const int f = (idx - kFragDXN) / 6, r = (idx - kFragDXN) % 6;
      for (int e = 0; e < 8; ++e) v[e] = w0x<NV>(W0, kperm(f >> 1, f & 1, r / 3, e), 32 + r % 3);
    } else if (idx < kFragDH) {  // A = W1: row c = j (j < 3), K = hidden kperm(mt, q, g, e)
      const int f = (idx - kFragL2) / 6, r = (idx - kFragL2) % 6;
      for (int e = 0; e < 8; ++e) v[e] = W1[(r % 3) * 64 + kperm(f >> 1, f & 1, r / 3, e)];
    } else {  // A = W1^T: row h = 32mt + j, K = c = e (< 3), lanes g = 0; two per u32x4
      const int i2 = 2 * (idx - kFragDH);
      for (int u = 0; u < 2; ++u) {
        const int h = 32 * ((i2 + u) >> 5) + ((i2 + u) & 31);
        v[4 * u] = W1[h];
        v[4 * u + 1] = W1[64 + h];
        v[4 * u + 2] = W1[128 + h];
        v[4 * u + 3] = 0.0f;
      }
    }
    frag[idx] = u32x4_t{pk_bf16(v[0], v[1]), pk_bf16(v[2], v[3]), pk_bf16(v[4], v[5]), pk_bf16(v[6], v[7])};
  }
}

// SDF (grid_sample) and Eq. 7 / Eq. 4 gradient at a sample position, for the geometry
// pass; the render kernel writes the same operations out itself (keeping the cell for its
// other outputs), so both get the same values.
VXN_DEV void sample_sdf_grad(const GridDesc& G, float px, float py, float pz, float& sdf, float (&gv)[3]) {
  const GsStencil gs = gs_stencil(G, px, py, pz);
  const Corners cn = corner_decomp(G, px, py, pz);
  if (!G.analytical && brick_interior(G, cn)) {
    Brick bk;
    load_brick(G, cn, bk);
    if (!gs_eval_brick(gs, cn, bk, sdf)) sdf = gs_eval(G, gs);
    eq7_gradient_brick(G, cn, bk, gv);
  } else {
    sdf = gs_eval(G, gs);
    grid_gradient(G, cn, gv);
  }
}

// Eq. 2 of one sample (neus_alpha.py:12-81) up to the clamp: q = (p_prev - p_next) /
// max(p_prev, 1e-5) with the iter_cos rule, and the inside-sphere factor.  A sample with
// q < 0 or outside the sphere has alpha = 0 and, since the alpha gradient passes only
// where 0 <= q <= 1 and inside, no gradient depends on its colour.
struct NeusQ {
  float tcos, off, fp, fn, pp, pn, den, q, inside;
};

VXN_DEV NeusQ neus_q(const float nrm[3], float dx, float dy, float dz, float dist, float sdf, float sv, float sr,
                     float px, float py, float pz, int iter_cos, float cos_anneal) {
  NeusQ r;
  r.tcos = (nrm[0] * dx + nrm[1] * dy) + nrm[2] * dz;
  float cosv = r.tcos;
  if (iter_cos) {
    const float ra = fmaxf(-r.tcos * 0.5f + 0.5f, 0.0f), rb = fmaxf(-r.tcos, 0.0f);
    cosv = -(ra * (1.0f - cos_anneal) + rb * cos_anneal);
  }
  r.off = 0.5f * dist * cosv;
  r.fp = sdf - r.off;
  r.fn = sdf + r.off;
  r.pp = sigmoid_ref(sv * r.fp);
  r.pn = sigmoid_ref(sv * r.fn);
  r.den = fmaxf(r.pp, 1e-5f);
  r.q = (r.pp - r.pn) / r.den;
  r.inside = (sr > 0.0f) ? ((sqrtf(px * px + py * py + pz * pz) < sr) ? 1.0f : 0.0f) : 1.0f;
  return r;
}

template <int MODE, int MLPM, int MULTI, int NV>
__global__ void __launch_bounds__(kRT, 2) render_kernel(RenderDev a) {
  using ML = MlpLayout<NV>;
  constexpr int VM = (NV + 7) / 8;  // view-embedding columns per thread in the FP32 weight gradient
  constexpr bool kBF = MLPM == kMlpBf16;
  constexpr bool kBFB = kBF && MODE != kFwd;
  // BF16 backward / training and BF16 serving: the hash features come from
  // hash_features_kernel (a.hfeat)
  constexpr bool kPre = kBF && (MODE != kFwd || MULTI);
  // BF16 backward / training: the SDF and its gradient come from geo_kernel (a.geo)
  __shared__ Red red;
  __shared__ __align__(16) float preh[64];  // b0 + W0[:, 32:32+NV] . view(d): the per-ray part of the hidden layer
  __shared__ int nlive_s[2];  // samples per wavefront with a non-zero output gradient
  __shared__ float stage[(MODE == kFwd || kBF) ? 1 : kStage];
  // BF16 form: weight fragments, the ray's SH values, and the transposed bf16 tiles of the
  // parameter-gradient products: inputs xt[35][s], hidden (then dL/dhidden) hdt[64][s],
  // output gradients ot[3][s]; s = sample of the ray (0..127).
  __shared__ u32x4_t bfrag[kBF ? kFragWords : 1];
  __shared__ float shl[32];
  __shared__ uint64_t live_s[2];
  __shared__ __align__(16) uint16_t xt[kBFB ? 35 * kTS : 8];
  __shared__ __align__(16) uint16_t hdt[kBFB ? 64 * kTS : 8];
  __shared__ __align__(16) uint16_t ot[kBFB ? 3 * kTS : 8];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const float* W0 = a.mlp;
  const float* b0 = a.mlp + ML::B0;
  const float* W1 = a.mlp + ML::W1;
  const float* b1 = a.mlp + ML::B1;
  const int64_t NS = a.R * (int64_t)a.S;
  // LDS staging for the parameter gradients: inputs Xs[128][kXS] (hash features and
  // normal), output grads Os[128][3], and per 16-hidden-unit chunk Hs/Ds[128][kCS].
  float* Xs = stage;
  float* Os = stage + kRT * kXS;
  float* Hs = Os + kRT * 3;
  float* Ds = Hs + kRT * kCS;
  // Per-thread ownership of parameter-gradient outputs within each hidden chunk:
  // hidden j = 16*chunk + (t & 15); input columns {g, g+8, ...} of the 35 staged ones
  // (g = t >> 4) and view-embedding columns {g, g+8, ...}; W1 (c, j) for t < 48; b1[c]
  // for t < 3.
  const int jl = t & 15, grp = t >> 4;
  float accX[4][5], accSH[4][VM], accW1[4], accB0[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
#pragma unroll
    for (int k = 0; k < 5; ++k) accX[c][k] = 0.0f;
#pragma unroll
    for (int m = 0; m < VM; ++m) accSH[c][m] = 0.0f;
    accW1[c] = 0.0f;
    accB0[c] = 0.0f;
  }
  float accb1 = 0.0f;
  double ds_acc = 0.0, loss_acc = 0.0, bce_acc = 0.0;
  // BF16 form: parameter-gradient accumulators; wave wid owns hidden rows 32*wid + crow.
  //   gw0[kt]: columns k = 32kt + j of [W0x | b0 | W0 SH columns]; gw1: W1^T columns c = j.
  const int jq = lane & 31, gq = lane >> 5;
  f32x16_t gw0[2] = {}, gw1 = {};
  float accb[3] = {0.0f, 0.0f, 0.0f};
  if constexpr (kBF && !MULTI) {
    build_mlp_fragments<NV>(W0, W1, bfrag, t);
    __syncthreads();
  }
  int frag_scene = -1;  // serving batches: scene whose MLP the fragments hold

  int64_t r0, r1, rstep;
  ray_range<MULTI>(a.R, a.n_rays_dev, r0, r1, rstep);
  for (int64_t ray = r0; ray < r1; ray += rstep) {
    const bool valid = t < a.S;
    const int64_t si = ray * a.S + t;
    const SceneDev* scn = nullptr;
    if constexpr (MULTI) {
      const int sx = a.ray_scene[ray];
      if (sx < 0 || sx >= a.n_scenes) {  // no such scene: zeros
        if (t < 3) a.rgb[3 * ray + t] = 0.0f;
        if (t == 0 && a.depth) a.depth[ray] = 0.0f;
        if (t == 0 && a.opacity) a.opacity[ray] = 0.0f;
        if (t < 3 && a.normal) a.normal[3 * ray + t] = 0.0f;
        continue;
      }
      scn = a.scenes + sx;
      if constexpr (kBF) {
        if (sx != frag_scene) {
          __syncthreads();
          build_mlp_fragments<NV>(scn->mlp, scn->mlp + ML::W1, bfrag, t);
          frag_scene = sx;
        }
      }
    }
    const GridDesc& G = MULTI ? scn->g : a.g;
    const HashDesc& Hd = MULTI ? scn->h : a.h;
    const float sv = MULTI ? scn->s : (a.var_dev ? s_of_variance(*a.var_dev, a.s_max) : (a.s_dev ? *a.s_dev : a.s));
    const float sr = MULTI ? scn->sphere_r : a.sphere_r;
    const float* W0 = MULTI ? scn->mlp : a.mlp;
    const float* b0 = W0 + ML::B0;
    const float* W1 = W0 + ML::W1;
    const float* b1 = W0 + ML::B1;
    const float ox = a.o[3 * ray], oy = a.o[3 * ray + 1], oz = a.o[3 * ray + 2];
    const float dx = a.d[3 * ray], dy = a.d[3 * ray + 1], dz = a.d[3 * ray + 2];
    const float mid = valid ? a.mid[si] : 0.0f;
    const float dist = valid ? a.dist[si] : 0.0f;
    // per-ray view embedding and its contribution to the hidden pre-activations
    float sh[NV];
    view_embedding<NV>(dx, dy, dz, sh);
    if (t < 64) {
      const float* wr = W0 + t * ML::WS + 32;
      float acc = b0[t];
#pragma unroll
      for (int i = 0; i < NV; ++i) acc = fmaf(wr[i], sh[i], acc);
      preh[t] = acc;
    }
    float shv[VM];  // this thread's view-embedding columns grp + 8m
#pragma unroll
    for (int m = 0; m < VM; ++m) shv[m] = (grp & 7) + 8 * m < NV ? sh[(grp & 7) + 8 * m] : 0.0f;
    if (kBF && t == 0) {
#pragma unroll
      for (int i = 0; i < NV; ++i) shl[i] = sh[i];
    }
    // ---------------------------------------------------------------- sample forward
    const float px = ox + dx * mid, py = oy + dy * mid, pz = oz + dz * mid;
    float sdf, gv[3];
    Corners cn;  // the sample's cell: the stencil's, then the corners output and the V_R marks
    if constexpr (kBFB && !MULTI) {  // the geometry pass's values (sample_sdf_grad, run ahead)
      sdf = valid ? a.geo[si] : 0.0f;
#pragma unroll
      for (int k = 0; k < 3; ++k) gv[k] = valid ? a.geo[(1 + k) * NS + si] : 0.0f;
    } else {
      // sample_sdf_grad's code, with the cell kept for the outputs below (written out here:
      // the FP32 serving render measured 18 % slower through the helper)
      const GsStencil gs = gs_stencil(G, px, py, pz);
      cn = corner_decomp(G, px, py, pz);
      if (!G.analytical && brick_interior(G, cn)) {
        Brick bk;
        load_brick(G, cn, bk);
        if (!gs_eval_brick(gs, cn, bk, sdf)) sdf = gs_eval(G, gs);
        eq7_gradient_brick(G, cn, bk, gv);
      } else {
        sdf = gs_eval(G, gs);
        grid_gradient(G, cn, gv);
      }
    }
    const float gr = norm3(gv);
    const float nrm[3] = {gv[0] / (gr + 1e-8f), gv[1] / (gr + 1e-8f), gv[2] / (gr + 1e-8f)};
    const float x0 = fminf(fmaxf(over_den(G, 0, px - G.mn[0]), 0.0f), 1.0f);
    const float x1 = fminf(fmaxf(over_den(G, 1, py - G.mn[1]), 0.0f), 1.0f);
    const float x2 = fminf(fmaxf(over_den(G, 2, pz - G.mn[2]), 0.0f), 1.0f);
    // Eq. 2 up to the clamp (needs only the geometry): ahead of the features where it decides
    // whether they are gathered (FP32 backward / training), after the MLP otherwise
    constexpr bool kEarlyQ = MODE != kFwd && !kPre;
    NeusQ nq;
    if constexpr (kEarlyQ) nq = neus_q(nrm, dx, dy, dz, dist, sdf, sv, sr, px, py, pz, a.iter_cos, a.cos_anneal);
    float x[35];  // 32 hash features, then the unit normal (MLP input columns 0-31, 48-50)
    if constexpr (!kPre) {
      // backward / training: a sample with q < 0 or outside the sphere has alpha = 0 and no
      // gradient through its colour, so its features are not gathered (zeros in their place)
      if (MODE == kFwd || (!(nq.q < 0.0f) && nq.inside != 0.0f)) {
#pragma unroll
        for (int l = 0; l < 16; ++l) hash_level_fwd<kBF>(Hd, l, x0, x1, x2, x[2 * l], x[2 * l + 1]);
      } else {
#pragma unroll
        for (int k = 0; k < 32; ++k) x[k] = 0.0f;
      }
    }
    x[32] = nrm[0];
    x[33] = nrm[1];
    x[34] = nrm[2];
    __syncthreads();  // preh ready
    float o3[3] = {b1[0], b1[1], b1[2]};
    uint64_t hmask = 0;  // BF16 form: relu'(hidden) per C element, bit 16 (2 mt + nt) + v
    if constexpr (!kBF) {
      // MLP forward, streaming the hidden units (one W0 row per iteration)
#pragma unroll 2
      for (int j = 0; j < 64; ++j) {
        const float* wr = W0 + j * ML::WS;
        float acc = preh[j];
#pragma unroll
        for (int i = 0; i < 32; ++i) acc = fmaf(wr[i], x[i], acc);
#pragma unroll
        for (int k = 0; k < 3; ++k) acc = fmaf(wr[32 + NV + k], x[32 + k], acc);
        const float hj = fmaxf(acc, 0.0f);
        o3[0] = fmaf(W1[j], hj, o3[0]);
        o3[1] = fmaf(W1[64 + j], hj, o3[1]);
        o3[2] = fmaf(W1[128 + j], hj, o3[2]);
      }
    } else {
      // Matrix cores, transposed: hidden^T [64 h x 64 samples] = W0x [64 x 48] . X^T,
      // samples of wave half nt = 0/1 in the N dimension (lane j <-> sample 32nt + j).
      uint32_t xp[24];
      if constexpr (!kPre) {
#pragma unroll
        for (int r = 0; r < 16; ++r) xp[r] = pk_bf16(x[2 * r], x[2 * r + 1]);
      } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) xp[r] = valid ? ((const uint32_t*)a.hfeat)[r * NS + si] : 0u;
      }
      xp[16] = pk_bf16(x[32], x[33]);
      xp[17] = pk_bf16(x[34], 0.0f);
#pragma unroll
      for (int r = 18; r < 24; ++r) xp[r] = 0u;
      if constexpr (kBFB) {
        // stage the inputs for the parameter gradients
#pragma unroll
        for (int k = 0; k < 35; ++k) xt[k * kTS + t] = (uint16_t)(xp[k >> 1] >> (16 * (k & 1)));
      }
      // one 32-sample tile nt at a time: hidden^T tile pair -> relu -> layer 2
      // out^T [3 (of 32) x 32 samples] = W1 . relu(hidden)^T
      float on[2][3];
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        u32x4_t hp[2][2];  // relu(hidden) as bf16, [mt][q]: the B operand of layer 2
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
          f32x16_t c;
#pragma unroll
          for (int q4 = 0; q4 < 4; ++q4) {
            const float4 p = *(const float4*)(preh + 32 * mt + 8 * q4 + 4 * gq);
            c[4 * q4] = p.x;
            c[4 * q4 + 1] = p.y;
            c[4 * q4 + 2] = p.z;
            c[4 * q4 + 3] = p.w;
          }
#pragma unroll
          for (int ks = 0; ks < 3; ++ks) {
            u32x4_t bx;
#pragma unroll
            for (int r = 0; r < 4; ++r) bx[r] = half_swap(xp[8 * ks + r], xp[8 * ks + 4 + r])[nt];
            c = mfma_bf16(bfrag[kFragFW + (3 * mt + ks) * 64 + lane], bx, c);
          }
#pragma unroll
          for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int v = 8 * q + 2 * r;
              const float h0 = c[v], h1 = c[v + 1];
              hmask |= (uint64_t)(h0 > 0.0f) << (16 * (2 * mt + nt) + v);
              hmask |= (uint64_t)(h1 > 0.0f) << (16 * (2 * mt + nt) + v + 1);
              const uint32_t pk = pk_bf16(fmaxf(h0, 0.0f), fmaxf(h1, 0.0f));
              hp[mt][q][r] = pk;
              if constexpr (kBFB) {
                uint16_t* col = hdt + 64 * wid + 32 * nt + jq;
                col[(32 * mt + crow(v, gq)) * kTS] = (uint16_t)pk;
                col[(32 * mt + crow(v + 1, gq)) * kTS] = (uint16_t)(pk >> 16);
              }
            }
        }
        f32x16_t oc = {};
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const u32x4_t A = jq < 3 ? bfrag[kFragL2 + (2 * mt + q) * 6 + 3 * gq + jq] : u32x4_t{0u, 0u, 0u, 0u};
            oc = mfma_bf16(A, hp[mt][q], oc);
          }
        on[nt][0] = oc[0];
        on[nt][1] = oc[1];
        on[nt][2] = oc[2];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c)
        o3[c] += __uint_as_float(half_swap(__float_as_uint(on[0][c]), __float_as_uint(on[1][c]))[0]);
    }
    float rgb[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) rgb[c] = sigmoid_ref(o3[c]);
    // ---------------------------------------------------------------- Eq. 2 alpha
    if constexpr (!kEarlyQ) nq = neus_q(nrm, dx, dy, dz, dist, sdf, sv, sr, px, py, pz, a.iter_cos, a.cos_anneal);
    const float tcos = nq.tcos, fp = nq.fp, fn = nq.fn, pp = nq.pp, pn = nq.pn, den = nq.den, q = nq.q;
    const float inside = nq.inside;
    const float araw = fminf(fmaxf(q, 0.0f), 1.0f);
    const float alpha = valid ? araw * inside : 0.0f;
    const float ov = (1.0f - alpha) + 1e-7f;
    const float oma = fminf(fmaxf(ov, 0.0f), 1.0f);
    // ---------------------------------------------------------------- transmittance
    const float T = (float)block_excl_prod((double)oma, red, lane, wid);
    const bool keep = !(T < 1e-4f);
    const float w = keep ? alpha * T : 0.0f;
    float comp[8] = {w * rgb[0], w * rgb[1], w * rgb[2], w * mid, w, w * nrm[0], w * nrm[1], w * nrm[2]};
    block_sums<8>(comp, red, lane, wid);
    const float Nr = sqrtf(comp[5] * comp[5] + comp[6] * comp[6] + comp[7] * comp[7]);

    if (MODE == kTrain && t == 0 && a.rgb) {
      a.rgb[3 * ray] = comp[0];
      a.rgb[3 * ray + 1] = comp[1];
      a.rgb[3 * ray + 2] = comp[2];
    }
    if (MODE == kFwd) {
      if (t == 0) {
        if (a.rgb) {
          a.rgb[3 * ray] = comp[0];
          a.rgb[3 * ray + 1] = comp[1];
          a.rgb[3 * ray + 2] = comp[2];
        }
        if (a.depth) a.depth[ray] = comp[3];
        if (a.opacity) a.opacity[ray] = comp[4];
        if (a.normal) {
          a.normal[3 * ray] = comp[5] / (Nr + 1e-8f);
          a.normal[3 * ray + 1] = comp[6] / (Nr + 1e-8f);
          a.normal[3 * ray + 2] = comp[7] / (Nr + 1e-8f);
        }
      }
      if (valid) {
        if (a.weights) a.weights[si] = w;
        if (a.alpha) a.alpha[si] = alpha;
        if (a.trans) a.trans[si] = T;
        if (a.sdf_s) a.sdf_s[si] = sdf;
        if (a.grad_s) {
          a.grad_s[3 * si] = gv[0];
          a.grad_s[3 * si + 1] = gv[1];
          a.grad_s[3 * si + 2] = gv[2];
        }
        if (a.normal_s) {
          a.normal_s[3 * si] = nrm[0];
          a.normal_s[3 * si + 1] = nrm[1];
          a.normal_s[3 * si + 2] = nrm[2];
        }
        if (a.corners) {
          if constexpr (kBFB && !MULTI) cn = corner_decomp(G, px, py, pz);
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            int cx, cy, cz;
            corner_xyz(cn, k, cx, cy, cz);
            a.corners[24 * si + 3 * k] = cx;
            a.corners[24 * si + 3 * k + 1] = cy;
            a.corners[24 * si + 3 * k + 2] = cz;
          }
        }
      }
    }
    if (a.vr) {
      // V_R marks: a sample in the same cell as the previous lane's marks nothing new.
      if constexpr (kBFB && !MULTI) cn = corner_decomp(G, px, py, pz);
      const int key = valid ? (int)lin3(G, cn.c0[0], cn.c0[1], cn.c0[2]) : -1;
      const int kp = __shfl_up(key, 1, 64);
      if (valid && (lane == 0 || kp != key)) vr_mark_cell(G, cn, a.vr);
    }
    if (MODE == kFwd) continue;

    // ---------------------------------------------------------------- upstream gradients
    float dr[3] = {0.0f, 0.0f, 0.0f}, dd = 0.0f, dop = 0.0f, dN[3] = {0.0f, 0.0f, 0.0f};
    if (MODE == kTrain) {
      // colour loss (trainer.py:213-236): Huber or MSE per channel, plain mean over the
      // batch, or with a foreground mask the mask-weighted sum over the mask's sum (the
      // plain mean when the mask sums to zero)
      const float* mk = a.target_mask;
      const float fw = (mk && a.loss_norm[1] != 0.0f) ? mk[ray] : 1.0f;
      const float nrm_l = mk ? a.loss_norm[0] : a.inv_n3;
      double ls = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float df = comp[c] - a.target[3 * ray + c];
        float l, dl;
        if (a.loss_type == VXN_LOSS_MSE) {
          l = df * df;
          dl = 2.0f * df;
        } else {
          const float ad = fabsf(df);
          const bool small = ad < a.huber_delta;
          l = small ? 0.5f * (ad * ad) / a.huber_delta : ad - 0.5f * a.huber_delta;
          const float sg = df > 0.0f ? 1.0f : (df < 0.0f ? -1.0f : 0.0f);
          dl = (small ? (ad / a.huber_delta) : 1.0f) * sg;
        }
        ls += (double)(l * fw);
        dr[c] = (dl * fw) * nrm_l;
      }
      if (t == 0) loss_acc += ls;
      // mask BCE (trainer.py:242-245): mask_weight * mean BCE(clamp(opacity, 1e-6, 1 - 1e-6))
      if (mk && a.mask_weight > 0.0f) {
        const float op = comp[4], m = mk[ray];
        const float xo = fminf(fmaxf(op, 1e-6f), 1.0f - 1e-6f);
        if (t == 0)
          bce_acc -= (double)(m * fmaxf(logf(xo), -100.0f) + (1.0f - m) * fmaxf(logf(1.0f - xo), -100.0f));
        const bool pass = op >= 1e-6f && op <= 1.0f - 1e-6f;
        dop = pass ? a.mask_weight * ((xo - m) / fmaxf((1.0f - xo) * xo, 1e-12f)) * a.inv_r : 0.0f;
      }
    } else {
      if (a.d_rgb) {
        dr[0] = a.d_rgb[3 * ray];
        dr[1] = a.d_rgb[3 * ray + 1];
        dr[2] = a.d_rgb[3 * ray + 2];
      }
      if (a.d_depth) dd = a.d_depth[ray];
      if (a.d_opacity) dop = a.d_opacity[ray];
      if (a.d_normal) {
        // normal = N / (|N| + 1e-8)
        const float gn[3] = {a.d_normal[3 * ray], a.d_normal[3 * ray + 1], a.d_normal[3 * ray + 2]};
        const float rr = Nr + 1e-8f;
        const float dotv = comp[5] * gn[0] + comp[6] * gn[1] + comp[7] * gn[2];
#pragma unroll
        for (int k = 0; k < 3; ++k) dN[k] = gn[k] / rr - (Nr > 0.0f ? comp[5 + k] * dotv / (Nr * rr * rr) : 0.0f);
      }
    }
    float gw = (dr[0] * rgb[0] + dr[1] * rgb[1]) + dr[2] * rgb[2] + dd * mid + dop +
               ((dN[0] * nrm[0] + dN[1] * nrm[1]) + dN[2] * nrm[2]);
    if (MODE == kBwd && a.d_weights && valid) gw += a.d_weights[si];
    if (!valid || !keep) gw = 0.0f;
    // suffix sums of g_i * alpha_i * T_i for the d(T)/d(oma) terms, in double
    double tot;
    const double qi = (double)gw * (double)alpha * (double)T;
    const double pre = block_incl_sum(qi, tot, red, lane, wid);
    const double suffix = tot - pre;
    float dA = gw * T;
    if (ov <= 1.0f && ov >= 0.0f) dA -= (float)(suffix / (double)oma);
    const float dAraw = valid ? dA * inside : 0.0f;
    const float dq = (q >= 0.0f && q <= 1.0f) ? dAraw : 0.0f;
    // every gradient of the sample is a multiple of w (colour, normal) or of dq (SDF,
    // cosine): zero both and the sample adds nothing (an empty-space sample, alpha clamped
    // to 0, has a nonzero dA from the samples behind it but dq = 0)
    const bool glive = w != 0.0f || dq != 0.0f;
    float dpp = dq / den;
    const float dpn = -dq / den;
    if (pp >= 1e-5f) dpp -= dq * q / den;
    const float dzp = dpp * (1.0f - pp) * pp;
    const float dzn = dpn * (1.0f - pn) * pn;
    ds_acc += (double)dzp * fp + (double)dzn * fn;
    const float dfp = dzp * sv, dfn = dzn * sv;
    const float dsdf = dfp + dfn;
    const float doff = dfn - dfp;
    float dcos = doff * 0.5f * dist;
    if (a.iter_cos) {
      const float ra = -tcos * 0.5f + 0.5f, rb = -tcos;
      dcos = dcos * (-((ra > 0.0f ? -0.5f : 0.0f) * (1.0f - a.cos_anneal) + (rb > 0.0f ? -1.0f : 0.0f) * a.cos_anneal));
    }
    float dn[3] = {dcos * dx + w * dN[0], dcos * dy + w * dN[1], dcos * dz + w * dN[2]};

    // ---------------------------------------------------------------- MLP backward
    float dout[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) dout[c] = valid ? (w * dr[c]) * (1.0f - rgb[c]) * rgb[c] : 0.0f;
    if constexpr (!kBF) {
    // Samples with a zero output gradient add nothing to the parameter gradients: stage
    // only the live ones, compacted to rows [0, nlive).
    const bool mlive = valid && (dout[0] != 0.0f || dout[1] != 0.0f || dout[2] != 0.0f);
    const uint64_t bal = __ballot(mlive);
    if (lane == 0) nlive_s[wid] = __popcll(bal);
    __syncthreads();
    const int nlive = nlive_s[0] + nlive_s[1];
    const int slot = (wid == 0 ? 0 : nlive_s[0]) + __popcll(bal & ((1ull << lane) - 1ull));
    const bool wave_live = bal != 0;
    if (mlive) {
#pragma unroll
      for (int i = 0; i < 35; ++i) Xs[slot * kXS + i] = x[i];
#pragma unroll
      for (int k = 0; k < 3; ++k) Os[slot * 3 + k] = dout[k];
    }
    float dxa[35];
#pragma unroll
    for (int i = 0; i < 35; ++i) dxa[i] = 0.0f;
#pragma unroll 1
    for (int cc = 0; cc < (nlive > 0 ? 4 : 0); ++cc) {
      // recompute this chunk's hidden units; dz0 = relu'(h) * (W1^T dout); dx += W0^T dz0
      if (wave_live) {
#pragma unroll 2
        for (int jj = 0; jj < 16; ++jj) {
          const int j = 16 * cc + jj;
          const float* wr = W0 + j * ML::WS;
          float acc = preh[j];
#pragma unroll
          for (int i = 0; i < 32; ++i) acc = fmaf(wr[i], x[i], acc);
#pragma unroll
          for (int k = 0; k < 3; ++k) acc = fmaf(wr[32 + NV + k], x[32 + k], acc);
          const float hj = fmaxf(acc, 0.0f);
          const float dh = (W1[j] * dout[0] + W1[64 + j] * dout[1]) + W1[128 + j] * dout[2];
          const float dzj = hj > 0.0f ? dh : 0.0f;
#pragma unroll
          for (int i = 0; i < 32; ++i) dxa[i] = fmaf(wr[i], dzj, dxa[i]);
#pragma unroll
          for (int k = 0; k < 3; ++k) dxa[32 + k] = fmaf(wr[32 + NV + k], dzj, dxa[32 + k]);
          if (mlive) {
            Hs[slot * kCS + jj] = hj;
            Ds[slot * kCS + jj] = dzj;
          }
        }
      }
      __syncthreads();
      // accumulate this chunk's parameter gradients over the ray's live samples
      float aX[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, dsum = 0.0f, aW1 = 0.0f;
      const int c1 = t < 48 ? (t >> 4) : 0;
#pragma unroll 4
      for (int s2 = 0; s2 < nlive; ++s2) {
        const float dv = Ds[s2 * kCS + jl];
        const float* xr = Xs + s2 * kXS + grp;
#pragma unroll
        for (int k = 0; k < 5; ++k)
          if (grp + 8 * k < 35) aX[k] = fmaf(dv, xr[8 * k], aX[k]);
        dsum += dv;
        aW1 = fmaf(Hs[s2 * kCS + jl], Os[s2 * 3 + c1], aW1);
      }
#pragma unroll
      for (int k = 0; k < 5; ++k) accX[cc][k] += aX[k];
#pragma unroll
      for (int m = 0; m < VM; ++m) accSH[cc][m] = fmaf(shv[m], dsum, accSH[cc][m]);
      accB0[cc] += dsum;
      accW1[cc] += aW1;
      __syncthreads();
    }
    if (t < 3)
      for (int s2 = 0; s2 < nlive; ++s2) accb1 += Os[s2 * 3 + t];
#pragma unroll
    for (int k = 0; k < 3; ++k) dn[k] += dxa[32 + k];
    if (valid && (!a.sparse || (!(q < 0.0f) && inside != 0.0f))) {
#pragma unroll
      for (int i = 0; i < 32; ++i) a.gfeat[i * NS + si] = dxa[i];
    }
    } else {
      // Parameter gradients are products over the ray's 128 samples (K); 16-sample chunks
      // without an output gradient are skipped.  Wave wid produces hidden rows 32 wid + crow.
      const bool mlive = valid && (dout[0] != 0.0f || dout[1] != 0.0f || dout[2] != 0.0f);
      const uint64_t bal = __ballot(mlive);
      if (lane == 0) live_s[wid] = bal;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        accb[c] += dout[c];
        ot[c * kTS + t] = bf16_bits(dout[c]);
      }
      __syncthreads();
      const uint64_t live[2] = {live_s[0], live_s[1]};
      const uint16_t* hrow = hdt + (32 * wid + jq) * kTS + 8 * gq;
      // dW1^T [32 h x 3 (of 32)] += relu(hidden) . dout
#pragma unroll 2
      for (int ks = 0; ks < 8; ++ks) {
        if (((live[ks >> 2] >> (16 * (ks & 3))) & 0xFFFFull) == 0) continue;
        const u32x4_t A = *(const u32x4_t*)(hrow + 16 * ks);
        const u32x4_t B = jq < 3 ? *(const u32x4_t*)(ot + jq * kTS + 16 * ks + 8 * gq) : u32x4_t{0u, 0u, 0u, 0u};
        gw1 = mfma_bf16(A, B, gw1);
      }
      __syncthreads();  // the hidden tile is consumed; it now receives dL/dhidden
      // dL/dhidden^T [64 x 64 samples] = relu' * (W1^T [64 x 3 (of 16)] . dout^T)
      const u32x2_t s0 = half_swap(pk_bf16(dout[0], dout[1]), 0u);
      const u32x2_t s1 = half_swap(pk_bf16(dout[2], 0.0f), 0u);
      float dxn[2][3];
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        u32x4_t dp[2][2];  // dL/dhidden of sample tile nt as bf16, [mt][q]
        uint16_t* col = hdt + 64 * wid + 32 * nt + jq;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
          u32x4_t A = {0u, 0u, 0u, 0u};
          if (gq == 0) {
            const u32x2_t w = ((const u32x2_t*)(bfrag + kFragDH))[32 * mt + jq];
            A = u32x4_t{w[0], w[1], 0u, 0u};
          }
          const f32x16_t c = mfma_bf16(A, u32x4_t{s0[nt], s1[nt], 0u, 0u}, f32x16_t{});
#pragma unroll
          for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int v = 8 * q + 2 * r, b = 16 * (2 * mt + nt) + v;
              const uint32_t pk =
                  pk_bf16(((hmask >> b) & 1) ? c[v] : 0.0f, ((hmask >> (b + 1)) & 1) ? c[v + 1] : 0.0f);
              dp[mt][q][r] = pk;
              col[(32 * mt + crow(v, gq)) * kTS] = (uint16_t)pk;
              col[(32 * mt + crow(v + 1, gq)) * kTS] = (uint16_t)(pk >> 16);
            }
        }
        // dL/dX^T [35 (of 64) x 32 samples] = W0x^T . dL/dhidden^T: the hash-feature rows
        // go straight to the records (row k, 32 consecutive samples per half-wave store),
        // the normal rows back to the sample's own lane
        f32x16_t c = {}, cn = {};
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            c = mfma_bf16(bfrag[kFragDX + (2 * mt + q) * 64 + lane], dp[mt][q], c);
            const u32x4_t A =
                jq < 3 ? bfrag[kFragDXN + (2 * mt + q) * 6 + 3 * gq + jq] : u32x4_t{0u, 0u, 0u, 0u};
            cn = mfma_bf16(A, dp[mt][q], cn);
          }
        const int sl = 64 * wid + 32 * nt + jq;
        if (sl < a.S) {
          float* rec = a.gfeat + ray * a.S + sl;
#pragma unroll
          for (int v = 0; v < 16; ++v) rec[crow(v, gq) * NS] = c[v];
        }
        dxn[nt][0] = cn[0];
        dxn[nt][1] = cn[1];
        dxn[nt][2] = cn[2];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k)
        dn[k] += __uint_as_float(half_swap(__float_as_uint(dxn[0][k]), __float_as_uint(dxn[1][k]))[0]);
      __syncthreads();  // dL/dhidden staged
      // [dW0x | db0 | dW0 view] [32 h x 64] += dL/dhidden . [X | 1 | view]; columns 32 + j of
      // the second tile: j < 3 the normal, 3 the bias, 4 .. 3 + NV the view embedding, the
      // rest zero
      uint32_t bc = 0u;
      if (jq == 3) bc = pk_bf16(1.0f, 1.0f);
      else if (jq >= 4 && jq < 4 + NV) bc = pk_bf16(shl[jq - 4], shl[jq - 4]);
      const u32x4_t Bc = {bc, bc, bc, bc};
#pragma unroll 2
      for (int ks = 0; ks < 8; ++ks) {
        if (((live[ks >> 2] >> (16 * (ks & 3))) & 0xFFFFull) == 0) continue;
        const u32x4_t A = *(const u32x4_t*)(hrow + 16 * ks);
        const u32x4_t B0 = *(const u32x4_t*)(xt + jq * kTS + 16 * ks + 8 * gq);
        const u32x4_t B1 = jq < 3 ? *(const u32x4_t*)(xt + (32 + jq) * kTS + 16 * ks + 8 * gq) : Bc;
        gw0[0] = mfma_bf16(A, B0, gw0[0]);
        gw0[1] = mfma_bf16(A, B1, gw0[1]);
      }
    }
    // ---------------------------------------------------------------- gradient records
    if (valid) {
      const float rr = gr + 1e-8f;
      const float dotv = (dn[0] * gv[0] + dn[1] * gv[1]) + dn[2] * gv[2];
      if (a.live) a.live[si] = glive ? 1 : 0;
      // (the BF16 form reads the geometry pass's flag: q and inside are not kept that long)
      if (!a.sparse || (kBFB ? (a.need[si] & 1) != 0 : (!(q < 0.0f) && inside != 0.0f))) {
        a.ggeo[si] = dsdf;
#pragma unroll
        for (int k = 0; k < 3; ++k)
          a.ggeo[(1 + k) * NS + si] = dn[k] / rr - (gr > 0.0f ? gv[k] * dotv / (gr * rr * rr) : 0.0f);
      }
    }
    __syncthreads();
  }

  if (MODE == kFwd) return;
  // ---------------------------------------------------------------- write partials
  float* P = a.part_mlp + (int64_t)blockIdx.x * ML::NP;
  if constexpr (kBF) {
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int h = 32 * wid + crow(v, gq), k = 32 * kt + jq;
        // k: 0..31 hash features, 32..34 normal (W0 columns 32 + NV ..), 35 bias,
        // 36 .. 35 + NV the view embedding (columns 32 .. 31 + NV)
        if (k < 32) P[h * ML::WS + k] = gw0[kt][v];
        else if (k < 35) P[h * ML::WS + NV + k] = gw0[kt][v];
        else if (k == 35) P[ML::B0 + h] = gw0[kt][v];
        else if (k < 36 + NV) P[h * ML::WS + k - 4] = gw0[kt][v];
      }
#pragma unroll
    for (int v = 0; v < 16; ++v)
      if (jq < 3) P[ML::W1 + jq * 64 + 32 * wid + crow(v, gq)] = gw1[v];
    block_sums<3>(accb, red, lane, wid);
    if (t < 3) P[ML::B1 + t] = accb[t];
  } else {
#pragma unroll
  for (int cc = 0; cc < 4; ++cc) {
    const int j = 16 * cc + jl;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const int col = grp + 8 * k;  // 0..31 hash features, 32..34 normal -> W0 columns 32 + NV ..
      if (col < 35) P[j * ML::WS + (col < 32 ? col : col + NV)] = accX[cc][k];
    }
#pragma unroll
    for (int m = 0; m < VM; ++m)
      if (grp + 8 * m < NV) P[j * ML::WS + 32 + grp + 8 * m] = accSH[cc][m];
    if (grp == 0) P[ML::B0 + j] = accB0[cc];
    if (t < 48) P[ML::W1 + (t >> 4) * 64 + j] = accW1[cc];
  }
  if (t < 3) P[ML::B1 + t] = accb1;
  }
  double dsw = wave_sum_d(ds_acc);
  if (lane == 0) red.d[wid] = dsw;
  __syncthreads();
  if (t == 0) {
    a.part_sc[3 * blockIdx.x] = red.d[0] + red.d[1];
    a.part_sc[3 * blockIdx.x + 1] = loss_acc;
    a.part_sc[3 * blockIdx.x + 2] = bce_acc;
  }
}

// ------------------------------------------------------------------------------------
// Wavefront segmented sum: lanes with equal keys in consecutive runs; every lane gets
// the sum over [run start, itself]; returns true on the run's last lane.  The lane
// shifts are DPP moves (one VALU instruction each, no LDS round trip): a Hillis-Steele
// scan over each 16-lane row (row_shr 1, 2, 4, 8), then the row carries of the GFX9 scan
// (row_bcast:15 adds lane 15 of rows 0 / 2 to rows 1 / 3, row_bcast:31 adds lane 31 to
// rows 2 and 3).  A lane takes a shifted value only when its run reaches back to the
// source lane, so a carried value is itself a sum over the same run.
// ------------------------------------------------------------------------------------
constexpr int kDppRowShr = 0x110;      // + n: row_shr:n
constexpr int kDppWaveShl1 = 0x130;    // lane j reads lane j + 1
constexpr int kDppWaveShr1 = 0x138;    // lane j reads lane j - 1
constexpr int kDppRowBcast15 = 0x142;  // lane 15 of each row to the next row
constexpr int kDppRowBcast31 = 0x143;  // lane 31 to rows 2 and 3

// The DPP move of v; lanes without a source, or in rows outside ROW_MASK, read 0.
template <int CTRL, int ROW_MASK>
VXN_DEV int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xF, true);
}
template <int CTRL, int ROW_MASK>
VXN_DEV float dpp_f(float v) {
  return __int_as_float(dpp_i<CTRL, ROW_MASK>(__float_as_int(v)));
}

// One step of the segmented sum: every lane takes the DPP-shifted value (full EXEC: a DPP
// source lane must be active), then the lanes of `mask` add it, under EXEC = mask; the step
// is skipped when no lane needs it.
constexpr int kSegCtrl[6] = {kDppRowShr + 1, kDppRowShr + 2, kDppRowShr + 4, kDppRowShr + 8, kDppRowBcast15,
                             kDppRowBcast31};
constexpr int kSegRows[6] = {0xF, 0xF, 0xF, 0xF, 0xA, 0xC};

template <int STEP, int NV>
VXN_DEV void seg_step(uint64_t mask, float (&v)[NV], bool c) {
  if (mask == 0) return;
  float u[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) u[k] = dpp_f<kSegCtrl[STEP], kSegRows[STEP]>(v[k]);
  if (c) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = v[k] + u[k];
  }
}

template <int NV>
VXN_DEV bool seg_sum(uint64_t key, float (&v)[NV], int lane) {
  const int klo = (int)(uint32_t)key, khi = (int)(uint32_t)(key >> 32);
  const uint64_t kp = (uint64_t)(uint32_t)dpp_i<kDppWaveShr1, 0xF>(klo) |
                      ((uint64_t)(uint32_t)dpp_i<kDppWaveShr1, 0xF>(khi) << 32);
  const uint64_t kn = (uint64_t)(uint32_t)dpp_i<kDppWaveShl1, 0xF>(klo) |
                      ((uint64_t)(uint32_t)dpp_i<kDppWaveShl1, 0xF>(khi) << 32);
  // run start of every lane: a max-scan of the starts (lanes without a source read 0)
  int st = (lane == 0 || kp != key) ? lane : 0;
  st = max(st, dpp_i<kDppRowShr + 1, 0xF>(st));
  st = max(st, dpp_i<kDppRowShr + 2, 0xF>(st));
  st = max(st, dpp_i<kDppRowShr + 4, 0xF>(st));
  st = max(st, dpp_i<kDppRowShr + 8, 0xF>(st));
  st = max(st, dpp_i<kDppRowBcast15, 0xA>(st));
  st = max(st, dpp_i<kDppRowBcast31, 0xC>(st));
  const int r = lane & 15;
  const bool c1 = r >= 1 && lane - 1 >= st, c2 = r >= 2 && lane - 2 >= st, c4 = r >= 4 && lane - 4 >= st,
             c8 = r >= 8 && lane - 8 >= st;
  const bool c15 = ((lane >> 4) & 1) != 0 && st <= (lane & ~15) - 1, c31 = lane >= 32 && st <= 31;
  // each step adds the shifted value on the lanes whose run reaches the source lane: one
  // DPP add per value with EXEC set to those lanes, and no step at all when no lane of the
  // wavefront needs it (short runs skip the longer shifts)
  seg_step<0>(__ballot(c1), v, c1);
  seg_step<1>(__ballot(c2), v, c2);
  seg_step<2>(__ballot(c4), v, c4);
  seg_step<3>(__ballot(c8), v, c8);
  seg_step<4>(__ballot(c15), v, c15);
  seg_step<5>(__ballot(c31), v, c31);
  return lane == 63 || kn != key;
}

// Sample position of record i (ray = i / S).
VXN_DEV int64_t ray_of(const RenderDev& a, int64_t i) { return (int64_t)vxn_fastdiv_apply(a.sdiv, (uint32_t)i); }

// The record passes walk the live samples in order when the render listed them: position k
// of the walk is sample live_idx[k], and past the walk's length the sample count R * S (out
// of range: no records).  Without the list the walk is every sample.
VXN_DEV int64_t walk_length(const RenderDev& a) { return a.live_idx ? (int64_t)*a.n_live : a.R * (int64_t)a.S; }
VXN_DEV int64_t walk_sample(const RenderDev& a, int64_t k, int64_t nwalk) {
  if (k >= nwalk) return a.R * (int64_t)a.S;
  return a.live_idx ? (int64_t)a.live_idx[k] : k;
}

VXN_DEV void sample_pos(const RenderDev& a, int64_t i, float& px, float& py, float& pz) {
  const int64_t ray = ray_of(a, i);
  const float m = a.mid[i];
  px = a.o[3 * ray] + a.d[3 * ray] * m;
  py = a.o[3 * ray + 1] + a.d[3 * ray + 1] * m;
  pz = a.o[3 * ray + 2] + a.d[3 * ray + 2] * m;
}

// Geometry pass of the BF16 backward / training (one sample per thread, ahead of the
// hash-feature gather): the sample's SDF and gradient (the render kernel reads them in
// place of its own stencil) and whether its colour matters (need): a sample with q < 0 or
// outside the sphere has alpha = 0 and no gradient through its colour, so its hash
// features are not gathered; with q = 0 its colour still enters its alpha gradient but its
// features get no gradient (w = 0), so the hash-table gradient's records skip it too.
// q comes from neus_q on these values, as in the render kernel, so both agree.
__global__ void __launch_bounds__(256) geo_kernel(RenderDev a) {
  const int64_t NS = a.R * (int64_t)a.S;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= NS) return;
  const int64_t ray = ray_of(a, i);
  const float ox = a.o[3 * ray], oy = a.o[3 * ray + 1], oz = a.o[3 * ray + 2];
  const float dx = a.d[3 * ray], dy = a.d[3 * ray + 1], dz = a.d[3 * ray + 2];
  const float mid = a.mid[i], dist = a.dist[i];
  const float px = ox + dx * mid, py = oy + dy * mid, pz = oz + dz * mid;
  float sdf, gv[3];
  sample_sdf_grad(a.g, px, py, pz, sdf, gv);
  a.geo[i] = sdf;
  a.geo[NS + i] = gv[0];
  a.geo[2 * NS + i] = gv[1];
  a.geo[3 * NS + i] = gv[2];
  const float gr = norm3(gv);
  const float nrm[3] = {gv[0] / (gr + 1e-8f), gv[1] / (gr + 1e-8f), gv[2] / (gr + 1e-8f)};
  const float sv = a.var_dev ? s_of_variance(*a.var_dev, a.s_max) : (a.s_dev ? *a.s_dev : a.s);
  const NeusQ nq = neus_q(nrm, dx, dy, dz, dist, sdf, sv, a.sphere_r, px, py, pz, a.iter_cos, a.cos_anneal);
  // bit 0: 0 <= q, inside (the colour and the SDF gradients can matter); bit 1: 0 < q,
  // inside (alpha > 0: the hash features can have a gradient)
  const bool in = nq.inside != 0.0f;
  a.need[i] = (uint8_t)(((!(nq.q < 0.0f) && in) ? 1 : 0) | ((nq.q > 0.0f && in) ? 2 : 0));
}

// Hash features of every sample, level-major and XCD-aware.  Work-groups are dispatched to
// the 8 XCDs round-robin (work-group b runs on XCD b % 8), so work-group b gathers level
// (b % 8) + 8 * pass for one 256-sample chunk, pass 0 for all chunks first: each XCD reads
// one level's table at a time (2^19 entries: 2 MB bf16) instead of every XCD streaming
// all 16 tables through its L2.  The values are those of hash_level_fwd at the render
// kernel's sample position (same operations), rounded to bf16 as the BF16 MLP's inputs.
template <int MULTI>
__global__ void __launch_bounds__(256) hash_features_kernel(RenderDev a, int64_t chunks) {
  const int64_t NS = a.R * (int64_t)a.S;
  const int64_t r = blockIdx.x >> 3;
  const int l = (int)(blockIdx.x & 7) + 8 * (int)(r / chunks);
  if (l >= 16) return;
  const int64_t i = (r % chunks) * (int64_t)blockDim.x + threadIdx.x;
  if (i >= NS) return;
  const GridDesc* g = &a.g;
  const HashDesc* h = &a.h;
  if constexpr (MULTI) {  // serving batch: the ray's scene, rays past the batch size skipped
    const int64_t ray = ray_of(a, i);
    if (ray >= min((int64_t)max(*a.n_rays_dev, 0), a.R)) return;
    const int si = a.ray_scene[ray];
    if (si < 0 || si >= a.n_scenes) return;
    g = &a.scenes[si].g;
    h = &a.scenes[si].h;
  } else if (l >= a.h.L) {
    return;
  } else if (a.need && !a.need[i]) {
    ((uint32_t*)a.hfeat)[l * NS + i] = 0u;  // colour unused: no gather
    return;
  }
  float px, py, pz;
  sample_pos(a, i, px, py, pz);
  const float x0 = fminf(fmaxf(over_den(*g, 0, px - g->mn[0]), 0.0f), 1.0f);
  const float x1 = fminf(fmaxf(over_den(*g, 1, py - g->mn[1]), 0.0f), 1.0f);
  const float x2 = fminf(fmaxf(over_den(*g, 2, pz - g->mn[2]), 0.0f), 1.0f);
  float f0, f1;
  hash_level_fwd<true>(*h, l, x0, x1, x2, f0, f1);
  ((uint32_t*)a.hfeat)[l * NS + i] = pk_bf16(f0, f1);
}

// One thread per sample; the hash levels are walked in order so a wavefront handles one
// level of 64 consecutive samples at a time.  Levels below `dedup_levels` merge runs.
__global__ void __launch_bounds__(256) hash_scatter_kernel(RenderDev a, int dedup_levels) {
  const int64_t NS = a.R * (int64_t)a.S;
  const int lane = threadIdx.x & 63;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool in = i < NS;
  float x0 = 0.0f, x1 = 0.0f, x2 = 0.0f;
  if (in) {
    float px, py, pz;
    sample_pos(a, i, px, py, pz);
    x0 = fminf(fmaxf(over_den(a.g, 0, px - a.g.mn[0]), 0.0f), 1.0f);
    x1 = fminf(fmaxf(over_den(a.g, 1, py - a.g.mn[1]), 0.0f), 1.0f);
    x2 = fminf(fmaxf(over_den(a.g, 2, pz - a.g.mn[2]), 0.0f), 1.0f);
  }
  for (int l = 0; l < a.h.L; ++l) {
    float d0 = in ? a.gfeat[(2 * l) * NS + i] : 0.0f;
    float d1 = in ? a.gfeat[(2 * l + 1) * NS + i] : 0.0f;
    const float lw = a.h.lw[l];
    if (lw == 0.0f) continue;
    if (lw < 1.0f) {
      d0 = d0 * lw;
      d1 = d1 * lw;
    }
    const bool live = d0 != 0.0f || d1 != 0.0f;
    if (__ballot(live) == 0) continue;
    const HashCell c = hash_cell(x0, x1, x2, a.h.res[l], a.h.mask);
    float v[16];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      v[2 * k] = d0 * c.w[k];
      v[2 * k + 1] = d1 * c.w[k];
    }
    bool tail = live;
    if (l < dedup_levels) {
      const float res = a.h.res[l];
      const uint64_t cx = (uint64_t)(int)floorf(x0 * res), cy = (uint64_t)(int)floorf(x1 * res),
                     cz = (uint64_t)(int)floorf(x2 * res);
      const uint64_t key = live ? ((cx << 42) | (cy << 21) | cz) : (~0ull - (uint64_t)lane);
      tail = seg_sum<16>(key, v, lane) && live;
    }
    if (__ballot(tail) == 0) continue;
    // Lane quad (4m .. 4m+3) adds, for sample m (then m+16, m+32, m+48), both features of
    // the corner pair that differs only in x: x and x+1 change only the low bits of the
    // hash unless x ends in binary 111, so the two entries usually share a 64-byte segment
    // and the four adjacent floats merge into one memory request.
    float* base = a.grad_table + (int64_t)l * a.h.T * 2;
    const int f = lane & 1, xs = (lane >> 1) & 1;
#pragma unroll
    for (int pass = 0; pass < 4; ++pass) {
      const int src = (lane >> 2) + 16 * pass;
      const bool st = __shfl((int)tail, src, 64) != 0;
      if (__ballot(st) == 0) continue;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float v00 = __shfl(v[2 * k], src, 64), v01 = __shfl(v[2 * k + 1], src, 64);
        const float v10 = __shfl(v[2 * (k + 4)], src, 64), v11 = __shfl(v[2 * (k + 4) + 1], src, 64);
        const uint32_t i0 = (uint32_t)__shfl((int)c.idx[k], src, 64);
        const uint32_t i1 = (uint32_t)__shfl((int)c.idx[k + 4], src, 64);
        const float val = xs ? (f ? v11 : v10) : (f ? v01 : v00);
        if (st) atomicAdd(base + (int64_t)(xs ? i1 : i0) * 2 + f, val);
      }
    }
  }
}

// ------------------------------------------------------------------------------------
// Binned hash-table gradient.  Device-scope float atomics sustain about 21 G requests/s
// however they are spread, which bounds hash_scatter_kernel.  Here the same records are
// binned by table region instead of added in place: (1) each wavefront counts, for one
// level and 512 consecutive samples, the records it produces per bin of 2048 consecutive
// entries (LDS counters of its own), one row of counts per wavefront; (2) a column scan
// over [level][row][bin] turns the counts into each row's first position in every bin;
// (3) the wavefronts write their records there; (4) four wavefronts per bin, each owning a
// quarter of its entries, sum the bin's records in LDS and add the sums to grad_table
// without float atomics.  A record holds the
// two corners of a cell that differ only in x (their entries differ by x ^ (x + 1) in the
// low bits, so they share a bin unless x ends in eleven 1 bits; then each corner gets a
// record of its own, its second half adding zero to the same entry): two entries within
// the bin and both features' weighted gradients of each.  Per live sample and level that
// is 4 records, with runs of consecutive samples in one cell summed first on the levels
// below `dedup_levels`, as in the atomic scatter.  Above them, where runs are short, a
// record is factored into one 16-byte store: the entries and flags in the first word, then
// x's cell fraction fx and the two features' gradients times the pair's y-z weight (the x
// corner takes (1 - fx) of them, the x + 1 corner fx), so those records need no separate
// entry array.  Every position comes from a tile's count row, a scan and LDS counters that
// only one wavefront touches, so the records' order in a bin, and the gradient, are the
// same on every run.
// ------------------------------------------------------------------------------------
// The live samples' indices in order: per block of kLiveBlock samples its live count, then
// each block's indices at the sum of the counts before it (the block adds them up itself;
// the last block writes the total).  A sample is listed where bit `bit` of its flag byte
// is set; masked to that bit, a 4-byte word's popcount counts its listed samples.
constexpr int kLiveBlock = 4096;  // 256 threads x 16 samples

VXN_DEV int live_count16(const uint8_t* __restrict__ live, int64_t s, int64_t NS, int bit, uint32_t (&w)[4]) {
  if (s + 16 <= NS) {
    const uint4 v = *reinterpret_cast<const uint4*>(live + s);
    w[0] = v.x;
    w[1] = v.y;
    w[2] = v.z;
    w[3] = v.w;
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t x = 0u;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (s + 4 * q + j < NS) x |= (uint32_t)live[s + 4 * q + j] << (8 * j);
      w[q] = x;
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) w[q] = (w[q] >> bit) & 0x01010101u;
  return __popc(w[0]) + __popc(w[1]) + __popc(w[2]) + __popc(w[3]);
}

// Exclusive scan of one value per thread over a 256-thread work-group; returns the total.
VXN_DEV int block_scan256(int v, int& excl, int* ws4) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) ws4[wid] = inc;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    pre += w < wid ? ws4[w] : 0;
    tot += ws4[w];
  }
  excl = pre + inc - v;
  return tot;
}

__global__ void __launch_bounds__(256) live_count_kernel(const uint8_t* __restrict__ live, int64_t NS, int bit,
                                                         int32_t* __restrict__ bcnt) {
  __shared__ int ws4[4];
  uint32_t w[4];
  const int c = live_count16(live, (int64_t)blockIdx.x * kLiveBlock + threadIdx.x * 16, NS, bit, w);
  int excl;
  const int tot = block_scan256(c, excl, ws4);
  if (threadIdx.x == 0) bcnt[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(256) live_list_kernel(const uint8_t* __restrict__ live, int64_t NS, int bit,
                                                        const int32_t* __restrict__ bcnt, int32_t* __restrict__ idx,
                                                        int32_t* __restrict__ n_live) {
  __shared__ int ws4[4];
  __shared__ int red[4];
  const int b = blockIdx.x;
  // this block's offset: the counts of the blocks before it
  int part = 0;
  for (int j = threadIdx.x; j < b; j += 256) part += bcnt[j];
  int e0;
  const int off = block_scan256(part, e0, red);
  uint32_t w[4];
  const int64_t s = (int64_t)b * kLiveBlock + threadIdx.x * 16;
  const int c = live_count16(live, s, NS, bit, w);
  int excl;
  const int tot = block_scan256(c, excl, ws4);
  int pos = off + excl;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if ((w[q] >> (8 * j)) & 0xFFu) idx[pos++] = (int32_t)(s + 4 * q + j);
  if (b == (int)gridDim.x - 1 && threadIdx.x == 0) *n_live = off + tot;
}

constexpr int kBinBits = 11;
constexpr int kBinEntries = 1 << kBinBits;
constexpr int kBinTileWaves = 4;            // wavefronts per work-group of the count / write passes
constexpr int kBinWaveChunks = 8;           // 64-sample chunks per wavefront
constexpr int kBinWaveSamples = 64 * kBinWaveChunks;
constexpr int kBinTile = kBinWaveSamples * kBinTileWaves;  // 2048 samples per work-group
constexpr int kMaxBinsPerLevel = 1024;      // tables up to 2^21 entries per level

struct BinDev {
  int nbl;          // bins per level: ceil(T / kBinEntries)
  int nbins;        // L * nbl, in (level, bin) order
  int ntile;        // sample tiles of kBinTile: work-groups per level of passes 1 and 3
  int nrows;        // count rows per level: one per tile (ntile)
  int nchunk;       // row chunks of the column scan
  int32_t* cnt;     // [L][nrows][nbl]: each tile's record count per bin, then (scanned in
                    // place) the position of its first record in each bin
  int32_t* csum;    // [L][nchunk][nbl]: column sums of row chunks, then their exclusive scan
  int32_t* base;    // [nbins + 1]: first record of each bin; base[nbins] = all records
  uint32_t* ridx;   // [records] entries within the bin: x corner | (x + 1 corner) << 16
  float4* rval;     // [records] dL/dfeature 0, 1 of the x corner, then of the x + 1 corner
  int fact_level;   // first level with factored records: rval only, {bits, fx, g0, g1} with
                    // bits = x corner | (x + 1 corner) << 11 | kFactFirst/kFactSecond
};
constexpr uint32_t kFactFirst = 1u << 22;   // split pair: only the x corner, (1 - fx) g
constexpr uint32_t kFactSecond = 2u << 22;  // split pair: only the x + 1 corner, fx g

// Level l of the 64 consecutive samples a wavefront holds (lane = sample i, the walk's
// samples in order): calls emit(tail, cell, v, d0, d1) with every lane of the wavefront
// when some lane is live: tail marks the lanes that emit records (the last lane of each
// run of samples in one cell below `dedup_levels`, every live sample above), v their 8
// corners' weighted gradients (run sums) and d0, d1 the level's feature gradients of the
// lane's own sample.  Walking the render's list of samples with a gradient, every listed
// sample is live; without the list a sample is live where its two gradients are not zero.
// Either way the count and write passes apply the same rule, so they agree.
// VALUES = false leaves v unset: only the tails are needed.
// The inputs of sample i at level l: its two feature gradients, mid-point depth and ray
// (zero past the last sample), loaded ahead of their use by the write pass.
struct LevelIn {
  float d0, d1, m, ox, oy, oz, rx, ry, rz;
};

template <bool GRAD = true>
VXN_DEV LevelIn level_in(const RenderDev& a, int64_t i, int l) {
  const int64_t NS = a.R * (int64_t)a.S;
  LevelIn x{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (i < NS) {
    const int64_t ray = ray_of(a, i);
    if constexpr (GRAD) {
      x.d0 = a.gfeat[(2 * l) * NS + i];
      x.d1 = a.gfeat[(2 * l + 1) * NS + i];
    }
    x.m = a.mid[i];
    x.ox = a.o[3 * ray];
    x.oy = a.o[3 * ray + 1];
    x.oz = a.o[3 * ray + 2];
    x.rx = a.d[3 * ray];
    x.ry = a.d[3 * ray + 1];
    x.rz = a.d[3 * ray + 2];
  }
  return x;
}

// hash_level_eval returns false (and sets nothing) when no lane of the wavefront emits.
template <bool VALUES>
VXN_DEV bool hash_level_eval(const RenderDev& a, int64_t i, const LevelIn& x, int lane, int l, int dedup_levels,
                             HashCell& c, float (&v)[16], float& d0, float& d1, bool& tail) {
  const int64_t NS = a.R * (int64_t)a.S;
  const bool in = i < NS;
  const float lw = a.h.lw[l];
  if (lw == 0.0f) return false;
  d0 = x.d0;
  d1 = x.d1;
  if (lw < 1.0f) {
    d0 = d0 * lw;
    d1 = d1 * lw;
  }
  // walking the live list, every listed sample emits its records (its gradients can still
  // be zero at a level: such records add zeros), so the count pass reads no gradients
  const bool live = a.live_idx ? in : (d0 != 0.0f || d1 != 0.0f);
  if (__ballot(live) == 0) return false;
  float x0 = 0.0f, x1 = 0.0f, x2 = 0.0f;
  if (in) {
    // sample_pos's arithmetic on the loaded ray
    const float px = x.ox + x.rx * x.m, py = x.oy + x.ry * x.m, pz = x.oz + x.rz * x.m;
    x0 = fminf(fmaxf(over_den(a.g, 0, px - a.g.mn[0]), 0.0f), 1.0f);
    x1 = fminf(fmaxf(over_den(a.g, 1, py - a.g.mn[1]), 0.0f), 1.0f);
    x2 = fminf(fmaxf(over_den(a.g, 2, pz - a.g.mn[2]), 0.0f), 1.0f);
  }
  c = hash_cell(x0, x1, x2, a.h.res[l], a.h.mask);
  if constexpr (VALUES) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      v[2 * k] = d0 * c.w[k];
      v[2 * k + 1] = d1 * c.w[k];
    }
  }
  tail = live;
  if (l < dedup_levels) {
    const float res = a.h.res[l];
    const uint64_t cx = (uint64_t)(int)floorf(x0 * res), cy = (uint64_t)(int)floorf(x1 * res),
                   cz = (uint64_t)(int)floorf(x2 * res);
    const uint64_t key = live ? ((cx << 42) | (cy << 21) | cz) : (~0ull - (uint64_t)lane);
    if constexpr (VALUES) {
      tail = seg_sum<16>(key, v, lane) && live;
    } else {
      const uint64_t kn = __shfl_down(key, 1, 64);
      tail = (lane == 63 || kn != key) && live;
    }
  }
  return true;
}

template <bool VALUES, class F>
VXN_DEV void hash_level_records(const RenderDev& a, int64_t i, int lane, int l, int dedup_levels, F&& emit) {
  HashCell c;
  float v[16];
  float d0, d1;
  bool tail;
  const LevelIn x = (VALUES || !a.live_idx) ? level_in<true>(a, i, l) : level_in<false>(a, i, l);
  if (hash_level_eval<VALUES>(a, i, x, lane, l, dedup_levels, c, v, d0, d1, tail))
    emit(tail, c, v, d0, d1);
}

// Passes 1 and 3 run one work-group per (level, 2048-sample tile), level-major: its four
// wavefronts walk 512 consecutive samples each, in 8 chunks of 64, and the tile owns one
// row of counters (LDS, then cnt[level][tile]).  Pass 3 takes the chunks in lockstep: the
// records of chunk c of all four wavefronts (at most kStageSlots) are ranked in each bin,
// sorted by bin through LDS, and stored as runs that continue the tile's previous run in
// that bin, so consecutive lanes store to consecutive slots.  A record's place in its bin
// follows (chunk, wavefront, lane, corner pair): ranks come from LDS atomics on a counter
// row that only the wavefront's own instructions touch, offsets from a scan, so every
// position - and the order in which pass 4 adds the records - is the same on every run.
// Staged records per chunk: 4 wavefronts x 64 lanes x 4 pairs; the rare split pairs beyond
// that are stored directly.
constexpr int kStageSlots = kBinTileWaves * 64 * 4;

__global__ void __launch_bounds__(256) hash_bin_count_kernel(RenderDev a, BinDev b, int dedup_levels) {
  extern __shared__ int32_t cnt_lds[];  // [nbl]
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int l = (int)(blockIdx.x / (unsigned)b.ntile), tile = (int)(blockIdx.x - (unsigned)l * b.ntile);
  for (int k = threadIdx.x; k < b.nbl; k += 256) cnt_lds[k] = 0;
  __syncthreads();
  const int64_t i0 = (int64_t)tile * kBinTile + (int64_t)wid * kBinWaveSamples;
  const int64_t nw = walk_length(a);
  for (int c = 0; c < kBinWaveChunks; ++c) {
    if (i0 + c * 64 >= nw) break;
    hash_level_records<false>(a, walk_sample(a, i0 + c * 64 + lane, nw), lane, l, dedup_levels,
                              [&](bool tail, const HashCell& cell, const float(&)[16], float, float) {
      if (tail) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int ba = (int)(cell.idx[q] >> kBinBits), bb = (int)(cell.idx[q + 4] >> kBinBits);
          atomicAdd(&cnt_lds[ba], 1);
          if (bb != ba) atomicAdd(&cnt_lds[bb], 1);
        }
      }
    });
  }
  __syncthreads();
  int32_t* out = b.cnt + ((int64_t)l * b.nrows + tile) * b.nbl;
  for (int k = threadIdx.x; k < b.nbl; k += 256) out[k] = cnt_lds[k];
}

// A work-group barrier that orders LDS only: loads from global memory issued before it stay
// in flight (the write pass prefetches the next chunk's inputs across its barriers).
VXN_DEV void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// FACT: the launch covers levels l0.. of factored records (no run sums, no entry array).
template <bool FACT>
__global__ void __launch_bounds__(256) hash_bin_write_kernel(RenderDev a, BinDev b, int dedup_levels, int l0) {
  extern __shared__ float4 wlds[];
  float4* stage = wlds;                                          // [kStageSlots] sorted records
  uint32_t* sridx = (uint32_t*)(stage + kStageSlots);            // [kStageSlots] their entry words
  uint16_t* sbin = (uint16_t*)(sridx + kStageSlots);             // [kStageSlots] their bins
  int32_t* wcnt = (int32_t*)(sbin + kStageSlots);                // [kBinTileWaves][nbl] ranks, then offsets
  int32_t* gpos = wcnt + kBinTileWaves * b.nbl;                  // [nbl] the tile's next slot in each bin
  int32_t* gb = gpos + b.nbl;                                    // [nbl] slot - stage index, this chunk
  int32_t* wsum = gb + b.nbl;                                    // [4] scan of the wavefront totals
  const int t = threadIdx.x, wid = t >> 6, lane = t & 63;
  const int nbl = b.nbl;
  const int lt = (int)(blockIdx.x / (unsigned)b.ntile), tile = (int)(blockIdx.x - (unsigned)lt * b.ntile);
  const int l = l0 + lt;
  const int32_t* row = b.cnt + ((int64_t)l * b.nrows + tile) * nbl;
  for (int k = t; k < nbl; k += 256) gpos[k] = row[k];
  constexpr bool fact = FACT;
  int32_t* mine = wcnt + wid * nbl;
  const int per = (nbl + 255) / 256;  // bins per thread in the scan (consecutive)
  const int64_t i0 = (int64_t)tile * kBinTile + (int64_t)wid * kBinWaveSamples + lane;
  constexpr uint32_t m = kBinEntries - 1;
  // the walk's positions i0 + 64 c; a tile past the walk's end has no records
  const int64_t nw = walk_length(a);
  if ((int64_t)tile * kBinTile >= nw) return;
  int64_t inxt = walk_sample(a, i0, nw);
  LevelIn nxt = level_in(a, inxt, l);
  for (int c = 0; c < kBinWaveChunks; ++c) {
    const LevelIn cur = nxt;
    const int64_t icur = inxt;
    if (c + 1 < kBinWaveChunks) {
      inxt = walk_sample(a, i0 + (c + 1) * 64, nw);
      nxt = level_in(a, inxt, l);
    }
    for (int k = lane; k < nbl; k += 64) mine[k] = 0;
    wave_sync();
    // plain scalar arrays: the HIP vector types would keep these in scratch memory
    float rvx[4], rvy[4], rvz[4], rvw[4], rwx[4], rwy[4];
    uint32_t rx[4], rx2[4];
    int rb[4], rb2[4], rk[4], rk2[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      rb[q] = -1;
      rb2[q] = -1;
    }
    HashCell cell;
    float v[16];
    float d0, d1;
    bool tail;
    if (hash_level_eval<!FACT>(a, icur, cur, lane, l, dedup_levels, cell, v, d0, d1, tail) && tail) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t ia = cell.idx[q], ib = cell.idx[q + 4];
        const int ba = (int)(ia >> kBinBits), bb = (int)(ib >> kBinBits);
        rb[q] = ba;
        rk[q] = atomicAdd(&mine[ba], 1);
        rx2[q] = 0u;
        rwx[q] = 0.0f;
        rwy[q] = 0.0f;
        if constexpr (FACT) {
          const float wy = (q & 2) ? cell.dy : 1.0f - cell.dy;
          const float wz = (q & 1) ? cell.dz : 1.0f - cell.dz;
          const float wyz = wy * wz;
          const uint32_t bits = bb == ba ? ((ia & m) | ((ib & m) << 11)) : ((ia & m) * 0x801u | kFactFirst);
          rvx[q] = __uint_as_float(bits);
          rvy[q] = cell.dx;
          rvz[q] = d0 * wyz;
          rvw[q] = d1 * wyz;
          rx[q] = 0u;
          rx2[q] = (ib & m) * 0x801u | kFactSecond;
        } else if (bb == ba) {
          rvx[q] = v[2 * q];
          rvy[q] = v[2 * q + 1];
          rvz[q] = v[2 * q + 8];
          rvw[q] = v[2 * q + 9];
          rx[q] = (ia & m) | ((ib & m) << 16);
        } else {
          rvx[q] = v[2 * q];
          rvy[q] = v[2 * q + 1];
          rvz[q] = 0.0f;
          rvw[q] = 0.0f;
          rx[q] = (ia & m) * 0x10001u;
          rx2[q] = (ib & m) * 0x10001u;
          rwx[q] = v[2 * q + 8];
          rwy[q] = v[2 * q + 9];
        }
        if (bb != ba) {
          rb2[q] = bb;
          rk2[q] = atomicAdd(&mine[bb], 1);
        }
      }
    }
    lds_barrier();
    // offsets: bins in order, then wavefronts in order within a bin
    int32_t tot = 0;
    for (int j = 0; j < per; ++j) {
      const int k = t * per + j;
      if (k < nbl)
#pragma unroll
        for (int w = 0; w < kBinTileWaves; ++w) tot += wcnt[w * nbl + k];
    }
    int32_t inc = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    if (lane == 63) wsum[wid] = inc;
    lds_barrier();
    int32_t run = inc - tot;
    for (int w = 0; w < wid; ++w) run += wsum[w];
    const int32_t ntot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    for (int j = 0; j < per; ++j) {
      const int k = t * per + j;
      if (k < nbl) {
        const int32_t g = gpos[k];
        gb[k] = g - run;
        int32_t o = run;
#pragma unroll
        for (int w = 0; w < kBinTileWaves; ++w) {
          const int32_t n = wcnt[w * nbl + k];
          wcnt[w * nbl + k] = o;
          o += n;
        }
        gpos[k] = g + (o - run);
        run = o;
      }
    }
    lds_barrier();
    // records past the stage's capacity (many split pairs in one chunk) go straight out
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (rb[q] >= 0) {
        const int slot = mine[rb[q]] + rk[q];
        const float4 rec = make_float4(rvx[q], rvy[q], rvz[q], rvw[q]);
        if (slot < kStageSlots) {
          stage[slot] = rec;
          sridx[slot] = rx[q];
          sbin[slot] = (uint16_t)rb[q];
        } else {
          const int32_t dst = gb[rb[q]] + slot;
          b.rval[dst] = rec;
          if (!fact) b.ridx[dst] = rx[q];
        }
      }
      if (rb2[q] >= 0) {
        const int slot = mine[rb2[q]] + rk2[q];
        const float4 rec = fact ? make_float4(__uint_as_float(rx2[q]), rvy[q], rvz[q], rvw[q])
                                : make_float4(rwx[q], rwy[q], 0.0f, 0.0f);
        if (slot < kStageSlots) {
          stage[slot] = rec;
          sridx[slot] = rx2[q];
          sbin[slot] = (uint16_t)rb2[q];
        } else {
          const int32_t dst = gb[rb2[q]] + slot;
          b.rval[dst] = rec;
          if (!fact) b.ridx[dst] = rx2[q];
        }
      }
    }
    lds_barrier();
    // a fixed number of store slots per thread: the next chunk's prefetched loads, issued
    // before these stores, are then waited for without waiting for the stores themselves
    const int nst = ntot < kStageSlots ? ntot : kStageSlots;
#pragma unroll
    for (int jj = 0; jj < kStageSlots / 256; ++jj) {
      const int j = t + jj * 256;
      if (j < nst) {
        const int32_t dst = gb[sbin[j]] + j;
        b.rval[dst] = stage[j];
        if (!fact) b.ridx[dst] = sridx[j];
      }
    }
  }
}

// LDS of the write pass for nbl bins per level.
static size_t bin_write_lds(int nbl) {
  return (size_t)kStageSlots * (sizeof(float4) + sizeof(uint32_t) + sizeof(uint16_t)) +
         ((size_t)kBinTileWaves * nbl + 2 * (size_t)nbl + 8) * sizeof(int32_t);
}

// Pass 2, the scan that turns the count rows into positions: a bin's records are stored in
// row (tile) order, bins in (level, bin) order.
// Column chunk sums, an exclusive scan of each column's chunks (one thread per column),
// one work-group's scan of the column totals (base), then each chunk writes its rows'
// positions in place.
constexpr int kColChunk = 64;    // rows per chunk
constexpr int kColThreads = 1024;

constexpr int kColBatch = 16;  // rows loaded ahead

__global__ void __launch_bounds__(256) bin_col_sums_kernel(BinDev b) {
  const int bin = blockIdx.x * 256 + threadIdx.x, ch = blockIdx.y, l = blockIdx.z;
  if (bin >= b.nbl) return;
  const int r1 = min(b.nrows, (ch + 1) * kColChunk);
  const int32_t* c = b.cnt + (int64_t)l * b.nrows * b.nbl + bin;
  int32_t s = 0;
  for (int r0 = ch * kColChunk; r0 < r1; r0 += kColBatch) {
    int32_t v[kColBatch];
#pragma unroll
    for (int j = 0; j < kColBatch; ++j) v[j] = r0 + j < r1 ? c[(int64_t)(r0 + j) * b.nbl] : 0;
#pragma unroll
    for (int j = 0; j < kColBatch; ++j) s += v[j];
  }
  b.csum[((int64_t)l * b.nchunk + ch) * b.nbl + bin] = s;
}

// One thread per column (level, bin): exclusive scan of the column's chunk sums in place,
// the column total into base[column].
__global__ void __launch_bounds__(256) bin_col_scan_kernel(BinDev b) {
  const int col = blockIdx.x * 256 + threadIdx.x;
  if (col >= b.nbins) return;
  const int l = col / b.nbl, bin = col - l * b.nbl;
  int32_t* cs = b.csum + (int64_t)l * b.nchunk * b.nbl + bin;
  int32_t run = 0;
  for (int c0 = 0; c0 < b.nchunk; c0 += kColBatch) {
    int32_t v[kColBatch];
#pragma unroll
    for (int j = 0; j < kColBatch; ++j) v[j] = c0 + j < b.nchunk ? cs[(int64_t)(c0 + j) * b.nbl] : 0;
#pragma unroll
    for (int j = 0; j < kColBatch; ++j) {
      if (c0 + j < b.nchunk) cs[(int64_t)(c0 + j) * b.nbl] = run;
      run += v[j];
    }
  }
  b.base[col] = run;
}

// One work-group: exclusive scan of the column totals in base[] in place (thread t takes
// `per` consecutive columns), the grand total at base[nbins].
__global__ void __launch_bounds__(kColThreads) bin_base_scan_kernel(BinDev b) {
  __shared__ int32_t ws[kColThreads / 64];
  const int per = (b.nbins + kColThreads - 1) / kColThreads;
  const int c0 = threadIdx.x * per;
  int32_t tot = 0;
  for (int k = 0; k < per; ++k)
    if (c0 + k < b.nbins) tot += b.base[c0 + k];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int32_t inc = tot;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) ws[wid] = inc;
  __syncthreads();
  int32_t pre = 0, all = 0;
  for (int w = 0; w < kColThreads / 64; ++w) {
    if (w < wid) pre += ws[w];
    all += ws[w];
  }
  int32_t run = pre + inc - tot;
  for (int k = 0; k < per; ++k) {
    if (c0 + k >= b.nbins) break;
    const int32_t t = b.base[c0 + k];
    b.base[c0 + k] = run;
    run += t;
  }
  if (threadIdx.x == 0) b.base[b.nbins] = all;
}

__global__ void __launch_bounds__(256) bin_col_positions_kernel(BinDev b) {
  const int bin = blockIdx.x * 256 + threadIdx.x, ch = blockIdx.y, l = blockIdx.z;
  if (bin >= b.nbl) return;
  const int r1 = min(b.nrows, (ch + 1) * kColChunk);
  int32_t* c = b.cnt + (int64_t)l * b.nrows * b.nbl + bin;
  int32_t run = b.base[l * b.nbl + bin] + b.csum[((int64_t)l * b.nchunk + ch) * b.nbl + bin];
  for (int r0 = ch * kColChunk; r0 < r1; r0 += kColBatch) {
    int32_t v[kColBatch];
#pragma unroll
    for (int j = 0; j < kColBatch; ++j) v[j] = r0 + j < r1 ? c[(int64_t)(r0 + j) * b.nbl] : 0;
#pragma unroll
    for (int j = 0; j < kColBatch; ++j) {
      if (r0 + j < r1) c[(int64_t)(r0 + j) * b.nbl] = run;
      run += v[j];
    }
  }
}

// Pass 4: one work-group of kSumWaves wavefronts per bin.  Wavefront w reads the bin's
// 64-record groups w, w + kSumWaves, ... (each record is read once) and adds them into its
// own copy of the bin's 2048 entries in LDS; the copies are then added in wavefront order,
// so the sum is the same on every run.  LDS float atomics run at a few cycles per lane, so a
// copy is updated with plain LDS read-add-write instead: in each round every pending
// half-record writes its lane id to its entry's tag, and a lane adds a half whose tag reads
// back its own id (one lane per entry per round; a lane whose two halves share the entry
// adds both, in order); the others retry in the next round.  A single wavefront owns a copy
// and its LDS instructions complete in order, so no update is lost.  A lane with nothing to
// add in a round reads and writes its own dummy slot (entry kBinEntries + lane), so a round
// is straight-line LDS code.  (Four groups' halves tagged together in one round, lane + 64
// x group, took more rounds than they saved.)  The bins run on XCD bin % 8 (work-group b on
// XCD b % 8).
constexpr int kSumWaves = 2;  // two copies per work-group: more, smaller work-groups per CU
constexpr int kSumSlots = kBinEntries + 64;

__global__ void __launch_bounds__(64 * kSumWaves) hash_bin_sum_kernel(BinDev b, float* __restrict__ grad_table,
                                                                     int64_t T, int overwrite) {
  extern __shared__ float2 sacc[];  // [kSumWaves][kSumSlots], then tags [kSumWaves][kSumSlots]
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int bin = blockIdx.x;
  if (bin >= b.nbins) return;
  float2* acc = sacc + wid * kSumSlots;
  uint8_t* tag = (uint8_t*)(sacc + kSumWaves * kSumSlots) + wid * kSumSlots;
  const int l = bin / b.nbl;
  const int64_t base = (int64_t)(bin - l * b.nbl) << kBinBits;
  const int ne = (int)max((int64_t)0, min((int64_t)kBinEntries, T - base));
  for (int e = lane; e < ne; e += 64) acc[e] = make_float2(0.0f, 0.0f);
  const int dummy = kBinEntries + lane;
  wave_sync();
  const int32_t r0 = b.base[bin], r1 = b.base[bin + 1];
  const bool fact = l >= b.fact_level;
  // U 64-record groups per iteration; the next iteration's groups are loaded before this
  // one's are added, so the loads of a wavefront stay in flight while it works in LDS.
  // Indices past the bin's end are clamped onto its last record (always a valid address
  // inside the loop) and that record's use is masked.
  constexpr int U = 2;
  constexpr int32_t kStep = U * 64 * kSumWaves;
  float4 nv[U];
  uint32_t nr[U];
  auto fetch = [&](int32_t k0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t kk = min(k0 + u * 64 * kSumWaves + lane, r1 - 1);
      nv[u] = b.rval[kk];
      nr[u] = fact ? 0u : b.ridx[kk];
    }
  };
  if (r0 + wid * 64 < r1) fetch(r0 + wid * 64);
  for (int32_t k0 = r0 + wid * 64; k0 < r1; k0 += kStep) {
    uint32_t r[U];
    float4 v[U];
    float4 cv[U];
    uint32_t cr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cv[u] = nv[u];
      cr[u] = nr[u];
    }
    if (k0 + kStep < r1) fetch(k0 + kStep);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (fact) {
        const float4 f = cv[u];
        const uint32_t bits = __float_as_uint(f.x);
        const float ox = 1.0f - f.y;
        const float a0 = ox * f.z, a1 = ox * f.w, c0 = f.y * f.z, c1 = f.y * f.w;
        r[u] = (bits & 0x7FFu) | (((bits >> 11) & 0x7FFu) << 16);
        const uint32_t fl = bits & (kFactFirst | kFactSecond);
        v[u] = fl == 0u ? make_float4(a0, a1, c0, c1)
                        : (fl == kFactFirst ? make_float4(a0, a1, 0.0f, 0.0f) : make_float4(c0, c1, 0.0f, 0.0f));
      } else {
        r[u] = cr[u];
        v[u] = cv[u];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e0 = (int)(r[u] & 0xFFFFu), e1 = (int)(r[u] >> 16);
      const bool ok = k0 + u * 64 * kSumWaves + lane < r1;
      bool p0 = ok, p1 = ok;
      while (__ballot(p0 || p1) != 0) {
        tag[p0 ? e0 : dummy] = (uint8_t)lane;
        tag[p1 ? e1 : dummy] = (uint8_t)lane;
        wave_sync();
        const int t0 = tag[p0 ? e0 : dummy], t1 = tag[p1 ? e1 : dummy];
        const bool w0 = p0 && t0 == lane, w1 = p1 && t1 == lane;
        // both halves on one entry: one read-add-write with the first half added first
        const bool same = w0 && w1 && e0 == e1;
        const int a0 = w0 ? e0 : dummy, a1 = (w1 && !same) ? e1 : dummy;
        float2 o0 = acc[a0], o1 = acc[a1];
        o0.x = o0.x + v[u].x;
        o0.y = o0.y + v[u].y;
        if (same) {
          o0.x = o0.x + v[u].z;
          o0.y = o0.y + v[u].w;
        }
        o1.x = o1.x + v[u].z;
        o1.y = o1.y + v[u].w;
        acc[a0] = o0;
        acc[a1] = o1;
        p0 = p0 && !w0;
        p1 = p1 && !w1;
        wave_sync();
      }
    }
  }
  __syncthreads();
  float2* g = (float2*)(grad_table + ((int64_t)l * T + base) * 2);
  for (int e = threadIdx.x; e < ne; e += 64 * kSumWaves) {
    float2 sum = sacc[e];
#pragma unroll
    for (int w = 1; w < kSumWaves; ++w) {
      const float2 o = sacc[w * kSumSlots + e];
      sum.x = sum.x + o.x;
      sum.y = sum.y + o.y;
    }
    if (overwrite) {
      g[e] = sum;
    } else if (sum.x != 0.0f || sum.y != 0.0f) {
      float2 o = g[e];
      o.x = o.x + sum.x;
      o.y = o.y + sum.y;
      g[e] = o;
    }
  }
}

constexpr size_t kSumLds = (size_t)kSumWaves * kSumSlots * (sizeof(float2) + 1);

// Brick slots of the Eq. 7 stencil around a cell with base corner c0: 8 cell corners
// (ox,oy,oz) then the 24 face neighbours at offset -1 / +2 along one axis.
VXN_DEV int slot_corner(int ox, int oy, int oz) { return 4 * ox + 2 * oy + oz; }
VXN_DEV int slot_face(int axis, int side, int u, int v) { return 8 + axis * 8 + side * 4 + 2 * u + v; }

// Inputs of the grid gradient for sample i (zero past the last sample), loaded ahead of
// their use by the binned passes.
struct GridIn {
  float dsdf, dg0, dg1, dg2, px, py, pz;
};

VXN_DEV GridIn grid_in(const RenderDev& a, int64_t i) {
  const int64_t NS = a.R * (int64_t)a.S;
  GridIn x{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (i < NS) {
    x.dsdf = a.ggeo[i];
    x.dg0 = a.ggeo[NS + i];
    x.dg1 = a.ggeo[2 * NS + i];
    x.dg2 = a.ggeo[3 * NS + i];
    sample_pos(a, i, x.px, x.py, x.pz);
  }
  return x;
}

// The brick of one sample (called by every lane of the wavefront): when its whole Eq. 7
// stencil (Eq. 4: its 8 cell corners) lies inside the grid, v gets the 32 stencil weights
// of the brick around cell c0 (slot_corner / slot_face order; the grid_sample weights are
// folded in when grid_sample picks the same cell) and the call returns true.  Stencils at
// the border and grid_sample cells other than the brick's are added to grad_grid here with
// atomics when EDGES (they are rare).  key is the cell (a lane-unique value without a brick).
// EDGES: the border stencils and the grid_sample stencils outside the brick are added
// with device atomics here (the atomic scatter, and the binned passes' ordered edge pass);
// without it they are only reported through `edge`.
template <bool VALUES, bool EDGES>
VXN_DEV bool grid_brick(const RenderDev& a, const GridIn& x, int lane, float (&v)[32], uint64_t& key, int (&c0)[3],
                        bool& edge) {
  const float dsdf = x.dsdf;
  const float dg[3] = {x.dg0, x.dg1, x.dg2};
  const bool live = dsdf != 0.0f || dg[0] != 0.0f || dg[1] != 0.0f || dg[2] != 0.0f;
  key = ~0ull - (uint64_t)lane;
  c0[0] = c0[1] = c0[2] = 0;
  edge = false;
  if (VALUES) {
#pragma unroll
    for (int k = 0; k < 32; ++k) v[k] = 0.0f;
  }
  if (__ballot(live) == 0) return false;
  const Corners cn = corner_decomp(a.g, x.px, x.py, x.pz);
  const GsStencil gs = gs_stencil(a.g, x.px, x.py, x.pz);
  const GridDesc& g = a.g;
  const bool ana = g.analytical != 0;
  // Eq. 7 bricks: the whole 4x4x4 brick is inside the grid and the divisors are 2*eps.
  // Eq. 4 touches only the 8 cell corners, distinct whenever c1 = c0 + 1 on every axis.
  const bool interior =
      ana ? (cn.c1[0] == cn.c0[0] + 1 && cn.c1[1] == cn.c0[1] + 1 && cn.c1[2] == cn.c0[2] + 1)
          : (cn.c0[0] >= 1 && cn.c0[0] + 2 <= g.X - 1 && cn.c0[1] >= 1 && cn.c0[1] + 2 <= g.Y - 1 &&
             cn.c0[2] >= 1 && cn.c0[2] + 2 <= g.Z - 1);
  const bool gs_same = gs.x0 == cn.c0[0] && gs.y0 == cn.c0[1] && gs.z0 == cn.c0[2] && gs.valid == 0xffu;
  // grid_sample's cell one step from the corner cell along one axis (its floor rounds the
  // other way near a cell face): its 8 corners are brick slots (cell corners and one face),
  // so an Eq. 7 brick takes them too
  const int gdx = gs.x0 - cn.c0[0], gdy = gs.y0 - cn.c0[1], gdz = gs.z0 - cn.c0[2];
  const bool fold = !ana && !gs_same && gs.valid == 0xffu &&
                    ((gdx * gdx == 1 && gdy == 0 && gdz == 0) || (gdx == 0 && gdy * gdy == 1 && gdz == 0) ||
                     (gdx == 0 && gdy == 0 && gdz * gdz == 1));
  edge = live && (!interior || (dsdf != 0.0f && !gs_same && !fold));
  if (EDGES && live && !interior) {
    grid_gradient_scatter(g, cn, dg, a.grad_grid);
    if (dsdf != 0.0f) gs_scatter(g, gs, dsdf, a.grad_grid);
  }
  const bool brick = live && interior;
  if (VALUES && brick && ana) {
    float cg[8];
    eq4_corner_grads(g, cn, dg, cg);
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = cg[k];
  }
  if (VALUES && brick && !ana) {
    // interior divisor of every corner: max((i_p - i_m) * eps, eps) = 2 * eps
    const float dvx = fmaxf(2.0f * g.eps[0], g.eps[0]), dvy = fmaxf(2.0f * g.eps[1], g.eps[1]),
                dvz = fmaxf(2.0f * g.eps[2], g.eps[2]);
    const float rx = 0.5f * (1.0f / g.eps[0]), ry = 0.5f * (1.0f / g.eps[1]), rz = 0.5f * (1.0f / g.eps[2]);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int ox = (k >> 2) & 1, oy = (k >> 1) & 1, oz = k & 1;
      const float vx = div_rn(cn.w[k] * dg[0], dvx, rx);
      const float vy = div_rn(cn.w[k] * dg[1], dvy, ry);
      const float vz = div_rn(cn.w[k] * dg[2], dvz, rz);
      // x stencil of corner k: +v at ox+1, -v at ox-1
      v[ox ? slot_face(0, 1, oy, oz) : slot_corner(1, oy, oz)] += vx;
      v[ox ? slot_corner(0, oy, oz) : slot_face(0, 0, oy, oz)] -= vx;
      v[oy ? slot_face(1, 1, ox, oz) : slot_corner(ox, 1, oz)] += vy;
      v[oy ? slot_corner(ox, 0, oz) : slot_face(1, 0, ox, oz)] -= vy;
      v[oz ? slot_face(2, 1, ox, oy) : slot_corner(ox, oy, 1)] += vz;
      v[oz ? slot_corner(ox, oy, 0) : slot_face(2, 0, ox, oy)] -= vz;
    }
  }
  if (brick && dsdf != 0.0f) {
    if (gs_same) {
      if (VALUES) {
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] += gs.w[k] * dsdf;
      }
    } else if (fold) {
      if (VALUES) {
        // corner k = (ox, oy, oz) of the cell shifted by +-1 along x, y or z
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int ox = (k >> 2) & 1, oy = (k >> 1) & 1, oz = k & 1;
          const float t = gs.w[k] * dsdf;
          if (gdx == 1) v[ox ? slot_face(0, 1, oy, oz) : slot_corner(1, oy, oz)] += t;
          if (gdx == -1) v[ox ? slot_corner(0, oy, oz) : slot_face(0, 0, oy, oz)] += t;
          if (gdy == 1) v[oy ? slot_face(1, 1, ox, oz) : slot_corner(ox, 1, oz)] += t;
          if (gdy == -1) v[oy ? slot_corner(ox, 0, oz) : slot_face(1, 0, ox, oz)] += t;
          if (gdz == 1) v[oz ? slot_face(2, 1, ox, oy) : slot_corner(ox, oy, 1)] += t;
          if (gdz == -1) v[oz ? slot_corner(ox, oy, 0) : slot_face(2, 0, ox, oy)] += t;
        }
      }
    } else if (EDGES) {
      gs_scatter(g, gs, dsdf, a.grad_grid);
    }
  }
  if (brick) {
    key = ((uint64_t)cn.c0[0] << 42) | ((uint64_t)cn.c0[1] << 21) | (uint64_t)cn.c0[2];
    c0[0] = cn.c0[0];
    c0[1] = cn.c0[1];
    c0[2] = cn.c0[2];
  }
  return brick;
}

__global__ void __launch_bounds__(256) grid_scatter_kernel(RenderDev a) {
  const int lane = threadIdx.x & 63;
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;  // position of the walk
  if (blockIdx.x * (int64_t)blockDim.x >= walk_length(a)) return;
  const GridIn x = grid_in(a, walk_sample(a, k, walk_length(a)));
  float v[32];
  uint64_t key;
  int c0[3];
  bool edge;
  const bool brick = grid_brick<true, true>(a, x, lane, v, key, c0, edge);
  const GridDesc& g = a.g;
  const bool ana = g.analytical != 0;
  const bool tail = seg_sum<32>(key, v, lane) && brick;
  const uint64_t tb = __ballot(tail);
  if (tb == 0) return;
  // Flush the runs' bricks column by column: the 32 slots form 12 (x, y) columns of
  // contiguous z (4 of length 4 through the cell, 8 of length 2 on the x and y faces).
  // An Eq. 4 brick holds only the cell corners: z = 0, 1 of the through-cell columns.
  // Passes of 16 runs; lane quad q takes the q-th run of the pass and lane 4q + e its
  // e-th z, so each atomic instruction writes whole columns and merges into few requests.
  const int64_t sx = (int64_t)g.Y * g.Z, sy = g.Z;
  const int64_t base_own = brick ? lin3(g, c0[0], c0[1], c0[2]) : 0;
  const int e = lane & 3, q = lane >> 2;
  const int ntail = __popcll(tb);
  for (int pass = 0; 16 * pass < ntail; ++pass) {
    const int want = 16 * pass + q;
    const bool act = want < ntail;
    // lane of the want-th run tail: the want-th set bit of tb
    int lo = 0, hi = 63;
    while (lo < hi) {
      const int md = (lo + hi) >> 1;
      const uint64_t msk = md == 63 ? ~0ull : ((1ull << (md + 1)) - 1ull);
      if (__popcll(tb & msk) >= want + 1) hi = md;
      else lo = md + 1;
    }
    const int src = act ? lo : lane;
    const int64_t base = (int64_t)(((uint64_t)(uint32_t)__shfl((int)(base_own >> 32), src, 64) << 32) |
                                   (uint64_t)(uint32_t)__shfl((int)base_own, src, 64));
    float* gp = a.grad_grid + base;
    // through-cell columns (ox, oy): z = -1, 0, 1, 2
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int ox = c >> 1, oy = c & 1;
      const float s0 = __shfl(v[slot_face(2, 0, ox, oy)], src, 64), s1 = __shfl(v[slot_corner(ox, oy, 0)], src, 64);
      const float s2 = __shfl(v[slot_corner(ox, oy, 1)], src, 64), s3 = __shfl(v[slot_face(2, 1, ox, oy)], src, 64);
      const float val = e == 0 ? s0 : (e == 1 ? s1 : (e == 2 ? s2 : s3));
      if (act && (!ana || e == 1 || e == 2)) atomicAdd(gp + ox * sx + oy * sy + (e - 1), val);
    }
    if (ana) continue;
    // face columns: x = -1 / 2 at y = u, y = -1 / 2 at x = u; z = 0, 1
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int side = c >> 1, u = c & 1;
      const float ax0 = __shfl(v[slot_face(0, side, u, 0)], src, 64), ax1 = __shfl(v[slot_face(0, side, u, 1)], src, 64);
      const float ay0 = __shfl(v[slot_face(1, side, u, 0)], src, 64), ay1 = __shfl(v[slot_face(1, side, u, 1)], src, 64);
      const int64_t off = side ? 2 : -1;
      if (act && e < 2) atomicAdd(gp + off * sx + u * sy + e, e == 0 ? ax0 : ax1);
      if (act && e >= 2) atomicAdd(gp + u * sx + off * sy + (e - 2), e == 2 ? ay0 : ay1);
    }
  }
}

// ------------------------------------------------------------------------------------
// Voxel-binned grid gradient: the bricks of grid_scatter_kernel's runs, sorted by the grid
// tile they touch and summed per tile in LDS, so every vertex of grad_grid is written by one
// plain read-add-write.  The border and grid_sample edge cases above stay atomic, but are
// added by one wavefront in sample order (grid_edge_kernel), so their order is fixed too.
//   1. grid_bin_count_kernel: one wavefront per count row of kGridRowSamples samples (8
//      chunks of 64, the runs of grid_scatter_kernel): per tile, the row's (run, tile) pairs.
//   2. the bin column scans of the hash gradient turn the rows into positions.
//   3. grid_bin_write_kernel: the same runs again; a run's brick (32 floats) goes to run
//      slot (chunk * 64 + rank of the run in its chunk), and {run slot, cell} to each
//      tile's list at the row's next position (LDS atomics on a row only the wavefront
//      touches, so the lists are the same on every run).
//   4. grid_bin_sum_kernel: four work-groups per tile; each wavefront owns an 8 x 8 x 16
//      block of the tile in LDS, queues the list entries whose bricks reach it (in list
//      order) and adds their slots, then adds its block into grad_grid.  Every vertex's sum
//      follows the list order, so grad_grid is the same on every run.
//   5. grid_edge_kernel: the samples the write pass flagged, in sample order.
// ------------------------------------------------------------------------------------
constexpr int kGtX = 32, kGtY = 32, kGtZ = 16;  // vertices per grid tile (64 KB of LDS)
constexpr int kGridRowSamples = 512;            // samples per count row: one wavefront
constexpr int kGridRowChunks = kGridRowSamples / 64;
constexpr int kGridMaxBins = 4096;              // tiles of a grid the binned passes take
constexpr int kGridMaxDim = 1024;               // cell coordinates are packed in 10 bits
constexpr int kStageStride = 33;                // LDS floats per staged brick

struct GridBinDev {
  int tx, ty, tz;   // tiles per axis
  int nbins;        // tx * ty * tz, tile (bx, by, bz) = bin (bx * ty + by) * tz + bz
  int nrows;        // count rows
  int32_t* cnt;     // [nrows][nbins] (pairs per row and tile, then the row's first position)
  int32_t* base;    // [nbins + 1] first list entry of each tile
  uint2* ref;       // [pairs] {run slot, cell of the run}, tile by tile
  float* val;       // [samples][32] brick of the run in each run slot
  uint64_t* edge;   // [nrows * kGridRowChunks] bit i of word w: sample 64 w + i has atomic stencils
};

// Last lane of a run of equal keys (the tail seg_sum reports).
VXN_DEV bool run_tail(uint64_t key, int lane) {
  const int klo = (int)(uint32_t)key, khi = (int)(uint32_t)(key >> 32);
  const uint64_t kn = (uint64_t)(uint32_t)dpp_i<kDppWaveShl1, 0xF>(klo) |
                      ((uint64_t)(uint32_t)dpp_i<kDppWaveShl1, 0xF>(khi) << 32);
  return lane == 63 || kn != key;
}

// Vertex range of a brick along one axis: Eq. 7 reaches c0 - 1 .. c0 + 2, Eq. 4 c0 .. c0 + 1.
VXN_DEV int brick_lo(bool ana, int c) { return ana ? c : c - 1; }
VXN_DEV int brick_hi(bool ana, int c) { return ana ? c + 1 : c + 2; }

// Calls f(bin) for every tile the brick at cell c reaches (1 to 8 of them).
template <class F>
VXN_DEV void brick_tiles(const GridBinDev& b, bool ana, const int (&c)[3], F&& f) {
  const int x0 = brick_lo(ana, c[0]) / kGtX, x1 = brick_hi(ana, c[0]) / kGtX;
  const int y0 = brick_lo(ana, c[1]) / kGtY, y1 = brick_hi(ana, c[1]) / kGtY;
  const int z0 = brick_lo(ana, c[2]) / kGtZ, z1 = brick_hi(ana, c[2]) / kGtZ;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int bx = x0 + ((q >> 2) & 1), by = y0 + ((q >> 1) & 1), bz = z0 + (q & 1);
    if (bx <= x1 && by <= y1 && bz <= z1) f((bx * b.ty + by) * b.tz + bz);
  }
}

__global__ void __launch_bounds__(256) grid_bin_count_kernel(RenderDev a, GridBinDev b) {
  extern __shared__ int32_t gcnt_lds[];  // [4][nbins]
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + wid;
  if (r >= b.nrows) return;
  int32_t* row = gcnt_lds + wid * b.nbins;
  for (int k = lane; k < b.nbins; k += 64) row[k] = 0;
  wave_sync();
  const bool ana = a.g.analytical != 0;
  const int64_t i0 = r * kGridRowSamples + lane;  // positions of the walk
  const int64_t nw = walk_length(a);
  GridIn nxt = grid_in(a, walk_sample(a, i0, nw));
  for (int c = 0; c < kGridRowChunks; ++c) {
    const GridIn cur = nxt;
    if (c + 1 < kGridRowChunks) nxt = grid_in(a, walk_sample(a, i0 + (c + 1) * 64, nw));
    float v[32];
    uint64_t key;
    int c0[3];
    bool edge;
    const bool brick = grid_brick<false, false>(a, cur, lane, v, key, c0, edge);
    if (run_tail(key, lane) && brick) brick_tiles(b, ana, c0, [&](int bin) { atomicAdd(&row[bin], 1); });
  }
  wave_sync();
  int32_t* out = b.cnt + r * b.nbins;
  for (int k = lane; k < b.nbins; k += 64) out[k] = row[k];
}

// LDS of the write pass: per wavefront its row of next positions and a 64-brick stage.
static size_t grid_write_lds(int nbins) {
  return 4 * ((size_t)nbins * sizeof(int32_t) + (size_t)64 * kStageStride * sizeof(float));
}

__global__ void __launch_bounds__(256) grid_bin_write_kernel(RenderDev a, GridBinDev b) {
  extern __shared__ int32_t gw_lds[];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + wid;
  if (r >= b.nrows) return;
  int32_t* pos = gw_lds + wid * b.nbins;                                   // [nbins]
  float* stage = (float*)(gw_lds + 4 * b.nbins) + wid * 64 * kStageStride;  // [64][kStageStride]
  const int32_t* prow = b.cnt + r * b.nbins;
  for (int k = lane; k < b.nbins; k += 64) pos[k] = prow[k];
  wave_sync();
  const bool ana = a.g.analytical != 0;
  const int64_t i0 = r * kGridRowSamples;  // positions of the walk
  const int64_t nw = walk_length(a);
  GridIn nxt = grid_in(a, walk_sample(a, i0 + lane, nw));
  for (int c = 0; c < kGridRowChunks; ++c) {
    const GridIn cur = nxt;
    if (c + 1 < kGridRowChunks) nxt = grid_in(a, walk_sample(a, i0 + (c + 1) * 64 + lane, nw));
    float v[32];
    uint64_t key;
    int c0[3];
    bool edge;
    const bool brick = grid_brick<true, false>(a, cur, lane, v, key, c0, edge);
    // the chunk's samples with atomic stencils, for the ordered edge pass
    const uint64_t eb = __ballot(edge);
    if (lane == 0) b.edge[r * kGridRowChunks + c] = eb;
    const bool tail = seg_sum<32>(key, v, lane) && brick;
    const uint64_t tb = __ballot(tail);
    if (tb == 0) continue;
    // run slots of this chunk: run0 + rank, rank = runs before this lane's in the chunk
    const int64_t run0 = i0 + c * 64;
    const int rank = __popcll(tb & ((1ull << lane) - 1ull));
    if (tail) {
#pragma unroll
      for (int k = 0; k < 32; ++k) stage[rank * kStageStride + k] = v[k];
      const uint32_t cell = (uint32_t)c0[0] | ((uint32_t)c0[1] << 10) | ((uint32_t)c0[2] << 20);
      const uint2 ent = make_uint2((uint32_t)(run0 + rank), cell);
      brick_tiles(b, ana, c0, [&](int bin) { b.ref[atomicAdd(&pos[bin], 1)] = ent; });
    }
    wave_sync();
    // the chunk's bricks are consecutive run slots: store them as whole 16-byte words
    const int nq = __popcll(tb) * 8;
    float4* dst = (float4*)(b.val + run0 * 32);
    for (int q = lane; q < nq; q += 64) {
      const float* sp = stage + (q >> 3) * kStageStride + (q & 7) * 4;
      dst[q] = make_float4(sp[0], sp[1], sp[2], sp[3]);
    }
    wave_sync();
  }
}

// Sum pass: four work-groups per tile, one per 16 x 16 x 16 quarter (x, y halves), of four
// wavefronts that each own an 8 x 8 x 16 block in LDS.  The work-group stages the tile's
// list 256 entries at a time (one per thread); each wavefront queues the staged entries
// whose bricks reach its block, then adds them two bricks per round (lanes 0-31 entry i,
// lanes 32-63 entry i + h of the queue), with the entries and brick values of kGsumBatch
// rounds loaded together, so a round waits only for its own LDS read-add-write.  Bricks
// closer than 4 cells on every axis may share vertices: their halves are added one after
// the other.
constexpr int kGsumWaves = 4;
constexpr int kGsumChunk = 64 * kGsumWaves;
constexpr int kGsumQueue = 128;
constexpr int kGsumBatch = 16;
constexpr int kGbX = 8, kGbY = 8;  // a wavefront's block: kGbX x kGbY x kGtZ vertices
constexpr int kGqX = 16, kGqY = 16;  // a work-group's quarter of the tile
static_assert((kGqX / kGbX) * (kGqY / kGbY) == kGsumWaves, "one block per wavefront");
static_assert((kGtX / kGqX) * (kGtY / kGqY) == 4, "four quarters per tile");
constexpr size_t kGsumLds = (size_t)kGsumWaves * kGbX * kGbY * kGtZ * sizeof(float) +
                            (size_t)kGsumChunk * sizeof(uint2) + (size_t)kGsumWaves * kGsumQueue * sizeof(uint2);

VXN_DEV int absdiff(int a, int b) { return a > b ? a - b : b - a; }

__global__ void __launch_bounds__(64 * kGsumWaves) grid_bin_sum_kernel(GridBinDev b, GridDesc g,
                                                                      float* __restrict__ grad_grid) {
  extern __shared__ float gsl[];
  constexpr int kBlock = kGbX * kGbY * kGtZ;
  const int t = threadIdx.x, wid = t >> 6, lane = t & 63;
  uint2* ents = (uint2*)(gsl + kGsumWaves * kBlock);    // [kGsumChunk] staged entries
  uint2* queue = ents + kGsumChunk + wid * kGsumQueue;  // this wavefront's queue
  const int bin = blockIdx.x >> 2, quarter = blockIdx.x & 3;
  const int bz = bin % b.tz, by = (bin / b.tz) % b.ty, bx = bin / (b.tz * b.ty);
  const int x0 = bx * kGtX + kGqX * (quarter >> 1) + kGbX * (wid >> 1);
  const int y0 = by * kGtY + kGqY * (quarter & 1) + kGbY * (wid & 1);
  const int z0 = bz * kGtZ;
  float* acc = gsl + wid * kBlock;  // [kGbX][kGbY][kGtZ]
  for (int k = lane; k < kBlock; k += 64) acc[k] = 0.0f;
  const bool ana = g.analytical != 0;
  // lane s (and s + 32) takes slot s of a brick: its offset from the brick's cell
  const int s = lane & 31, half = lane >> 5;
  int sdx, sdy, sdz;
  if (s < 8) {
    sdx = (s >> 2) & 1;
    sdy = (s >> 1) & 1;
    sdz = s & 1;
  } else {
    const int f = s - 8, axis = f >> 3, off = ((f >> 2) & 1) ? 2 : -1, u = (f >> 1) & 1, w = f & 1;
    sdx = axis == 0 ? off : u;
    sdy = axis == 0 ? u : (axis == 1 ? off : w);
    sdz = axis == 2 ? off : w;
  }
  const int beg = b.base[bin], end = b.base[bin + 1];
  // adds the queued bricks [0, qn) into the block
  auto drain = [&](int qn) {
    wave_sync();
    const int h = (qn + 1) >> 1;
    for (int i0 = 0; i0 < h; i0 += kGsumBatch) {
      uint32_t km[kGsumBatch];
      float val[kGsumBatch];
      bool apart[kGsumBatch];
#pragma unroll
      for (int q = 0; q < kGsumBatch; ++q) {
        const int i = i0 + q;
        const bool two = i + h < qn;
        const uint2 ea = i < h ? queue[i] : make_uint2(0u, 0u);
        const uint2 eb = two ? queue[i + h] : make_uint2(0u, 0u);
        const bool act = i < h && (half == 0 || two);
        const uint2 e = half ? eb : ea;
        km[q] = e.y;
        val[q] = act ? b.val[(int64_t)e.x * 32 + s] : 0.0f;
        apart[q] = !two || absdiff((int)(ea.y & 1023u), (int)(eb.y & 1023u)) > 3 ||
                   absdiff((int)((ea.y >> 10) & 1023u), (int)((eb.y >> 10) & 1023u)) > 3 ||
                   absdiff((int)(ea.y >> 20), (int)(eb.y >> 20)) > 3;
      }
#pragma unroll
      for (int q = 0; q < kGsumBatch; ++q) {
        if (i0 + q < h) {
          const int lx = (int)(km[q] & 1023u) + sdx - x0, ly = (int)((km[q] >> 10) & 1023u) + sdy - y0,
                    lz = (int)(km[q] >> 20) + sdz - z0;
          const bool ok = val[q] != 0.0f && (unsigned)lx < (unsigned)kGbX && (unsigned)ly < (unsigned)kGbY &&
                          (unsigned)lz < (unsigned)kGtZ;
          const int idx = (lx * kGbY + ly) * kGtZ + lz;
          if (apart[q]) {
            if (ok) acc[idx] += val[q];
          } else {
            if (ok && half == 0) acc[idx] += val[q];
            wave_sync();
            if (ok && half == 1) acc[idx] += val[q];
          }
          wave_sync();
        }
      }
    }
  };
  uint2 pe = beg + t < end ? b.ref[beg + t] : make_uint2(0u, 0u);
  for (int c0 = beg; c0 < end; c0 += kGsumChunk) {
    __syncthreads();
    ents[t] = pe;
    __syncthreads();
    if (c0 + kGsumChunk + t < end) pe = b.ref[c0 + kGsumChunk + t];  // the next chunk in flight
    const int n = min(kGsumChunk, end - c0);
    int qn = 0;
    for (int w0 = 0; w0 < n; w0 += 64) {
      const int j = w0 + lane;
      const uint2 e = j < n ? ents[j] : make_uint2(0u, 0u);
      const int cx = (int)(e.y & 1023u), cy = (int)((e.y >> 10) & 1023u);
      const bool touch = j < n && brick_hi(ana, cx) >= x0 && brick_lo(ana, cx) < x0 + kGbX &&
                         brick_hi(ana, cy) >= y0 && brick_lo(ana, cy) < y0 + kGbY;
      const uint64_t m = __ballot(touch);
      if (touch) queue[qn + __popcll(m & ((1ull << lane) - 1ull))] = e;
      qn += __popcll(m);
      if (qn > kGsumQueue - 64) {
        drain(qn);
        qn = 0;
      }
    }
    if (qn > 0) drain(qn);
  }
  // the block into grad_grid (z runs of kGtZ floats): every load first, then the stores
  constexpr int kPer = kBlock / 64;
  float v[kPer], o[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int k = lane + 64 * q;
    const int vx = x0 + k / (kGbY * kGtZ), vy = y0 + (k / kGtZ) % kGbY, vz = z0 + k % kGtZ;
    v[q] = acc[k];
    if (!(vx < g.X && vy < g.Y && vz < g.Z)) v[q] = 0.0f;
    o[q] = v[q] != 0.0f ? grad_grid[lin3(g, vx, vy, vz)] : 0.0f;
  }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int k = lane + 64 * q;
    const int vx = x0 + k / (kGbY * kGtZ), vy = y0 + (k / kGtZ) % kGbY, vz = z0 + k % kGtZ;
    if (v[q] != 0.0f) grad_grid[lin3(g, vx, vy, vz)] = o[q] + v[q];
  }
}

// Tiles of the binned grid gradient for an X x Y x Z grid, or 0 when it does not apply.
// Pass 5: one wavefront lists the samples of the write pass's edge words in sample order
// (in LDS, up to kEdgeList at a time), then adds their atomic stencils 64 samples per
// instruction, lane k taking the k-th listed sample; so those adds reach every vertex in
// sample order on every run.  A handful of samples per step take this path (the grid's
// border cells, and samples whose grid_sample cell is not one of the brick's).
constexpr int kEdgeBatch = 16;   // edge words per lane loaded at once
constexpr int kEdgeList = 2048;  // listed samples per flush

__global__ void __launch_bounds__(64) grid_edge_kernel(RenderDev a, const uint64_t* __restrict__ edge,
                                                       int64_t nwords) {
  __shared__ int32_t list[kEdgeList];  // positions of the walk
  const int lane = threadIdx.x;
  const int64_t nw = walk_length(a);
  int n = 0;  // listed samples (wave-uniform)
  auto flush = [&]() {
    wave_sync();
    for (int b = 0; b < n; b += 64) {
      GridIn x{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      if (b + lane < n) x = grid_in(a, walk_sample(a, (int64_t)list[b + lane], nw));
      float v[32];
      uint64_t key;
      int c0[3];
      bool e;
      grid_brick<false, true>(a, x, lane, v, key, c0, e);
    }
    wave_sync();
    n = 0;
  };
  for (int64_t w0 = 0; w0 < nwords; w0 += 64 * kEdgeBatch) {
    uint64_t mine[kEdgeBatch];
#pragma unroll
    for (int u = 0; u < kEdgeBatch; ++u) {
      const int64_t w = w0 + u * 64 + lane;
      mine[u] = w < nwords ? edge[w] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < kEdgeBatch; ++u) {
      if (__ballot(mine[u] != 0ull) == 0ull) continue;
      // exclusive scan of the lanes' sample counts: lane order is sample order
      const int c = __popcll(mine[u]);
      int inc = c;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
      }
      const int tot = __shfl(inc, 63, 64);
      if (n + tot > kEdgeList) flush();
      int pos = n + inc - c;
      uint64_t m = mine[u];
      const int64_t s0 = (w0 + u * 64 + lane) * 64;
      while (m != 0ull) {
        const int bit = __ffsll((unsigned long long)m) - 1;
        m &= m - 1ull;
        list[pos++] = (int32_t)(s0 + bit);
      }
      n += tot;
    }
  }
  if (n > 0) flush();
}

static int grid_bins(int X, int Y, int Z) {
  if (X > kGridMaxDim || Y > kGridMaxDim || Z > kGridMaxDim) return 0;
  const int64_t n = (int64_t)((X + kGtX - 1) / kGtX) * ((Y + kGtY - 1) / kGtY) * ((Z + kGtZ - 1) / kGtZ);
  return n <= kGridMaxBins ? (int)n : 0;
}

// Fixed-order reduction of the per-work-group partials: work-group b2 sums the blocks
// {b2, b2 + 64, ...} for 256 consecutive parameters; a second pass sums the 64 rows.
constexpr int kRedRows = 64;

__global__ void __launch_bounds__(256) render_reduce1_kernel(const float* __restrict__ part_mlp, int nb,
                                                             float* __restrict__ rows, int np) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int r = blockIdx.y;
  if (k >= np) return;
  float acc = 0.0f;
  for (int b = r; b < nb; b += kRedRows) acc += part_mlp[(int64_t)b * np + k];
  rows[(int64_t)r * np + k] = acc;
}

// Clears n floats (the atomic hash scatter's table gradient when it is to be overwritten;
// that path also takes tables that are not 8-byte aligned).
__global__ void __launch_bounds__(256) zero_f32_kernel(float* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// The variance outputs of the reductions (nullable var: none).
struct VarianceOut {
  const float* var;
  float s_max;
  float* s_out;
  float* d_var;
  float amp_k;
  int32_t* amplified;
};

__global__ void __launch_bounds__(256) render_reduce2_kernel(const float* __restrict__ rows,
                                                             const double* __restrict__ part_sc, int nb,
                                                             float* __restrict__ grad_mlp, float* __restrict__ grad_s,
                                                             double* __restrict__ loss, float inv_n3, int np,
                                                             const float* __restrict__ loss_norm, float mask_weight,
                                                             float inv_r, double* __restrict__ loss_parts,
                                                             VarianceOut vo) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < np) {
    double acc = 0.0;
    for (int r = 0; r < kRedRows; ++r) acc += (double)rows[(int64_t)r * np + k];
    grad_mlp[k] = (float)acc;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x < 64) {  // one wavefront: fixed-order strided sums
    double ds = 0.0, ls = 0.0, lb = 0.0;
    for (int b = threadIdx.x; b < nb; b += 64) {
      ds += part_sc[3 * b];
      ls += part_sc[3 * b + 1];
      lb += part_sc[3 * b + 2];
    }
    ds = wave_sum_d(ds);
    ls = wave_sum_d(ls);
    lb = wave_sum_d(lb);
    if (threadIdx.x == 0) {
      const float gs = (float)ds;
      if (grad_s) grad_s[0] = gs;
      if (vo.var) {
        // dL/dv of s = clamp(exp(10 v), 1e-4, s_max) (models/variance.py:38-41), then
        // amplify_s_gradient (training/scheduler.py:14-33), in PyTorch's operation order
        const float e = expf(vo.var[0] * 10.0f);
        if (vo.s_out) vo.s_out[0] = fminf(fmaxf(e, 1e-4f), vo.s_max);
        if (vo.d_var) {
          const float inside = (e >= 1e-4f && e <= vo.s_max) ? 1.0f : 0.0f;
          float dv = ((gs * e) * inside) * 10.0f;
          if (vo.amplified) {
            const bool neg = dv < 0.0f;
            if (neg) dv = dv * vo.amp_k;
            vo.amplified[0] = neg ? 1 : 0;
          }
          vo.d_var[0] = dv;
        }
      }
      const double lc = ls * (double)(loss_norm ? loss_norm[0] : inv_n3);
      const double lm = lb * (double)inv_r;
      if (loss) loss[0] = lc + (double)mask_weight * lm;
      if (loss_parts) {
        loss_parts[0] = lc;
        loss_parts[1] = lm;
      }
    }
  }
}

// The colour loss's normalisation with a foreground mask (trainer.py:223-231): out[0] =
// 1 / (3 * sum(mask)) and out[1] = 1 when the mask has weight, else 1 / (3 R) and 0.
__global__ void __launch_bounds__(256) mask_norm_kernel(const float* __restrict__ mask, int64_t R,
                                                        float* __restrict__ out) {
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < R; i += 256) acc += (double)mask[i];
  acc = wave_sum_d(acc);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double fg = 3.0 * (((red[0] + red[1]) + red[2]) + red[3]);
    out[0] = fg > 0.0 ? (float)(1.0 / fg) : (float)(1.0 / (3.0 * (double)R));
    out[1] = fg > 0.0 ? 1.0f : 0.0f;
  }
}

static int device_cus() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cached[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cached[dev] = n;
  }
  return cached[dev];
}

static int render_blocks(int64_t R) {
  const int64_t cap = (int64_t)device_cus() * 4;
  return (int)(R < cap ? (R > 0 ? R : 1) : cap);
}

static int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

// Shape of the binned hash gradient for ns samples over L levels of T = 2^log2_T
// entries; use = false selects the atomic scatter (more bins per level than the LDS
// counters hold, or records beyond int32).
struct BinPlan {
  bool use;
  int nbl, nbins, ntile, nrows, nchunk;
  int64_t n_cnt, n_csum, max_rec;
};

// Records: 4 per sample and level, 8 on levels whose x can end in kBinBits 1 bits (res >=
// kBinEntries - 1: the x corner pair may straddle two bins).
static BinPlan bin_plan(int64_t ns, int L, int log2_T, const int32_t* level_res) {
  BinPlan p{};
  if (ns <= 0 || L <= 0 || L > VXN_MAX_LEVELS || log2_T < 0 || log2_T > 30) return p;
  const int64_t T = (int64_t)1 << log2_T;
  p.nbl = (int)((T + kBinEntries - 1) / kBinEntries);
  p.nbins = L * p.nbl;
  const int64_t ntile = (ns + kBinTile - 1) / kBinTile;
  int per = 0;
  for (int l = 0; l < L; ++l) per += level_res[l] >= kBinEntries - 1 ? 8 : 4;
  p.max_rec = ns * per;
  p.use = p.nbl <= kMaxBinsPerLevel && p.max_rec < ((int64_t)1 << 31) && ntile * L < ((int64_t)1 << 31) &&
          ntile * (int64_t)p.nbins < ((int64_t)1 << 31);
  if (!p.use) return p;
  p.ntile = (int)ntile;
  p.nrows = p.ntile;
  p.nchunk = (p.nrows + kColChunk - 1) / kColChunk;
  p.n_cnt = (int64_t)L * p.nrows * p.nbl;
  p.n_csum = (int64_t)L * p.nchunk * p.nbl;
  return p;
}

// Workspace layout: MLP partials | scalar partials | reduction rows | gfeat | ggeo | hfeat
// | bin count rows | column chunk sums | bin bases | record entries | record values
// | grid tile count rows | their column chunk sums | tile bases | tile lists | bricks.
struct RenderWs {
  int64_t off_sc, off_norm, off_rows, off_feat, off_geo, off_hf, off_cnt, off_csum, off_base, off_ridx, off_rval;
  int64_t off_gcnt, off_gcsum, off_gbase, off_gref, off_gval, off_gedge, off_live, off_lidx, off_lcnt, off_nlive,
      off_gsd, off_need, off_lidx2, off_lcnt2, off_nlive2, total;
  BinPlan bins;
  int gbins;   // grid tiles of the binned grid gradient (0: the atomic grid scatter)
  int grows;   // its count rows
};

static RenderWs render_ws(int64_t R, int S, int L, int log2_T, const int32_t* level_res, int np, int gbins) {
  RenderWs w;
  const int64_t nb = render_blocks(R);
  const int64_t ns = R * (int64_t)S;
  w.off_sc = align256(nb * np * (int64_t)sizeof(float));
  w.off_norm = align256(w.off_sc + nb * 3 * (int64_t)sizeof(double));
  w.off_rows = align256(w.off_norm + 2 * (int64_t)sizeof(float));
  w.off_feat = align256(w.off_rows + (int64_t)kRedRows * np * (int64_t)sizeof(float));
  w.off_geo = align256(w.off_feat + 32 * ns * (int64_t)sizeof(float));
  w.off_hf = align256(w.off_geo + kGeo * ns * (int64_t)sizeof(float));
  w.off_cnt = align256(w.off_hf + 16 * ns * (int64_t)sizeof(uint32_t));
  w.bins = bin_plan(ns, L, log2_T, level_res);
  int64_t end;
  if (w.bins.use) {
    w.off_csum = align256(w.off_cnt + w.bins.n_cnt * 4);
    w.off_base = align256(w.off_csum + w.bins.n_csum * 4);
    w.off_ridx = align256(w.off_base + ((int64_t)w.bins.nbins + 1) * 4);
    w.off_rval = align256(w.off_ridx + w.bins.max_rec * 4);
    end = align256(w.off_rval + w.bins.max_rec * 16);
  } else {
    w.off_csum = w.off_base = w.off_ridx = w.off_rval = w.off_cnt;
    end = w.off_cnt;
  }
  // the geometry pass's per-sample SDF and gradient, and the need flags
  w.off_gsd = end;
  w.off_need = align256(w.off_gsd + ns * 16);
  end = align256(w.off_need + ns);
  // the live-sample list of the record passes
  if (w.bins.use) {
    w.off_live = end;
    w.off_lidx = align256(w.off_live + ns);
    w.off_lcnt = align256(w.off_lidx + ns * 4);
    w.off_nlive = align256(w.off_lcnt + ((ns + kLiveBlock - 1) / kLiveBlock) * 4);
    // the grid gradient's own list (BF16 form: the hash-table gradient's list is shorter)
    w.off_lidx2 = align256(w.off_nlive + 4);
    w.off_lcnt2 = align256(w.off_lidx2 + ns * 4);
    w.off_nlive2 = align256(w.off_lcnt2 + ((ns + kLiveBlock - 1) / kLiveBlock) * 4);
    end = align256(w.off_nlive2 + 4);
  } else {
    w.off_live = w.off_lidx = w.off_lcnt = w.off_nlive = w.off_lidx2 = w.off_lcnt2 = w.off_nlive2 = end;
  }
  w.gbins = gbins;
  w.grows = (int)((ns + kGridRowSamples - 1) / kGridRowSamples);
  w.off_gcnt = end;
  if (gbins > 0) {
    const int64_t gchunk = (w.grows + kColChunk - 1) / kColChunk;
    w.off_gcsum = align256(w.off_gcnt + (int64_t)w.grows * gbins * 4);
    w.off_gbase = align256(w.off_gcsum + gchunk * gbins * 4);
    w.off_gref = align256(w.off_gbase + ((int64_t)gbins + 1) * 4);
    w.off_gval = align256(w.off_gref + ns * 8 * 8);  // a run reaches at most 8 tiles
    w.off_gedge = align256(w.off_gval + ns * 32 * 4);
    end = align256(w.off_gedge + (int64_t)w.grows * kGridRowChunks * 8);
  } else {
    w.off_gcsum = w.off_gbase = w.off_gref = w.off_gval = w.off_gedge = end;
  }
  w.total = end;
  return w;
}

static RenderWs render_ws(const vxn_render_args& A) {
  const int gbins = A.grid_binned ? grid_bins(A.X, A.Y, A.Z) : 0;
  return render_ws(A.n_rays, A.n_samples, A.n_levels, A.log2_T, A.level_res,
                   A.view_embed == VXN_VIEW_PE ? MlpLayout<27>::NP : MlpLayout<16>::NP, gbins);
}

static int build_render_dev(const vxn_render_args& A, RenderDev& a, int mode) {
  a = RenderDev{};  // every optional pointer starts null
  if (int e = check_grid(A.grid, A.X, A.Y, A.Z, A.box)) return e;
  if (int e = check_gradient_mode(A.gradient_mode)) return e;
  VXN_REQUIRE(A.n_rays >= 0, VXN_ERR_SHAPE, "n_rays must be >= 0");
  VXN_REQUIRE(A.n_samples >= 1 && A.n_samples <= kRT, VXN_ERR_UNSUPPORTED, "n_samples must be in [1, %d] (got %d)",
              kRT, A.n_samples);
  VXN_REQUIRE(A.n_levels == 16 && A.n_features == 2, VXN_ERR_UNSUPPORTED,
              "the fused render implements the 16-level x 2-feature hash feeding the radiance MLP");
  VXN_REQUIRE(A.mlp_mode == VXN_MLP_FP32 || A.mlp_mode == VXN_MLP_BF16_MFMA, VXN_ERR_VALUE, "unknown mlp_mode %d",
              A.mlp_mode);
  VXN_REQUIRE(A.view_embed == VXN_VIEW_SH || A.view_embed == VXN_VIEW_PE, VXN_ERR_VALUE, "unknown view_embed %d",
              A.view_embed);
  VXN_REQUIRE(A.rays_o && A.rays_d && A.mid_z && A.dists && A.table && A.mlp, VXN_ERR_NULL, "null input");
  VXN_REQUIRE(A.n_rays * (int64_t)A.n_samples < ((int64_t)1 << 31), VXN_ERR_SHAPE,
              "n_rays * n_samples must be below 2^31 (got %lld)", (long long)(A.n_rays * (int64_t)A.n_samples));
  HashDesc h;
  if (int e = make_hash_desc(h, A.table, A.table_dtype, A.n_levels, A.log2_T, A.n_features, A.level_res, A.level_w))
    return e;
  a.o = A.rays_o;
  a.d = A.rays_d;
  a.mid = A.mid_z;
  a.dist = A.dists;
  a.R = A.n_rays;
  a.S = A.n_samples;
  a.sdiv = vxn_fastdiv_make((uint32_t)A.n_samples);
  a.g = make_grid_desc(A.grid, A.X, A.Y, A.Z, A.box);
  a.g.analytical = A.gradient_mode;
  a.h = h;
  a.mlp = A.mlp;
  a.s = A.s;
  a.s_dev = A.s_dev;
  a.var_dev = A.variance;
  a.s_max = A.s_max;
  a.sphere_r = A.sphere_radius;
  a.cos_anneal = A.cos_anneal_ratio;
  a.iter_cos = A.use_iter_cos;
  a.rgb = A.rgb;
  a.depth = A.depth;
  a.normal = A.normal;
  a.opacity = A.opacity;
  a.weights = A.weights;
  a.alpha = A.alpha;
  a.trans = A.transmittance;
  a.sdf_s = A.sdf_samples;
  a.grad_s = A.grad_samples;
  a.normal_s = A.normal_samples;
  a.corners = A.corner_indices;
  a.d_rgb = A.d_rgb;
  a.d_depth = A.d_depth;
  a.d_normal = A.d_normal;
  a.d_opacity = A.d_opacity;
  a.d_weights = A.d_weights;
  a.target = A.target_rgb;
  a.huber_delta = A.huber_delta;
  a.inv_n3 = A.n_rays > 0 ? (float)(1.0 / (3.0 * (double)A.n_rays)) : 0.0f;
  a.inv_r = A.n_rays > 0 ? (float)(1.0 / (double)A.n_rays) : 0.0f;
  a.loss_type = A.loss_type;
  a.target_mask = mode == kTrain ? A.target_mask : nullptr;
  a.mask_weight = A.mask_weight;
  a.loss_norm = nullptr;
  a.grad_grid = A.grad_grid;
  a.grad_table = A.grad_table;
  a.vr = A.vr_bitmask;
  a.part_mlp = nullptr;
  a.part_sc = nullptr;
  a.gfeat = nullptr;
  a.ggeo = nullptr;
  a.hfeat = nullptr;
  a.scenes = nullptr;
  a.ray_scene = nullptr;
  a.n_rays_dev = nullptr;
  a.n_scenes = 0;
  VXN_REQUIRE(A.variance == nullptr || A.s_max > 0.0f, VXN_ERR_VALUE, "s_max must be > 0 with a variance");
  if (mode != kFwd) {
    VXN_REQUIRE(A.grad_grid && A.grad_table && A.grad_mlp && A.grad_s, VXN_ERR_NULL, "null gradient output");
    const RenderWs w = render_ws(A);
    VXN_REQUIRE(A.workspace != nullptr && A.workspace_bytes >= w.total, VXN_ERR_WORKSPACE,
                "render workspace needs %lld bytes", (long long)w.total);
    char* ws = (char*)A.workspace;
    a.part_mlp = (float*)ws;
    a.part_sc = (double*)(ws + w.off_sc);
    a.loss_norm = (const float*)(ws + w.off_norm);
    a.gfeat = (float*)(ws + w.off_feat);
    a.ggeo = (float*)(ws + w.off_geo);
    a.hfeat = ws + w.off_hf;
    if (w.bins.use) a.live = (uint8_t*)(ws + w.off_live);
    if (A.mlp_mode == VXN_MLP_BF16_MFMA) {  // the geometry pass runs ahead of the feature gather
      a.geo = (float*)(ws + w.off_gsd);
      a.need = (uint8_t*)(ws + w.off_need);
    }
  }
  if (mode == kTrain) {
    VXN_REQUIRE(A.target_rgb != nullptr, VXN_ERR_NULL, "target_rgb is null");
    VXN_REQUIRE(A.huber_delta > 0.0f, VXN_ERR_VALUE, "huber_delta must be > 0");
    VXN_REQUIRE(A.loss_type == VXN_LOSS_HUBER || A.loss_type == VXN_LOSS_MSE, VXN_ERR_VALUE,
                "unknown loss_type %d", A.loss_type);
    VXN_REQUIRE(A.mask_weight >= 0.0f, VXN_ERR_VALUE, "mask_weight must be >= 0");
  }
  if (mode != kFwd && A.scatter_stream) {
    hipEvent_t const* se = (hipEvent_t const*)A.scatter_events;
    VXN_REQUIRE(se != nullptr && se[0] != nullptr && se[1] != nullptr, VXN_ERR_NULL,
                "scatter_stream needs scatter_events[2]");
  }
  return 0;
}

// Hash levels whose cells are wide enough for consecutive samples to share them (up to
// res 2048: merging runs of 1.05-1.2 samples there still beat factored records, whose
// repeated entries within a record group cost the sum pass extra tag rounds).
static int dedup_levels(const HashDesc& h) {
  int n = 0;
  while (n < h.L && h.res[n] <= 2048.0f) ++n;
  return n;
}

template <int MODE, int NV>
static int launch_render_nv(const vxn_render_args* p, void* stream) {
  VXN_REQUIRE(p != nullptr, VXN_ERR_NULL, "args is null");
  RenderDev a{};
  if (int e = build_render_dev(*p, a, MODE)) return e;
  hipStream_t s = (hipStream_t)stream;
  if (p->n_rays == 0) {
    if (MODE != kFwd && p->grad_table_overwrite) {
      const int64_t nf = ((int64_t)a.h.L << p->log2_T) * 2;
      const unsigned nz = (unsigned)((nf + 255) / 256 < 65536 ? (nf + 255) / 256 : 65536);
      hipLaunchKernelGGL(zero_f32_kernel, dim3(nz), dim3(256), 0, s, p->grad_table, nf);
      VXN_LAUNCH_CHECK();
    }
    return 0;
  }
  const int nb = render_blocks(p->n_rays);
  hipEvent_t const* ev = (hipEvent_t const*)p->kernel_events;
  const int64_t ns = p->n_rays * (int64_t)p->n_samples;
  // second stream: the grid gradient follows the render kernel there, beside the
  // hash-table gradient's passes
  hipStream_t ss = MODE != kFwd ? (hipStream_t)p->scatter_stream : nullptr;
  hipEvent_t const* se = (hipEvent_t const*)p->scatter_events;
  RenderWs w{};
  BinDev b{};
  bool binned_hash = false;
  int dl = 0;
  if (MODE != kFwd) {
    w = render_ws(*p);
    binned_hash = w.bins.use && ((uintptr_t)p->grad_table & 7) == 0;
    dl = dedup_levels(a.h);
    if (binned_hash) {
      char* ws = (char*)p->workspace;
      b.nbl = w.bins.nbl;
      b.nbins = w.bins.nbins;
      b.ntile = w.bins.ntile;
      b.nrows = w.bins.nrows;
      b.nchunk = w.bins.nchunk;
      b.cnt = (int32_t*)(ws + w.off_cnt);
      b.csum = (int32_t*)(ws + w.off_csum);
      b.base = (int32_t*)(ws + w.off_base);
      b.ridx = (uint32_t*)(ws + w.off_ridx);
      b.rval = (float4*)(ws + w.off_rval);
      b.fact_level = dl;
    }
  }
  // pass 1 and 2 of the binned hash-table gradient: record counts per tile and bin, then
  // every tile's first position in each bin
  auto count_and_scan = [&](hipStream_t st) -> int {
    const unsigned nwg = (unsigned)(a.h.L * b.ntile);
    hipLaunchKernelGGL(hash_bin_count_kernel, dim3(nwg), dim3(256), (size_t)b.nbl * sizeof(int32_t), st, a, b, dl);
    VXN_LAUNCH_CHECK();
    const dim3 cgrid((unsigned)((b.nbl + 255) / 256), (unsigned)b.nchunk, (unsigned)a.h.L);
    hipLaunchKernelGGL(bin_col_sums_kernel, cgrid, dim3(256), 0, st, b);
    VXN_LAUNCH_CHECK();
    hipLaunchKernelGGL(bin_col_scan_kernel, dim3((unsigned)((b.nbins + 255) / 256)), dim3(256), 0, st, b);
    VXN_LAUNCH_CHECK();
    hipLaunchKernelGGL(bin_base_scan_kernel, dim3(1), dim3(kColThreads), 0, st, b);
    VXN_LAUNCH_CHECK();
    hipLaunchKernelGGL(bin_col_positions_kernel, cgrid, dim3(256), 0, st, b);
    VXN_LAUNCH_CHECK();
    return 0;
  };
  // the samples listed in order by their byte flags: the record passes and the grid
  // gradient walk only them
  // (bit `bit` of each flag byte; list k = 0 for the record passes, 1 for the grid gradient)
  const int32_t* grid_idx = nullptr;
  const int32_t* grid_n = nullptr;
  auto list_samples = [&](hipStream_t st, const uint8_t* flags, int bit, int k) -> int {
    char* wsp = (char*)p->workspace;
    const unsigned nlb = (unsigned)((ns + kLiveBlock - 1) / kLiveBlock);
    int32_t* lcnt = (int32_t*)(wsp + (k ? w.off_lcnt2 : w.off_lcnt));
    int32_t* lidx = (int32_t*)(wsp + (k ? w.off_lidx2 : w.off_lidx));
    int32_t* nlive = (int32_t*)(wsp + (k ? w.off_nlive2 : w.off_nlive));
    hipLaunchKernelGGL(live_count_kernel, dim3(nlb), dim3(256), 0, st, flags, ns, bit, lcnt);
    VXN_LAUNCH_CHECK();
    hipLaunchKernelGGL(live_list_kernel, dim3(nlb), dim3(256), 0, st, flags, ns, bit, (const int32_t*)lcnt, lidx,
                       nlive);
    VXN_LAUNCH_CHECK();
    if (k) {
      grid_idx = lidx;
      grid_n = nlive;
    } else {
      a.live_idx = lidx;
      a.n_live = nlive;
    }
    return 0;
  };
  // BF16 backward / training: the geometry pass's need flags cover every sample with a
  // gradient (w != 0 needs alpha > 0, dq != 0 needs 0 <= q <= 1, both inside the sphere;
  // only w != 0 gives the hash features a gradient), so the record passes walk the list
  // of bit 1 and the grid gradient that of bit 0, and the hash-table gradient's count and
  // scans, which read no gradients, run before the render kernel (beside the feature
  // gather and the render kernel when scatter_stream is given: se[1] marks them)
  bool early_list = false;
  a.sparse = binned_hash ? 1 : 0;
  if (ev) VXN_HIP(hipEventRecord(ev[0], s));
  if (MODE != kFwd && p->mlp_mode == VXN_MLP_BF16_MFMA) {
    hipLaunchKernelGGL(geo_kernel, dim3((unsigned)((ns + 127) / 128)), dim3(128), 0, s, a);
    VXN_LAUNCH_CHECK();
    if (binned_hash) {
      a.live = nullptr;  // the render kernel's flags are not needed
      early_list = true;
      hipStream_t ls = s;
      if (ss) {
        VXN_HIP(hipEventRecord(se[0], s));
        VXN_HIP(hipStreamWaitEvent(ss, se[0], 0));
        ls = ss;
      }
      if (int e = list_samples(ls, a.need, 1, 0)) return e;
      if (int e = count_and_scan(ls)) return e;
      if (int e = list_samples(ls, a.need, 0, 1)) return e;
      if (ss) VXN_HIP(hipEventRecord(se[1], ss));
    }
    const int64_t chunks = (ns + 255) / 256;
    const unsigned nb_hf = (unsigned)(8 * ((a.h.L + 7) / 8) * chunks);
    hipLaunchKernelGGL(hash_features_kernel<0>, dim3(nb_hf), dim3(256), 0, s, a, chunks);
    VXN_LAUNCH_CHECK();
  }
  if (ev) VXN_HIP(hipEventRecord(ev[1], s));
  if (MODE == kTrain && a.target_mask) {
    hipLaunchKernelGGL(mask_norm_kernel, dim3(1), dim3(256), 0, s, a.target_mask, p->n_rays, (float*)a.loss_norm);
    VXN_LAUNCH_CHECK();
  }
  // the binned grid gradient writes grad_grid with plain read-add-writes: work that the
  // caller forks from fork_event (atomic adds into grad_grid) starts after it
  const int gbins = w.gbins;
  if (p->fork_event && gbins == 0) VXN_HIP(hipEventRecord((hipEvent_t)p->fork_event, s));
  if (p->mlp_mode == VXN_MLP_BF16_MFMA)
    hipLaunchKernelGGL((render_kernel<MODE, kMlpBf16, 0, NV>), dim3(nb), dim3(kRT), 0, s, a);
  else
    hipLaunchKernelGGL((render_kernel<MODE, kMlpF32, 0, NV>), dim3(nb), dim3(kRT), 0, s, a);
  VXN_LAUNCH_CHECK();
  if (ev) VXN_HIP(hipEventRecord(ev[2], s));
  if (MODE != kFwd) {
    const int sb = (int)((ns + 255) / 256);
    if (early_list) {
      // the counts and positions from the second stream (se[1] is recorded again below)
      if (ss) VXN_HIP(hipStreamWaitEvent(s, se[1], 0));
    } else if (binned_hash) {
      // the samples with a gradient (the render kernel's flags: w != 0 or dq != 0)
      if (int e = list_samples(s, (const uint8_t*)a.live, 0, 0)) return e;
    }
    // the grid gradient's passes walk their own list where there is one
    RenderDev ag = a;
    if (grid_idx) {
      ag.live_idx = grid_idx;
      ag.n_live = grid_n;
    }
    // the grid gradient needs only the render kernel's per-sample geometry gradients
    hipStream_t gs = s;
    if (ss) {
      VXN_HIP(hipEventRecord(se[0], s));
      VXN_HIP(hipStreamWaitEvent(ss, se[0], 0));
      gs = ss;
    }
    if (gbins > 0) {
      char* ws = (char*)p->workspace;
      GridBinDev gb;
      gb.tx = (a.g.X + kGtX - 1) / kGtX;
      gb.ty = (a.g.Y + kGtY - 1) / kGtY;
      gb.tz = (a.g.Z + kGtZ - 1) / kGtZ;
      gb.nbins = w.gbins;
      gb.nrows = w.grows;
      gb.cnt = (int32_t*)(ws + w.off_gcnt);
      gb.base = (int32_t*)(ws + w.off_gbase);
      gb.ref = (uint2*)(ws + w.off_gref);
      gb.val = (float*)(ws + w.off_gval);
      gb.edge = (uint64_t*)(ws + w.off_gedge);
      // the column scans of the hash gradient on the tile count rows (one "level")
      BinDev cb = {};
      cb.nbl = w.gbins;
      cb.nbins = w.gbins;
      cb.nrows = w.grows;
      cb.nchunk = (w.grows + kColChunk - 1) / kColChunk;
      cb.cnt = gb.cnt;
      cb.csum = (int32_t*)(ws + w.off_gcsum);
      cb.base = gb.base;
      const unsigned gwg = (unsigned)((w.grows + 3) / 4);
      hipLaunchKernelGGL(grid_bin_count_kernel, dim3(gwg), dim3(256), (size_t)4 * w.gbins * sizeof(int32_t), gs, ag,
                         gb);
      VXN_LAUNCH_CHECK();
      const dim3 cgrid((unsigned)((cb.nbl + 255) / 256), (unsigned)cb.nchunk, 1u);
      hipLaunchKernelGGL(bin_col_sums_kernel, cgrid, dim3(256), 0, gs, cb);
      VXN_LAUNCH_CHECK();
      hipLaunchKernelGGL(bin_col_scan_kernel, dim3((unsigned)((cb.nbins + 255) / 256)), dim3(256), 0, gs, cb);
      VXN_LAUNCH_CHECK();
      hipLaunchKernelGGL(bin_base_scan_kernel, dim3(1), dim3(kColThreads), 0, gs, cb);
      VXN_LAUNCH_CHECK();
      hipLaunchKernelGGL(bin_col_positions_kernel, cgrid, dim3(256), 0, gs, cb);
      VXN_LAUNCH_CHECK();
      hipLaunchKernelGGL(grid_bin_write_kernel, dim3(gwg), dim3(256), grid_write_lds(w.gbins), gs, ag, gb);
      VXN_LAUNCH_CHECK();
      if (p->grid_wait_event) VXN_HIP(hipStreamWaitEvent(gs, (hipEvent_t)p->grid_wait_event, 0));
      hipLaunchKernelGGL(grid_bin_sum_kernel, dim3((unsigned)(4 * w.gbins)), dim3(64 * kGsumWaves), kGsumLds, gs, gb,
                         a.g, a.grad_grid);
      VXN_LAUNCH_CHECK();
      // the edge pass adds with atomics: forked atomic adds into grad_grid may run beside it
      if (p->fork_event) VXN_HIP(hipEventRecord((hipEvent_t)p->fork_event, gs));
      hipLaunchKernelGGL(grid_edge_kernel, dim3(1), dim3(64), 0, gs, ag, (const uint64_t*)gb.edge,
                         (int64_t)w.grows * kGridRowChunks);
      VXN_LAUNCH_CHECK();
    } else {
      hipLaunchKernelGGL(grid_scatter_kernel, dim3(sb), dim3(256), 0, gs, ag);
      VXN_LAUNCH_CHECK();
    }
    if (ev) VXN_HIP(hipEventRecord(ev[3], gs));
    if (ss) VXN_HIP(hipEventRecord(se[1], ss));
    if (binned_hash) {
      if (!early_list) {
        if (int e = count_and_scan(s)) return e;
      }
      // levels below dl carry run sums (record values + entry array), the rest factored records
      if (dl > 0) {
        hipLaunchKernelGGL(hash_bin_write_kernel<false>, dim3((unsigned)(dl * b.ntile)), dim3(256),
                           bin_write_lds(b.nbl), s, a, b, dl, 0);
        VXN_LAUNCH_CHECK();
      }
      if (dl < a.h.L) {
        hipLaunchKernelGGL(hash_bin_write_kernel<true>, dim3((unsigned)((a.h.L - dl) * b.ntile)), dim3(256),
                           bin_write_lds(b.nbl), s, a, b, dl, dl);
        VXN_LAUNCH_CHECK();
      }
      hipLaunchKernelGGL(hash_bin_sum_kernel, dim3((unsigned)b.nbins), dim3(64 * kSumWaves), kSumLds, s, b,
                         p->grad_table,
                         (int64_t)1 << p->log2_T, p->grad_table_overwrite ? 1 : 0);
      VXN_LAUNCH_CHECK();
    } else {
      if (p->grad_table_overwrite) {
        const int64_t nf = ((int64_t)a.h.L << p->log2_T) * 2;
        const unsigned nz = (unsigned)((nf + 255) / 256 < 65536 ? (nf + 255) / 256 : 65536);
        hipLaunchKernelGGL(zero_f32_kernel, dim3(nz), dim3(256), 0, s, p->grad_table, nf);
        VXN_LAUNCH_CHECK();
      }
      hipLaunchKernelGGL(hash_scatter_kernel, dim3(sb), dim3(256), 0, s, a, dl);
      VXN_LAUNCH_CHECK();
    }
    if (ev) VXN_HIP(hipEventRecord(ev[4], s));
    float* rows = (float*)((char*)p->workspace + w.off_rows);
    constexpr int np = MlpLayout<NV>::NP;
    hipLaunchKernelGGL(render_reduce1_kernel, dim3((np + 255) / 256, kRedRows), dim3(256), 0, s, a.part_mlp, nb,
                       rows, np);
    VXN_LAUNCH_CHECK();
    hipLaunchKernelGGL(render_reduce2_kernel, dim3((np + 1 + 255) / 256), dim3(256), 0, s, rows, a.part_sc, nb,
                       p->grad_mlp, p->grad_s, MODE == kTrain ? p->loss : nullptr, a.inv_n3, np,
                       a.target_mask ? a.loss_norm : nullptr, a.target_mask ? a.mask_weight : 0.0f, a.inv_r,
                       MODE == kTrain ? p->loss_parts : nullptr,
                       VarianceOut{p->variance, p->s_max, p->s_out, p->d_variance, p->s_amp_k, p->amplified});
    VXN_LAUNCH_CHECK();
    if (ss) VXN_HIP(hipStreamWaitEvent(s, se[1], 0));
    if (ev) VXN_HIP(hipEventRecord(ev[5], s));
  }
  return 0;
}

template <int MODE>
static int launch_render(const vxn_render_args* p, void* stream) {
  VXN_REQUIRE(p != nullptr, VXN_ERR_NULL, "args is null");
  return p->view_embed == VXN_VIEW_PE ? launch_render_nv<MODE, 27>(p, stream) : launch_render_nv<MODE, 16>(p, stream);
}

}  // namespace vxn

using namespace vxn;

extern "C" int vxn_render_workspace_bytes(const vxn_render_args* a, int64_t* bytes) {
  VXN_REQUIRE(a != nullptr && bytes != nullptr, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(a->n_rays >= 0 && a->n_samples >= 0, VXN_ERR_SHAPE, "bad shape");
  *bytes = render_ws(*a).total;
  return 0;
}

extern "C" int vxn_render_forward(const vxn_render_args* a, void* stream) { return launch_render<kFwd>(a, stream); }
extern "C" int vxn_render_backward(const vxn_render_args* a, void* stream) { return launch_render<kBwd>(a, stream); }
extern "C" int vxn_render_train(const vxn_render_args* a, void* stream) { return launch_render<kTrain>(a, stream); }

// ------------------------------------------------------------------------------------
// Serving batches
// ------------------------------------------------------------------------------------

extern "C" int vxn_serve_scene_bytes(int64_t* bytes) {
  VXN_REQUIRE(bytes != nullptr, VXN_ERR_NULL, "bytes is null");
  *bytes = (int64_t)sizeof(SceneDev);
  return 0;
}

extern "C" int vxn_serve_pack_scene(const vxn_scene_desc* d, void* out, int64_t out_bytes) {
  VXN_REQUIRE(d != nullptr && out != nullptr, VXN_ERR_NULL, "null argument");
  VXN_REQUIRE(out_bytes >= (int64_t)sizeof(SceneDev), VXN_ERR_WORKSPACE, "a scene record needs %lld bytes",
              (long long)sizeof(SceneDev));
  if (int e = check_grid(d->grid, d->X, d->Y, d->Z, d->box)) return e;
  if (int e = check_gradient_mode(d->gradient_mode)) return e;
  VXN_REQUIRE(d->n_levels == 16 && d->n_features == 2, VXN_ERR_UNSUPPORTED,
              "the fused render implements the 16-level x 2-feature hash feeding the 51-input MLP");
  VXN_REQUIRE(d->mlp != nullptr && d->table != nullptr, VXN_ERR_NULL, "null table or mlp");
  if (d->occ_mask) VXN_REQUIRE(d->occ_res >= 1, VXN_ERR_SHAPE, "occ_res must be >= 1");
  SceneDev sc{};
  sc.g = make_grid_desc(d->grid, d->X, d->Y, d->Z, d->box);
  sc.g.analytical = d->gradient_mode;
  if (int e = make_hash_desc(sc.h, d->table, d->table_dtype, d->n_levels, d->log2_T, d->n_features, d->level_res,
                             d->level_w))
    return e;
  sc.oc = make_occ_desc(d->occ_mask, d->occ_mask ? d->occ_res : 1, d->occ_box);
  sc.mlp = d->mlp;
  sc.s = d->s;
  sc.sphere_r = d->sphere_radius;
  memcpy(out, &sc, sizeof(SceneDev));
  return 0;
}

static int serve_shape(const vxn_serve_render_args& A, int& steps, int& per, int& S) {
  VXN_REQUIRE(A.capacity >= 0, VXN_ERR_SHAPE, "capacity must be >= 0");
  VXN_REQUIRE(A.n_samples >= 2, VXN_ERR_SHAPE, "n_samples must be >= 2");
  VXN_REQUIRE(A.up_sample_steps >= 0 && A.n_importance >= 0, VXN_ERR_VALUE, "negative importance settings");
  steps = (A.n_importance > 0 && A.up_sample_steps > 0) ? A.up_sample_steps : 0;
  per = steps ? (A.n_importance / steps > 1 ? A.n_importance / steps : 1) : 0;
  S = A.n_samples + steps * per;
  VXN_REQUIRE(S <= kRT, VXN_ERR_UNSUPPORTED, "at most %d samples per ray are supported (got %d)", kRT, S);
  VXN_REQUIRE(A.capacity * (int64_t)S < ((int64_t)1 << 31), VXN_ERR_SHAPE, "capacity * samples must be below 2^31");
  VXN_REQUIRE(per <= 64 && steps <= 20, VXN_ERR_UNSUPPORTED, "at most 64 samples per step and 20 steps");
  return 0;
}

extern "C" int vxn_serve_render_workspace_bytes(const vxn_serve_render_args* a, int64_t* bytes) {
  VXN_REQUIRE(a != nullptr && bytes != nullptr, VXN_ERR_NULL, "null argument");
  int steps, per, S;
  if (int e = serve_shape(*a, steps, per, S)) return e;
  const int64_t plane = align256(a->capacity * (int64_t)S * (int64_t)sizeof(float));
  *bytes = (a->mlp_mode == VXN_MLP_BF16_MFMA ? 3 + 16 : 3) * plane;  // z, dists, mid (+ 16 feature planes)
  return 0;
}

extern "C" int vxn_serve_render(const vxn_serve_render_args* p, void* stream) {
  VXN_REQUIRE(p != nullptr, VXN_ERR_NULL, "args is null");
  const vxn_serve_render_args& A = *p;
  int steps, per, S;
  if (int e = serve_shape(A, steps, per, S)) return e;
  if (A.capacity == 0) return 0;
  VXN_REQUIRE(A.scenes && A.ray_scene && A.n_rays && A.rays_o && A.rays_d && A.t_lin && A.rgb, VXN_ERR_NULL,
              "null argument");
  VXN_REQUIRE(A.n_scenes >= 1, VXN_ERR_VALUE, "n_scenes must be >= 1");
  VXN_REQUIRE(!steps || A.u_det, VXN_ERR_NULL, "u_det is required");
  VXN_REQUIRE(A.bound_radius > 0.0f, VXN_ERR_VALUE, "bound_radius must be > 0");
  VXN_REQUIRE(A.mlp_mode == VXN_MLP_FP32 || A.mlp_mode == VXN_MLP_BF16_MFMA, VXN_ERR_VALUE, "unknown mlp_mode %d",
              A.mlp_mode);
  const int64_t plane = align256(A.capacity * (int64_t)S * (int64_t)sizeof(float));
  const bool bf = A.mlp_mode == VXN_MLP_BF16_MFMA;
  const int64_t need = (bf ? 3 + 16 : 3) * plane;
  VXN_REQUIRE(A.workspace != nullptr && A.workspace_bytes >= need, VXN_ERR_WORKSPACE,
              "serve workspace needs %lld bytes", (long long)need);
  float* z = (float*)A.workspace;
  float* dists = (float*)((char*)A.workspace + plane);
  float* mid = (float*)((char*)A.workspace + 2 * plane);
  const SceneDev* scenes = (const SceneDev*)A.scenes;
  hipStream_t s = (hipStream_t)stream;
  if (int e = launch_serve_upsample(scenes, A.n_scenes, A.ray_scene, A.n_rays, A.capacity, A.rays_o, A.rays_d,
                                    A.n_samples, per, steps, A.t_lin, A.u_det, A.bound_radius, z, dists, mid, s))
    return e;
  RenderDev a{};
  a.o = A.rays_o;
  a.d = A.rays_d;
  a.mid = mid;
  a.dist = dists;
  a.R = A.capacity;
  a.S = S;
  a.sdiv = vxn_fastdiv_make((uint32_t)S);
  a.cos_anneal = A.cos_anneal_ratio;
  a.iter_cos = A.use_iter_cos;
  a.rgb = A.rgb;
  a.depth = A.depth;
  a.opacity = A.opacity;
  a.normal = A.normal;
  a.scenes = scenes;
  a.ray_scene = A.ray_scene;
  a.n_rays_dev = A.n_rays;
  a.n_scenes = A.n_scenes;
  const int nb = render_blocks(A.capacity);
  if (bf) {
    // the BF16 form gathers the hash features level-major, XCD-aware, ahead of the render
    a.hfeat = (char*)A.workspace + 3 * plane;
    const int64_t chunks = (A.capacity * (int64_t)S + 255) / 256;
    VXN_REQUIRE(16 * chunks < ((int64_t)1 << 31), VXN_ERR_SHAPE, "batch too large");
    hipLaunchKernelGGL(hash_features_kernel<1>, dim3((unsigned)(16 * chunks)), dim3(256), 0, s, a, chunks);
    VXN_LAUNCH_CHECK();
  }
  if (bf)
    hipLaunchKernelGGL((render_kernel<kFwd, kMlpBf16, 1, 16>), dim3(nb), dim3(kRT), 0, s, a);
  else
    hipLaunchKernelGGL((render_kernel<kFwd, kMlpF32, 1, 16>), dim3(nb), dim3(kRT), 0, s, a);
  VXN_LAUNCH_CHECK();
  return 0;
}
