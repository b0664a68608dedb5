/* C ABI of libvoxneus_hip.so: hand-written gfx950 (MI355X) kernels for the VoxNeuS
 * voxel-SDF rendering and regularisation path.
 *
 * Conventions (all entry points):
 *   - return 0 on success or a negative VXN_ERR_* code; vxn_last_error() returns the
 *     calling thread's last message;
 *   - every pointer argument that names device memory must be a HIP device pointer;
 *     arguments documented "host" are read on the host before launch;
 *   - all work is enqueued on `stream` (a hipStream_t; 0 = the null stream) and is
 *     asynchronous: no entry point synchronises, allocates or frees device memory.
 *     Scratch memory is passed in by the caller as `workspace`, sized by the matching
 *     *_workspace_bytes() query;
 *   - gradient outputs documented "accumulated" are added to (callers zero them first);
 *     all other outputs are overwritten.
 *
 * Citations name the reference interface each entry point replaces (paths relative to
 * the reference checkout's Code/ directory).
 */
#ifndef VOXNEUS_HIP_H
#define VOXNEUS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VXN_OK = 0,
  VXN_ERR_NULL = -1,        /* a required pointer is null */
  VXN_ERR_SHAPE = -2,       /* a size or shape is out of the supported range */
  VXN_ERR_VALUE = -3,       /* an argument value is invalid */
  VXN_ERR_UNSUPPORTED = -4, /* a configuration this library does not implement */
  VXN_ERR_WORKSPACE = -5,   /* workspace missing or too small */
  VXN_ERR_HIP = -6          /* a HIP runtime call failed */
};

enum { VXN_DTYPE_F32 = 0, VXN_DTYPE_BF16 = 1 };
enum { VXN_MLP_FP32 = 0, VXN_MLP_BF16_MFMA = 1 };

#define VXN_MAX_LEVELS 32
#define VXN_MLP_IN 51
#define VXN_MLP_HIDDEN 64
#define VXN_MLP_PARAMS (64 * 51 + 64 + 3 * 64 + 3) /* W0, b0, W1, b1 flattened in that order */
/* View embedding of the radiance MLP (VoxNeuS view_embed_type, models/voxneus.py:78-82):
 * degree-4 spherical harmonics (16 values, W0 [64,51]) or the NeRF positional encoding
 * with 4 frequencies and the input (27 values, W0 [64,62], VXN_MLP_PARAMS_PE). */
#define VXN_VIEW_SH 0
#define VXN_VIEW_PE 1
#define VXN_MLP_PARAMS_PE (64 * 62 + 64 + 3 * 64 + 3)
/* Colour loss of the training step (trainer.py:213-236, loss_type). */
#define VXN_LOSS_HUBER 0
#define VXN_LOSS_MSE 1

/* Surface gradient of the dense grid (DenseSDFGrid.query_normals gradient_mode,
 * models/sdf_grid.py:122-160): the Eq. 7 interpolated central difference or the Eq. 4
 * derivative of the trilinear interpolant (the ablation baseline).                   */
#define VXN_GRADIENT_INTERPOLATED 0
#define VXN_GRADIENT_ANALYTICAL 1

const char* vxn_version(void);
const char* vxn_last_error(void);

/* ---- Dense SDF grid queries -------------------------------------------------------
 * grid: [X,Y,Z] float32, z fastest.  box6 (host): xyz_min[3], xyz_max[3].            */

/* Trilinear SDF (F.grid_sample, border, align_corners) + the surface gradient in
 * gradient_mode (VXN_GRADIENT_*), its normalisation, and optionally the 8 corner indices
 * [n,8,3] int64 and trilinear weights [n,8].  Replaces DenseSDFGrid.query_sdf +
 * query_normals (models/sdf_grid.py:90-160, geometry/gradients.py:103-220). Outputs are
 * nullable.                                                                           */
int vxn_grid_query_fwd(const float* grid, int X, int Y, int Z, const float* box6, const float* xyz, int64_t n,
                       float* sdf, float* grad, float* normal, int64_t* corners, float* corner_weights,
                       int32_t gradient_mode, void* stream);

/* Backward of the above w.r.t. the grid: d_sdf [n] and/or d_grad [n,3] (dL w.r.t. the
 * un-normalised gradient) are scattered into grad_grid (accumulated).               */
int vxn_grid_query_bwd(const float* grid, int X, int Y, int Z, const float* box6, const float* xyz, int64_t n,
                       const float* d_sdf, const float* d_grad, float* grad_grid, int32_t gradient_mode,
                       void* stream);

/* Eq. 7 on continuous grid coordinates [n,3] (already in [0,res-1]); eps3 (host) is the
 * voxel size.  Replaces _backend.gradient_interp_forward / _backward
 * (cuda/gradient_interp.cu:154-212, bound by cuda/adam_upd.cpp:118-125).  The backward
 * accumulates into grad_grid.                                                        */
int vxn_gradient_interp_fwd(const float* grid, int X, int Y, int Z, const float* coords, int64_t n,
                            const float* eps3, float* out, void* stream);
int vxn_gradient_interp_bwd(const float* grad_out, const float* coords, int64_t n, float* grad_grid, int X, int Y,
                            int Z, const float* eps3, void* stream);

/* Eq. 4 analytical gradient [n,3] and its unit normal [n,3] (both nullable) on continuous
 * grid coordinates.  Replaces compute_analytical_gradient (geometry/gradients.py:142-220);
 * the backward accumulates dL/dgrad [n,3] into grad_grid, as autograd through its corner
 * gather does.                                                                        */
int vxn_analytical_gradient_fwd(const float* grid, int X, int Y, int Z, const float* coords, int64_t n,
                                const float* eps3, float* grad, float* normal, void* stream);
int vxn_analytical_gradient_bwd(const float* grad_out, const float* coords, int64_t n, float* grad_grid, int X,
                                int Y, int Z, const float* eps3, void* stream);

/* Corner indices [n,8,3] int64 and trilinear weights [n,8] (both nullable) of continuous
 * grid coordinates [n,3].  Replaces get_voxel_corners_and_weights
 * (geometry/coordinates.py:55-124).                                                   */
int vxn_voxel_corners(const float* coords, int64_t n, int X, int Y, int Z, int64_t* corners, float* corner_weights,
                      void* stream);

/* SDF value and unit normal (in gradient_mode) at every point of an NX*NY*NZ lattice
 * spanning lattice_box6 (host), written [NX,NY,NZ] and [NX,NY,NZ,3].  The mesh-extraction
 * query of geometry/mesh.py:80-114 (compute_mesh_normals) over a full lattice.       */
int vxn_lattice_query(const float* grid, int X, int Y, int Z, const float* box6, int NX, int NY, int NZ,
                      const float* lattice_box6, float* sdf, float* normal, int32_t gradient_mode, void* stream);

/* Progressive grid super-resolution: grid [X,Y,Z] -> out [NX,NY,NZ] by trilinear
 * interpolation with aligned corners (F.interpolate, mode="trilinear",
 * align_corners=True), in PyTorch's CPU operation order.  Replaces upsample_sdf_grid
 * (geometry/upsample.py:12-39) under DenseSDFGrid.scale_volume_grid
 * (models/sdf_grid.py:183-195); out must not overlap grid.                           */
int vxn_upsample_grid(const float* grid, int X, int Y, int Z, float* out, int NX, int NY, int NZ, void* stream);

/* Post-training smoothing of the grid before mesh extraction: a separable Gaussian
 * filter with edge replication, one 1-D correlation per axis (x, then y, then z), each
 * accumulated in double and rounded to float32 (scipy.ndimage.gaussian_filter with
 * mode="nearest").  weights (host, 2*radius+1 doubles, symmetric, radius <= 32) are the
 * correlation weights (scipy's normalised Gaussian of radius int(4 sigma + 0.5)).
 * Replaces apply_gaussian_filter_to_grid (geometry/mesh.py:17-31), the optional
 * pre-filter of extract_mesh_marching_cubes (mesh.py:56-57); out must not overlap grid.
 * Workspace: vxn_gaussian_filter_workspace_bytes (one grid of float32).               */
int vxn_gaussian_filter_workspace_bytes(int X, int Y, int Z, int64_t* bytes);
int vxn_gaussian_filter(const float* grid, int X, int Y, int Z, const double* weights, int radius, float* out,
                        void* workspace, int64_t workspace_bytes, void* stream);

/* ---- Mesh extraction -----------------------------------------------------------------
 * Marching cubes of the grid at `level` (value < level is inside), replacing the
 * scikit-image call of extract_mesh_marching_cubes (geometry/mesh.py:34-77): one vertex
 * per grid edge whose end values straddle the level, at the linear-interpolation
 * crossing, mapped to world space as index * (xyz_max - xyz_min) / (res - 1) + xyz_min
 * (mesh.py:71-73); triangles oriented with normals towards increasing values, from a
 * case table built on faces so the mesh is crack-free (vxn_mc_case_table returns it:
 * tri [256][16] edge triples, -1 terminated, ntri [256]; host only).
 * Two calls: vxn_mc_count writes counts[0] = vertices and counts[1] = triangles (device
 * int32) and leaves offsets in the workspace; vxn_mc_emit, on the same grid and
 * workspace, writes verts [V,3] float32 and faces [F,3] int32 (entries beyond the
 * capacities are dropped).                                                            */
int vxn_mc_case_table(int8_t* tri, uint8_t* ntri);
int vxn_mc_workspace_bytes(int X, int Y, int Z, int64_t* bytes);
int vxn_mc_count(const float* grid, int X, int Y, int Z, float level, void* workspace, int64_t workspace_bytes,
                 int32_t* counts, void* stream);
int vxn_mc_emit(const float* grid, int X, int Y, int Z, const float* box6, float level, void* workspace,
                int64_t workspace_bytes, float* verts, int64_t max_verts, int32_t* faces, int64_t max_faces,
                void* stream);

/* ---- Multiresolution hash encoding ------------------------------------------------
 * table: [L, 2^log2T, F] float32 (dtype 0, 8-byte aligned) or bfloat16 (dtype 1, 16-byte
 * aligned; VXN_ERR_VALUE otherwise); F must be 2.
 * level_res (host, L entries): floor(base * b^l) as computed by the reference;
 * level_w (host, nullable): the NeuS2 lambda weights (1 = active).
 * x01: [n,3] normalised coordinates, clamped to [0,1] inside.  Replaces
 * PyTorchHashGrid.forward (models/radiance_hash.py:71-166).  The backward accumulates
 * float32 gradients into grad_table [L,T,F] regardless of the table dtype.           */
int vxn_hash_fwd(const void* table, int dtype, int L, int log2T, int F, const int32_t* level_res,
                 const float* level_w, const float* x01, int64_t n, float* feats, void* stream);
int vxn_hash_bwd(int L, int log2T, int F, const int32_t* level_res, const float* level_w, const float* x01, int64_t n,
                 const float* d_feats, float* grad_table, void* stream);

/* ---- Progressive importance upsampler ---------------------------------------------
 * Replaces RaySampler.sample_multistep_sdf with up_sample / sample_pdf / cat_z_vals
 * (rendering/sampler.py:127-204, 243-320, 11-66) and OccupancyGrid.is_occupied
 * (rendering/occupancy_grid.py:121-147).                                             */
typedef struct vxn_sampler_args {
  const float* rays_o; /* [R,3] */
  const float* rays_d; /* [R,3] */
  const float* near;   /* [R] */
  const float* far;    /* [R] */
  int64_t n_rays;
  const float* grid;
  int32_t X, Y, Z;
  float box[6];
  const uint8_t* occ_mask; /* [occ_res^3] bool, nullable = no occupancy modulation */
  int32_t occ_res;
  float occ_box[6];
  int32_t n_samples, n_importance, up_sample_steps, perturb;
  const float* t_lin;  /* device [n_samples]: torch.linspace(0, 1, n_samples) */
  const float* u_det;  /* device [per_step]: torch.linspace(0.5/p, 1-0.5/p, p) */
  const float* t_rand; /* device [R, n_samples] stratified jitter (perturb only) */
  const float* u_rand; /* device [steps, R, per_step] inverse-CDF draws (perturb only) */
  float* z_vals;       /* [R, S], S = n_samples + steps*per_step */
  float* dists;        /* [R, S] */
  float* mid_z;        /* [R, S] */
  int32_t* inds;       /* nullable [steps, R, per_step]: searchsorted indices */
  float bound_radius;  /* when near and far are both null: near/far from this bounding sphere
                          (near_far_from_sphere, rendering/rays.py:102-128), computed in-kernel */
  uint32_t* vr_bitmask; /* nullable [ceil(X*Y*Z/32)]: bit set for all 8 corners of the cell of
                           every final sample (at o + d * mid_z), as vxn_render_args.vr_bitmask */
  uint32_t* vr_cells;   /* nullable [ceil(X*Y*Z/32)], all zero on entry, used with vr_bitmask when
                           Z is a multiple of 32: each sample sets one bit, its cell's (left
                           set), and a pass then writes vr_bitmask as those cells' corners: the
                           same bits as above, but vr_bitmask is overwritten, not added to (no
                           reference counterpart: fewer atomics for the same V_R) */
} vxn_sampler_args;

int vxn_upsample(const vxn_sampler_args* a, void* stream);

/* ---- Fused per-ray render (forward, backward, training step) ----------------------
 * One ray per work-group: per-sample SDF + Eq. 7 normal, hash features, the SH (or
 * positional-encoding) view embedding, the 51->64->3 (62->64->3) radiance MLP, NeuS Eq. 2
 * alpha, front-to-back transmittance and compositing.  Serving scenes use the SH form.  Replaces VoxNeuS.render passes 2-3 with forward_points and
 * volume_render_from_samples (models/voxneus.py:115-167, 226-258;
 * rendering/neus_alpha.py:12-81; rendering/volume_render.py:12-153).                 */
typedef struct vxn_render_args {
  const float* rays_o; /* [R,3] */
  const float* rays_d; /* [R,3] */
  const float* mid_z;  /* [R,S] */
  const float* dists;  /* [R,S] */
  int64_t n_rays;
  int32_t n_samples; /* S <= 128 */
  const float* grid;
  int32_t X, Y, Z;
  float box[6];
  const void* table; /* [L,T,F]; 8-byte aligned (float32) or 16-byte aligned (bfloat16) */
  int32_t table_dtype;
  int32_t n_levels, log2_T, n_features;
  int32_t level_res[VXN_MAX_LEVELS];
  float level_w[VXN_MAX_LEVELS];
  const float* mlp; /* [VXN_MLP_PARAMS] float32: W0 [64,51], b0 [64], W1 [3,64], b1 [3]; with
                       view_embed = VXN_VIEW_PE [VXN_MLP_PARAMS_PE], W0 [64,62] */
  int32_t mlp_mode;
  float s;
  float sphere_radius;
  float cos_anneal_ratio;
  int32_t use_iter_cos;
  /* forward outputs (nullable); the training entry also writes rgb when it is given */
  float* rgb;            /* [R,3] */
  float* depth;          /* [R] */
  float* normal;         /* [R,3] */
  float* opacity;        /* [R] */
  float* weights;        /* [R,S] */
  float* alpha;          /* [R,S] */
  float* transmittance;  /* [R,S] */
  float* sdf_samples;    /* [R,S] */
  float* grad_samples;   /* [R,S,3] */
  float* normal_samples; /* [R,S,3] */
  int64_t* corner_indices; /* [R,S,8,3] */
  /* backward inputs (nullable = zero) */
  const float* d_rgb;     /* [R,3] */
  const float* d_depth;   /* [R] */
  const float* d_normal;  /* [R,3] */
  const float* d_opacity; /* [R] */
  const float* d_weights; /* [R,S] */
  /* fused training loss: Huber(rgb, target_rgb, delta), mean over R*3 */
  const float* target_rgb; /* [R,3] */
  float huber_delta;
  /* gradient outputs */
  float* grad_grid;  /* [X,Y,Z], accumulated */
  float* grad_table; /* [L,T,F] float32, accumulated */
  float* grad_mlp;   /* [VXN_MLP_PARAMS] (PE: [VXN_MLP_PARAMS_PE]), overwritten */
  float* grad_s;     /* [1] dL/ds, overwritten */
  double* loss;      /* [1] Huber loss (training step only), overwritten */
  uint32_t* vr_bitmask; /* nullable [ceil(X*Y*Z/32)]: bit set for all 8 corners of every sample */
  void* workspace;
  int64_t workspace_bytes;
  /* nullable profiling hook: hipEvent_t[6] recorded on the stream before the hash-feature
     gather (BF16 form: the geometry pass and the gather) and after it, the render kernel, the grid gradient (the atomic scatter or the
     binned passes), the hash-table gradient (its binned passes or the atomic scatter) and
     the reductions (forward: only the first three) */
  void* const* kernel_events;
  int32_t gradient_mode; /* VXN_GRADIENT_*: the normal of every sample (VoxNeuS gradient_mode) */
  /* nullable hipEvent_t, recorded on `stream` after the hash-feature gather, just before
     the render kernel (with grid_binned: after the binned grid gradient's tile sums, whose
     plain read-add-writes of grad_grid must not overlap other writers, on the stream that
     runs them; the ordered edge pass that follows adds with atomics): a caller can start
     independent work on another stream from it, such as atomic adds into grad_grid (the
     fused training step starts its regularisers there; grad_grid is then the same on every
     run only up to the order of those atomics) */
  void* fork_event;
  /* nullable device float: the NeuS sharpness s read on the device in place of `s`, so a
     caller that keeps s on the device (s = clamp(exp(10 v)), models/variance.py:31) needs
     no host read */
  const float* s_dev;
  /* VXN_VIEW_SH (0) or VXN_VIEW_PE: the view embedding the MLP's W0 columns 32.. expect;
     grad_mlp then has VXN_MLP_PARAMS or VXN_MLP_PARAMS_PE entries */
  int32_t view_embed;
  /* training step only (trainer.py:213-245): the colour loss VXN_LOSS_HUBER (delta
     huber_delta) or VXN_LOSS_MSE, a mean over the rays' 3 channels; with target_mask
     (nullable device [R], the foreground mask) the mask-weighted sum over 3 * sum(mask)
     (the plain mean when the mask sums to 0), plus mask_weight * the mean binary cross
     entropy of clamp(opacity, 1e-6, 1 - 1e-6) against the mask.  loss receives the total;
     loss_parts (nullable device [2]) the colour loss and the mask BCE. */
  int32_t loss_type;
  const float* target_mask;
  float mask_weight;
  double* loss_parts;
  /* backward and training entries: nullable second stream (default priority) for the
     grid gradient (binned or atomic), which then runs beside the hash-table gradient.
     scatter_events (hipEvent_t[2], required with it) order it after the render kernel and
     make `stream` wait for it before the call returns, so every output is complete on
     `stream` as without it.  kernel_events[3] is then recorded on scatter_stream after the
     grid gradient, and the hash-table gradient's span is kernel_events[2] -> [4]. */
  void* scatter_stream;
  void* const* scatter_events;
  /* nullable device float: the variance v of SingleVarianceNetwork (models/variance.py:
     31-41).  With it the kernels evaluate s = clamp(exp(10 v), 1e-4, s_max) themselves (in
     place of s and s_dev); backward and training entries then also write s_out (nullable,
     that s) and d_variance (nullable) = dL/ds * e * [1e-4 <= e <= s_max] * 10 with
     e = exp(10 v), and with amplified (nullable) apply amplify_s_gradient
     (training/scheduler.py:14-33): a negative d_variance is multiplied by s_amp_k and
     amplified[0] records whether it was. */
  const float* variance;
  float s_max;
  float* s_out;
  float* d_variance;
  float s_amp_k;
  int32_t* amplified;
  /* backward and training entries: 1 = grad_table is overwritten (every entry, zero where
     no sample reaches it) instead of accumulated, so the caller need not clear it */
  int32_t grad_table_overwrite;
  /* backward and training entries: 1 = the grid gradient is voxel-binned instead of the
     atomic scatter (gradient_interp.cu:86-152): each run's brick of stencil weights is
     listed under the 32 x 32 x 16-vertex grid tiles it reaches, and each tile's sums are
     added into grad_grid with one plain read-add-write per vertex; the stencils at the grid
     border (and of the rare samples whose grid_sample cell is not their corner cell) stay
     atomic but are added by one wavefront in sample order, so the grid gradient of the
     call is the same on every run.  Grids up
     to 1024 vertices per axis and 4096 tiles; otherwise (and with 0) the atomic scatter.
     grad_grid must have no other writer during the call except atomic adds forked from
     fork_event (recorded after the tile sums). */
  int32_t grid_binned;
  /* nullable hipEvent_t: the binned grid gradient's read-add-writes of grad_grid (and its
     edge pass) wait for it, so other work adding into grad_grid on another stream (the
     fused training step's regularisers, started right after the sampler) is ordered before
     them instead of after fork_event */
  void* grid_wait_event;
} vxn_render_args;

/* Workspace of the backward and training entries (the forward needs none): per-sample
 * records of the hash-feature and geometry gradients, and the binned hash-table gradient
 * (4 records of 20 bytes per sample and level, 8 on levels with res >= 2047, plus the
 * per-bin offsets): about 1.4 GB for 2^20 samples over 16 levels.  Tables beyond 2^21
 * entries per level (16 levels) take the atomic scatter and no record space.  With
 * grid_binned, 200 more bytes per sample (a brick and up to 8 list entries per run) and
 * the tile count rows.  The size depends on n_rays, n_samples, n_levels, log2_T,
 * level_res, the grid size and grid_binned.                             */
int vxn_render_workspace_bytes(const vxn_render_args* a, int64_t* bytes);
int vxn_render_forward(const vxn_render_args* a, void* stream);
int vxn_render_backward(const vxn_render_args* a, void* stream);
int vxn_render_train(const vxn_render_args* a, void* stream);

/* ---- Active vertex set V_R and closed-form regularisers ---------------------------*/

/* Mark corner indices [n,3] int64 in a bitmask of ceil(X*Y*Z/32) words (accumulated).
 * The marking half of extract_active_vertices (regularization/active_vertices.py:12-50). */
int vxn_vr_mark_corners(const int64_t* corners, int64_t n, int X, int Y, int Z, uint32_t* bitmask, void* stream);

/* Compact the bitmask into ascending linear indices ix*Y*Z+iy*Z+iz (the torch.unique
 * order).  Bit 31 of an entry is set when its world position is NOT inside the relaxed
 * sphere |x| < relax_radius (training/trainer.py:260-266).  counts (device, [2]) gets
 * {all active, active inside}.  vr_list must hold X*Y*Z entries.                      */
int vxn_vr_workspace_bytes(int X, int Y, int Z, int64_t* bytes);
int vxn_vr_compact(const uint32_t* bitmask, int X, int Y, int Z, const float* box6, float relax_radius,
                   int32_t* vr_list, int32_t* counts, void* workspace, int64_t workspace_bytes, void* stream);

/* Fused closed-form eikonal (Eq. 10-11) and curvature (Eq. 12-14) gradients over V_R,
 * accumulated into grad_grid with weights w_eik / w_curv, plus both losses reduced in
 * FP64 (means over interior vertices) into losses[2] (device).  |V_R| is counts[1] when
 * non-zero (relaxed-sphere filter applied) else counts[0] (the filter's keep-all case).
 * eps3 (host) is the voxel size.  Replaces compute_eikonal_loss_and_gradient and
 * compute_curvature_loss_and_gradient (regularization/eikonal.py:13-171,
 * regularization/curvature.py:13-165).                                               */
int vxn_reg_workspace_bytes(int X, int Y, int Z, int64_t* bytes);
int vxn_reg_eik_curv(const float* grid, int X, int Y, int Z, const float* eps3, const int32_t* vr_list,
                     const int32_t* counts, float w_eik, float w_curv, float* grad_grid, double* losses,
                     void* workspace, int64_t workspace_bytes, void* stream);

/* Reference-shaped single regularisers over active [V,3] int64 vertices; eps3 (host) is
 * the voxel size.  Replace _backend.eikonal_grad / curvature_grad
 * (cuda/eikonal_grad.cu:71-101, cuda/curvature_grad.cu:73-103); accumulate into
 * grad_grid.                                                                          */
int vxn_eikonal_grad(const float* grid, int X, int Y, int Z, const int64_t* active, int64_t V, float* grad_grid,
                     const float* eps3, float weight, void* stream);
int vxn_curvature_grad(const float* grid, int X, int Y, int Z, const int64_t* active, int64_t V, float* grad_grid,
                       const float* eps3, float weight, void* stream);

/* The same regularisers straight from the V_R bitmask the render marks, for the fused
 * step: counts[2] (device) receives {|V_R|, |V_R inside the relaxed sphere|} (box6 and
 * relax_radius as vxn_vr_compact), then work-groups take tiles of 8 x 8 (x, y) rows by
 * 64 z, test V_R membership in the bitmask, add each member's eikonal and curvature terms
 * into the tile's LDS accumulator (the tile and a one-vertex border), and add that to
 * grad_grid with one device atomic per reached vertex per tile.  Same
 * losses, gradient and |V_R| normalisation as vxn_reg_eik_curv on the compacted list;
 * replaces
 * extract_active_vertices + the relaxed-sphere filter + both regularisers
 * (regularization/active_vertices.py:12-50, trainer.py:260-266, eikonal.py:13-171,
 * curvature.py:13-165).  Workspace: vxn_reg_vr_workspace_bytes.                         */
int vxn_reg_vr_workspace_bytes(int X, int Y, int Z, int64_t* bytes);
int vxn_reg_eik_curv_vr(const float* grid, int X, int Y, int Z, const float* eps3, const float* box6,
                        float relax_radius, const uint32_t* bitmask, float w_eik, float w_curv, float* grad_grid,
                        double* losses, int32_t* counts, void* workspace, int64_t workspace_bytes, void* stream);

/* No counterpart in the reference.  vxn_reg_eik_curv_vr for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving or training loops, harnesses
 * that forbid hipMalloc): the same kernels and results, with the partial sums in a static
 * device array of 176 KiB per GPU (held until the process ends) instead of a workspace.
 * Grids of at most VXN_REG_NOALLOC_MAX_VOXELS vertices (512^3); larger ones return
 * VXN_ERR_SHAPE.  Allocates nothing; every call on one GPU shares the array, so calls
 * that overlap on one GPU (other streams, other threads) race.                          */
#define VXN_REG_NOALLOC_MAX_VOXELS (1ll << 27)
int vxn_reg_eik_curv_vr_noalloc(const float* grid, int X, int Y, int Z, const float* eps3, const float* box6,
                                float relax_radius, const uint32_t* bitmask, float w_eik, float w_curv,
                                float* grad_grid, double* losses, int32_t* counts, void* stream);

/* ---- Alpha and compositing of per-sample values the caller holds ------------------ */

/* volume_render_from_samples (rendering/volume_render.py:57-153): per ray of n_samples
 * (<= 128) samples, the Eq. 2 alpha from sdf / normals / ray_dirs / dists / s (neus_alpha.py:
 * 12-81; iter_cos annealing with use_iter_cos), the unit-sphere mask from pts (nullable),
 * transmittance, the T < 1e-4 cutoff and the composites rgb [R,3], depth [R,1] (from
 * z_vals), normal [R,3], opacity [R,1], weights / alpha / transmittance [R,S] (outputs
 * nullable).  With alpha_in [R,S] (and sdf = null) the alpha is an input: composite_weights
 * (volume_render.py:12-40).  colors, normals and z_vals are nullable (read as zero).
 * The backward takes the upstream gradients (nullable = zero) of every output and writes
 * d_colors, d_sdf, d_normals, d_alpha_in (nullable) and d_s (one float, overwritten, summed
 * in a fixed order; workspace vxn_composite_bwd_workspace_bytes).                       */
typedef struct vxn_composite_args {
  int64_t n_rays;
  int32_t n_samples;
  const float* colors;   /* [R,S,3] */
  const float* sdf;      /* [R,S] */
  const float* normals;  /* [R,S,3] */
  const float* ray_dirs; /* [R,3], or [R,S,3] with dirs_per_sample */
  int32_t dirs_per_sample;
  const float* z_vals;   /* [R,S] */
  const float* dists;    /* [R,S] */
  const float* s;        /* device scalar */
  const float* pts;      /* [R,S,3] */
  const float* alpha_in; /* [R,S] */
  float sphere_radius;
  float cos_anneal_ratio;
  int32_t use_iter_cos;
  float* rgb;
  float* depth;
  float* normal;
  float* opacity;
  float* weights;
  float* alpha;
  float* transmittance;
  const float* d_rgb;
  const float* d_depth;
  const float* d_normal;
  const float* d_opacity;
  const float* d_weights;
  const float* d_alpha;
  const float* d_transmittance;
  float* d_colors;
  float* d_sdf;
  float* d_normals;
  float* d_alpha_in;
  float* d_s;
} vxn_composite_args;
int vxn_composite_fwd(const vxn_composite_args* a, void* stream);
int vxn_composite_bwd_workspace_bytes(int64_t n_rays, int64_t* bytes);
int vxn_composite_bwd(const vxn_composite_args* a, void* workspace, int64_t workspace_bytes, void* stream);

/* compute_neus_alpha (rendering/neus_alpha.py:12-81) elementwise over n samples: sdf [n],
 * normals / dirs [n,3], dists [n], s (device scalar); the backward writes d_sdf, d_normals
 * (nullable) and d_s (nullable, one float, overwritten; workspace
 * vxn_neus_alpha_workspace_bytes).                                                      */
int vxn_neus_alpha_fwd(const float* sdf, const float* normals, const float* dirs, const float* dists, const float* s,
                       int64_t n, float cos_anneal_ratio, int32_t use_iter_cos, float* alpha, void* stream);
int vxn_neus_alpha_workspace_bytes(int64_t n, int64_t* bytes);
int vxn_neus_alpha_bwd(const float* sdf, const float* normals, const float* dirs, const float* dists, const float* s,
                       int64_t n, float cos_anneal_ratio, int32_t use_iter_cos, const float* d_alpha, float* d_sdf,
                       float* d_normals, float* d_s, void* workspace, int64_t workspace_bytes, void* stream);

/* The radiance MLP one point at a time: rgb [n,3] = sigmoid(W1 relu(W0 x + b0) + b1) with
 * x = [feats (n,32), view (n,view_dim: 16 SH or 27 positional encoding), normals (n,3)]
 * and mlp = W0 [64][35 + view_dim], b0, W1 [3][64], b1 flattened; mlp_mode
 * VXN_MLP_BF16_MFMA takes the BF16 form's rounding points.  Replaces RadianceMLP.forward
 * (models/radiance_mlp.py:81-99) on VoxNeuS.forward_points (models/voxneus.py:115-167).
 * The backward writes dL/dfeats, dL/dview, dL/dnormals (each nullable) and overwrites
 * d_mlp with the gradient summed over the points in a fixed order; workspace
 * vxn_point_mlp_bwd_workspace_bytes.                                                    */
int vxn_point_mlp_fwd(const float* feats, const float* view, int view_dim, const float* normals, const float* mlp,
                      int64_t n, int mlp_mode, float* rgb, void* stream);
int vxn_point_mlp_bwd_workspace_bytes(int64_t n, int view_dim, int64_t* bytes);
int vxn_point_mlp_bwd(const float* feats, const float* view, int view_dim, const float* normals, const float* mlp,
                      int64_t n, int mlp_mode, const float* d_rgb, float* d_feats, float* d_view, float* d_normals,
                      float* d_mlp, void* workspace, int64_t workspace_bytes, void* stream);

/* s-gradient amplification on the device: grad_v *= k if grad_v < 0.  amplified
 * (device, nullable) receives 1 or 0.  Replaces amplify_s_gradient
 * (training/scheduler.py:14-33) without its host read-back.                           */
int vxn_s_amplify(float* grad_v, float k, int32_t* amplified, void* stream);

/* Post-upsample eikonal clamp: dst[i] += g[i] * c with g = src * weight and
 * c = (1 / max|g|) * clamp when max|g| > clamp, else c = 1 (a NaN maximum leaves c = 1),
 * in the reference's float32 operation order.  src is the eikonal gradient computed with
 * weight 1 into its own buffer; dst is grid.grad.  dst is read and written without
 * atomics, so nothing may write it concurrently; dst and src must not overlap.  Replaces
 * the cooldown branch of VoxNeuSTrainer.train_step (training/trainer.py:309-316), which
 * trainer.py:131-164 arms for 500 steps after each grid super-resolution.  Workspace:
 * vxn_eik_clamp_add_workspace_bytes.                                                   */
int vxn_eik_clamp_add_workspace_bytes(int64_t* bytes);
int vxn_eik_clamp_add(float* dst, const float* src, int64_t n, float weight, float clamp, void* workspace,
                      int64_t workspace_bytes, void* stream);

/* ---- Occupancy grid -----------------------------------------------------------------
 * OccupancyGrid.update_from_sdf (rendering/occupancy_grid.py:62-119), called every 16
 * steps by the trainer (trainer.py:171-173) through VoxNeuS.update_occupancy_grid
 * (models/voxneus.py:105-113): at cell centres of the res^3 occupancy grid over
 * occ_box6 (host), the SDF grid's logistic density s sigmoid(s d)(1 - sigmoid(s d)),
 * raised to 2 threshold within 2 |voxel step| of the surface, replaces the cell's
 * density by max(decay * previous, new) (the new value alone when first_update) and
 * sets mask = density > threshold.  density [res^3] float32, mask [res^3] uint8.
 * With idx_x/y/z (device int64 [n_cells] each, as the reference's torch.randint draws)
 * only those cells are updated (cells drawn twice get the same value; indices outside
 * [0, res) are skipped) and a workspace of vxn_occupancy_update_workspace_bytes is
 * needed; with null indices every cell is updated once (a full sweep, no workspace).  */
int vxn_occupancy_update_workspace_bytes(int64_t n_cells, int64_t* bytes);
int vxn_occupancy_update(float* density, uint8_t* mask, int32_t res, const float* occ_box6, const float* grid,
                         int32_t X, int32_t Y, int32_t Z, const float* box6, float s, float decay, float threshold,
                         int32_t first_update, const int64_t* idx_x, const int64_t* idx_y, const int64_t* idx_z,
                         int64_t n_cells, void* workspace, int64_t workspace_bytes, void* stream);

/* ---- Optimizer updates ------------------------------------------------------------
 * float32 tensors of n elements, 16-byte aligned; step is the update count including
 * this one (the reference's state["step"] after its increment).                      */

/* Dense and masked (grad != 0 only) Adam of the native backend: _backend.adam_upd /
 * masked_adam_upd (cuda/adam_upd.cpp:48-72, cuda/adam_upd_kernel.cu:6-82), with
 * step_size = lr sqrt(1 - beta2^t) / (1 - beta1^t) formed in float as there.  The masked
 * form is SparseGridAdam.step on a HIP tensor (training/optimizer.py:75-90); n_updated
 * (device, nullable) accumulates the number of elements updated (its
 * last_updated_count, without the host read-back).                                   */
int vxn_adam_upd(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n, int32_t step,
                 float beta1, float beta2, float lr, float eps, void* stream);
int vxn_masked_adam_upd(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n, int32_t step,
                        float beta1, float beta2, float lr, float eps, int32_t* n_updated, void* stream);

/* SparseGridAdam.step with weight decay (training/optimizer.py:92-113: the masked
 * PyTorch formula, eps added after dividing by sqrt(1 - beta2^t)).  The scalars are the
 * Python floats (double), rounded to float where they meet a tensor, as PyTorch does.  */
int vxn_sparse_adam_decay(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                          int64_t step, double beta1, double beta2, double lr, double eps, double weight_decay,
                          int32_t* n_updated, void* stream);

/* torch.optim.Adam (no weight decay, no amsgrad) over up to 8 tensors in one launch: the
 * network optimizer of create_optimizer (training/optimizer.py:143-150; hash table, MLP
 * and variance, each its own parameter group and learning rate), in the order of
 * PyTorch's multi-tensor update.  A tensor whose lr_flag (device int32, nullable) is
 * nonzero uses lr * lr_flag_scale: the s-amplified variance learning rate
 * (trainer.py:332) chosen on the device.  param_min clamps the updated parameter from
 * below (the variance floor, trainer.py:346-353; -INFINITY for none).  param_bf16
 * (nullable) receives the updated parameter rounded to bfloat16 (a bf16 working copy of
 * an FP32 master).                                                                     */
typedef struct vxn_adam_tensor {
  float* param;
  const float* grad;
  float* exp_avg;
  float* exp_avg_sq;
  int64_t n;
  double lr;
  const int32_t* lr_flag;
  double lr_flag_scale;
  float param_min;
  void* param_bf16;
} vxn_adam_tensor;

int vxn_adam_step(const vxn_adam_tensor* tensors, int32_t n_tensors, int64_t step, double beta1, double beta2,
                  double eps, void* stream);

/* ---- Serving: many resident scenes, packed request batches -------------------------
 * The reference renders one scene in 2048-ray chunks (scripts/render.py:35-54:
 * near_far_from_sphere + VoxNeuS.render, no_grad, det sampler) and queries SDF and
 * normals in 65536-point batches (geometry/mesh.py:80-114); it has no request loop.
 * Here a batch packs rays (or points) of many requests on many resident scenes: a device
 * table holds one packed scene record per scene (vxn_serve_pack_scene), a per-ray scene
 * index selects it, and the batch size is read from device memory so that one captured
 * HIP graph serves batches of any size up to the capacity.  Each ray's result does not
 * depend on the other rays of the batch.                                              */
typedef struct vxn_scene_desc {
  const float* grid; /* [X,Y,Z] float32 */
  int32_t X, Y, Z;
  float box[6];               /* xyz_min, xyz_max */
  const uint8_t* occ_mask;    /* nullable [r,r,r] bool occupancy (sampler modulation) */
  int32_t occ_res;
  float occ_box[6];
  const void* table;          /* [L,T,2] float32 (8-byte aligned) or bfloat16 (16-byte aligned) */
  int32_t table_dtype;
  int32_t n_levels, log2_T, n_features;
  int32_t level_res[VXN_MAX_LEVELS];
  float level_w[VXN_MAX_LEVELS];
  const float* mlp;           /* [VXN_MLP_PARAMS] float32 */
  float s;                    /* NeuS sharpness */
  float sphere_radius;        /* compositing inside-sphere radius (1.0 in the reference) */
  int32_t gradient_mode;      /* VXN_GRADIENT_* of the scene's model */
} vxn_scene_desc;

/* Size of one packed scene record of the device table. */
int vxn_serve_scene_bytes(int64_t* bytes);
/* Validates a scene and writes its packed record into host memory `out` (the caller
 * copies records to the device table, record i at byte offset i * vxn_serve_scene_bytes). */
int vxn_serve_pack_scene(const vxn_scene_desc* scene, void* out, int64_t out_bytes);

typedef struct vxn_serve_render_args {
  const void* scenes;        /* device table of packed scene records */
  int32_t n_scenes;
  const int32_t* ray_scene;  /* [capacity] scene of each ray (outside [0, n_scenes): zeros) */
  const int32_t* n_rays;     /* device scalar: rays in this batch (clamped to capacity) */
  int64_t capacity;
  const float* rays_o;       /* [capacity,3] */
  const float* rays_d;       /* [capacity,3] */
  int32_t n_samples, n_importance, up_sample_steps; /* 64, 64, 4 in the reference */
  const float* t_lin;        /* [n_samples] linspace(0, 1) */
  const float* u_det;        /* [n_importance / steps] the deterministic sample_pdf u */
  float bound_radius;        /* near/far from this sphere (1.0 in the reference) */
  int32_t mlp_mode;
  float cos_anneal_ratio;
  int32_t use_iter_cos;
  float* rgb;                /* [capacity,3] */
  float* depth;              /* nullable [capacity] */
  float* opacity;            /* nullable [capacity] */
  float* normal;             /* nullable [capacity,3] */
  void* workspace;           /* z, dists, mid of the batch */
  int64_t workspace_bytes;
} vxn_serve_render_args;

int vxn_serve_render_workspace_bytes(const vxn_serve_render_args* a, int64_t* bytes);
/* Sampler + fused render forward over a packed batch (two kernels, no allocation, no
 * synchronisation: capturable in a HIP graph). */
int vxn_serve_render(const vxn_serve_render_args* a, void* stream);
/* SDF and unit normal at a packed batch of points: pt_scene [capacity], n_points device
 * scalar, xyz [capacity,3] -> sdf [capacity], normal [capacity,3] (nullable). */
int vxn_serve_query(const void* scenes, int32_t n_scenes, const int32_t* pt_scene, const int32_t* n_points,
                    int64_t capacity, const float* xyz, float* sdf, float* normal, void* stream);

/* Moves rows between request ranges of a device arena and the batch buffers of a step.
 * segs (device, [max_segs][3] int32): arena row offset, row count, scene; n_segs (device
 * scalar) segments are used, in order, filling batch rows 0.. up to the capacity.
 * to_batch = 1: batch0/1[i] = arena0/1[row(i)] (w0/w1 floats per row), batch_scene[i] =
 * the segment's scene and *batch_n = the row total; to_batch = 0: arena0/1[row(i)] =
 * batch0/1[i].  Rows outside [0, arena_rows) read as zero (scene -1) and are not written. */
int vxn_serve_move(const int32_t* segs, const int32_t* n_segs, int32_t max_segs, int64_t capacity,
                   int64_t arena_rows, int32_t to_batch, float* arena0, float* batch0, int32_t w0, float* arena1,
                   float* batch1, int32_t w1, int32_t* batch_scene, int32_t* batch_n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VOXNEUS_HIP_H */
