"""The fused VoxNeuS training step (fwd + bwd + regularisers) on one GPU.

``FusedTrainStep.step`` follows ``VoxNeuSTrainer.train_step`` (training/trainer.py:188-326)
in order - ray bounds, multi-step sampling, render, Huber colour loss, backward, V_R
with the relaxed-sphere filter, closed-form eikonal + curvature gradients, and the k=5
negative s-gradient amplification - without any host synchronisation: losses, counts
and the amplification flag stay on the device until the caller reads them.

``amplify_s_gradient`` mirrors training/scheduler.py:14-33 (it returns a bool, which
needs one device read; ``FusedTrainStep`` uses the device-side form instead).
"""

from __future__ import annotations

import ctypes
import dataclasses
from dataclasses import dataclass
from typing import Dict, Optional

import torch

from . import _native as N
from . import functional as Fn


def near_far_from_sphere(rays_o, rays_d, radius=1.0, default_near=0.05, default_far=4.0):
    """rendering/rays.py:102-128 (elementwise, on the rays' device)."""
    a = torch.sum(rays_d ** 2, dim=-1, keepdim=True)
    b = 2.0 * torch.sum(rays_o * rays_d, dim=-1, keepdim=True)
    mid = 0.5 * (-b) / a
    near = torch.clamp(mid - radius, min=default_near)
    far = torch.clamp(mid + radius, min=near + 0.1)
    return near, far


def get_rays(H: int, W: int, K, c2w: torch.Tensor, normalize_dir: bool = True, convention: str = "opencv"):
    """rendering/rays.py:11-69: rays through the pixel centres (+0.5) of a pinhole camera,
    (H, W, 3) origins and directions on c2w's device (ray setup, upstream of the path)."""
    dev = c2w.device
    K = torch.as_tensor(K, dtype=torch.float32, device=dev)
    i, j = torch.meshgrid(torch.arange(W, dtype=torch.float32, device=dev) + 0.5,
                          torch.arange(H, dtype=torch.float32, device=dev) + 0.5, indexing="xy")
    fx, fy, cx, cy = K[0, 0], K[1, 1], K[0, 2], K[1, 2]
    if convention.lower() == "opencv":
        dirs = torch.stack([(i - cx) / fx, (j - cy) / fy, torch.ones_like(i)], dim=-1)
    else:
        dirs = torch.stack([(i - cx) / fx, -(j - cy) / fy, -torch.ones_like(i)], dim=-1)
    rays_d = torch.matmul(dirs, c2w[:3, :3].T)
    if normalize_dir:
        rays_d = rays_d / torch.linalg.norm(rays_d, dim=-1, keepdim=True)
    return c2w[:3, 3].expand(rays_d.shape), rays_d


def amplify_s_gradient(variance_param: torch.nn.Parameter, k: float = 5.0) -> bool:
    """training/scheduler.py:14-33: grad <- k*grad when grad < 0; returns whether it fired."""
    if variance_param.grad is None:
        return False
    flag = torch.zeros(1, device=variance_param.device, dtype=torch.int32)
    g = variance_param.grad
    Fn.s_amplify(g.reshape(1), float(k), flag)
    return bool(flag.item())


def s_grad_to_v(grad_s: torch.Tensor, variance: torch.Tensor, s_max: float) -> torch.Tensor:
    """dL/dv from dL/ds for s = clamp(exp(10 v), 1e-4, s_max) (models/variance.py:38-41)."""
    e = torch.exp(variance.detach() * 10.0)
    inside = ((e >= 1e-4) & (e <= s_max)).to(e.dtype)
    return grad_s.reshape(()) * e * inside * 10.0


@dataclass
class StepConfig:
    n_samples: int = 64
    n_importance: int = 64
    up_sample_steps: int = 4
    perturb: bool = False
    sphere_radius: float = 1.0
    huber_delta: float = 0.1
    # trainer.py:56-59, 213-245: "Huber" or "MSE" colour loss; with use_mask and a target
    # mask passed to step(), the foreground-weighted colour loss plus mask_weight * BCE
    loss_type: str = "Huber"
    use_mask: bool = False
    mask_weight: float = 0.1
    w_eik: float = 0.1
    w_curv: float = 5e-4
    s_amp_k: float = 5.0
    amplify: bool = True
    relax_radius: float = 1.2
    mlp_mode: int = N.VXN_MLP_FP32
    active_levels: Optional[float] = None
    gradient_mode: str = "interpolated"  # or "analytical" (Eq. 4), as VoxNeuS(gradient_mode=...)
    view_embed_type: str = "sh"  # or "pe" (positional encoding; the MLP's W0 is [64, 62])
    # mark V_R in the sampler and run the regularisers on a second stream, beside the
    # render and the hash-table gradient
    concurrent_regularisers: bool = True
    # start them just before the render kernel, or with the binned grid gradient after it
    # (else right after the sampler, beside the hash-feature gather, the binned grid sum
    # waiting for them through vxn_render_args.grid_wait_event; measured slower on the
    # 8192-ray step: BF16 1.63 -> 1.72 ms, FP32 2.50 -> 2.58 ms)
    fork_after_features: bool = True
    # run the grid gradient on a third stream, beside the hash-table gradient (8192-ray
    # step: BF16 form 1.77 -> 1.73 ms with the atomic scatter, 1.68 ms binned; FP32 form
    # 2.57 -> 2.55 ms atomic)
    concurrent_scatter: bool = True
    # the grid gradient voxel-binned (vxn_render_args.grid_binned: the same on every run, one
    # plain read-add-write per vertex) instead of the atomic scatter.  None: binned for the
    # BF16 form, atomic for the FP32 form, the faster of the two for each beside the
    # hash-table gradient (FP32: 2.73 ms binned against 2.55 ms atomic)
    binned_grid_gradient: Optional[bool] = None
    # max-abs of the weighted eikonal gradient during a post-upsample cooldown
    # (trainer.py:118, scheduler.eik_grad_clamp)
    eik_grad_clamp: float = 0.1
    # NeuS iter_cos annealing of the alpha's cosine (neus_alpha.py:46-57; the trainer
    # passes min(step / 10000, 1) as the ratio, trainer.py:192)
    use_iter_cos: bool = False
    cos_anneal_ratio: float = 0.0


class FusedTrainStep:
    """Owns the device workspaces of one scene's training step.

    Parameters are plain device tensors: ``grid`` [X,Y,Z] f32, ``table`` [L,T,2] f32 or
    bf16, ``mlp`` [VXN_MLP_PARAMS] f32 (W0, b0, W1, b1 flattened; [VXN_MLP_PARAMS_PE] with
    the positional-encoding view embedding), ``variance`` scalar.
    Gradients land in ``grad_grid`` [X,Y,Z], ``grad_table`` [L,T,2] f32, ``grad_mlp``,
    ``grad_v`` (after amplification); ``loss`` / ``reg_losses`` are FP64 device scalars.
    """

    def __init__(self, grid: torch.Tensor, table: torch.Tensor, mlp: torch.Tensor, variance: torch.Tensor,
                 level_res, box=(-1.0, -1.0, -1.0, 1.0, 1.0, 1.0), s_max: float = 300.0,
                 occ_mask: Optional[torch.Tensor] = None, cfg: Optional[StepConfig] = None, occ_box=None):
        self.grid, self.table, self.mlp, self.variance = grid, table, mlp, variance
        self.level_res = list(level_res)
        self.box = tuple(float(v) for v in box)
        # the occupancy grid's own bounds (the model's, occupancy_grid.py:33-47)
        self.occ_box = self.box if occ_box is None else tuple(float(v) for v in occ_box)
        self.s_max = float(s_max)
        # a private copy: the schedules change the weights of this step's configuration
        self.cfg = dataclasses.replace(cfg) if cfg is not None else StepConfig()
        cfg = self.cfg
        dev = grid.device
        if occ_mask is None:
            # the reference always samples with an occupancy grid; a fresh one is all True
            # (occupancy_grid.py:51-54), i.e. factor 1 inside the box and 0.1 outside it
            occ_mask = torch.ones((128, 128, 128), device=dev, dtype=torch.bool)
        self.occ_mask = occ_mask
        self.occ_density = torch.zeros(occ_mask.shape, device=dev, dtype=torch.float32)
        self.occ_updates = 0
        X, Y, Z = grid.shape
        self.grad_grid = torch.zeros_like(grid)
        self.grad_table = torch.zeros(table.shape, device=dev, dtype=torch.float32)
        if mlp.numel() != N.mlp_params(N.view_embed_code(cfg.view_embed_type)):
            raise ValueError(f"mlp has {mlp.numel()} parameters; view_embed_type {cfg.view_embed_type!r} needs "
                             f"{N.mlp_params(N.view_embed_code(cfg.view_embed_type))}")
        self.grad_mlp = torch.zeros(mlp.numel(), device=dev, dtype=torch.float32)
        self.grad_s = torch.zeros(1, device=dev, dtype=torch.float32)
        self.grad_v = torch.zeros(1, device=dev, dtype=torch.float32)
        self.loss = torch.zeros(1, device=dev, dtype=torch.float64)
        self.loss_parts = torch.zeros(2, device=dev, dtype=torch.float64)  # colour loss, mask BCE
        self.rgb = None  # the step's rendered colours [R, 3] (for the PSNR of the metrics)
        self.amplified = torch.zeros(1, device=dev, dtype=torch.int32)
        self.bitmask = torch.zeros(Fn.vr_bitmask_words((X, Y, Z)), device=dev, dtype=torch.int32)
        # the sampler's per-cell V_R marks (whole bitmask words per grid column only)
        self.vr_cells = torch.zeros_like(self.bitmask) if Z % 32 == 0 else None
        self.vr_counts = torch.zeros(2, device=dev, dtype=torch.int32)
        nb = ctypes.c_int64(0)
        N.check(N.lib().vxn_reg_vr_workspace_bytes(X, Y, Z, ctypes.byref(nb)))
        self.reg_ws = torch.empty(int(nb.value), device=dev, dtype=torch.uint8)
        self.reg_losses = torch.zeros(2, device=dev, dtype=torch.float64)
        self.eps3 = Fn.voxel_size_of(self.box, (X, Y, Z))
        self._render_ws = None
        self._args = None
        self._args_key = None
        self.s_value = torch.zeros((), device=dev, dtype=torch.float32)  # s of the last step, on the device
        self.kernel_events = None  # optional N.KernelEvents(6) recorded between the render kernels
        self.side = N.SideStream(dev) if cfg.concurrent_regularisers else None
        self.scatter = N.SideStream(dev) if cfg.concurrent_scatter else None
        # steps left in the post-upsample cooldown (trainer.py:119, 163): while > 0 the
        # eikonal gradient goes to its own buffer and is clamped before it is added
        self.post_upsample_cooldown = 0
        self._g_eik = None
        self._clamp_ws = None
        self.level_w = None
        self.set_active_levels(cfg.active_levels)

    def set_active_levels(self, active_levels: Optional[float]) -> None:
        """The NeuS2 bandwidth lambda of the following steps (radiance_hash.py:90-102); None
        keeps every level.  The level weights go into each launch's arguments."""
        from .hash_grid import level_weights
        self.cfg.active_levels = active_levels
        self.level_w = level_weights(len(self.level_res), active_levels)

    def _render_args(self, rays_o, rays_d, mid, dists, target, target_mask=None):
        """The render arguments of this step.  The static part is built once.  The kernels
        read the variance v on the device and evaluate s = clamp(exp(10 v), 1e-4, s_max)
        (variance.py:31) themselves; the reductions write s, dL/dv and the amplification
        flag, so building the arguments reads nothing back from the GPU."""
        cfg = self.cfg
        R, S = mid.shape
        key = (R, S, mid.device, self.grid.data_ptr(), self.table.data_ptr(), self.mlp.data_ptr(),
               self.variance.data_ptr(), self._grid_binned())
        a = self._args
        if a is None or self._args_key != key:
            rc = Fn.RenderConfig(self.box, self.level_res, self.level_w, cfg.sphere_radius, cfg.cos_anneal_ratio,
                                 cfg.use_iter_cos, mlp_mode=cfg.mlp_mode,
                                 gradient_mode=N.gradient_mode_code(cfg.gradient_mode),
                                 view_embed=N.view_embed_code(cfg.view_embed_type),
                                 grid_binned=self._grid_binned())
            a = N.RenderArgs()
            Fn.fill_render_args(a, rc, self.grid, self.table, self.mlp, 0.0, rays_o, rays_d, mid, dists)
            a.variance = N.dptr(self.variance, torch.float32, "variance")
            a.s_max, a.s_out = float(self.s_max), N.dptr(self.s_value)
            a.d_variance = N.dptr(self.grad_v)
            a.grad_table_overwrite = 1
            a.huber_delta = cfg.huber_delta
            a.loss_type = N.loss_type_code(cfg.loss_type)
            a.mask_weight = float(cfg.mask_weight)
            a.loss_parts = N.dptr(self.loss_parts)
            a.grad_grid, a.grad_table = N.dptr(self.grad_grid), N.dptr(self.grad_table)
            a.grad_mlp, a.grad_s, a.loss = N.dptr(self.grad_mlp), N.dptr(self.grad_s), N.dptr(self.loss)
            need = ctypes.c_int64(0)
            N.check(N.lib().vxn_render_workspace_bytes(ctypes.byref(a), ctypes.byref(need)))
            if self._render_ws is None or self._render_ws.numel() < need.value:
                self._render_ws = torch.empty(int(need.value), device=mid.device, dtype=torch.uint8)
            a.workspace, a.workspace_bytes = self._render_ws.data_ptr(), self._render_ws.numel()
            self._args, self._args_key = a, key
        # per-step settings of the schedules
        for i in range(len(self.level_res)):
            a.level_w[i] = 1.0 if self.level_w is None else float(self.level_w[i])
        a.cos_anneal_ratio = float(cfg.cos_anneal_ratio)
        a.use_iter_cos = int(cfg.use_iter_cos)
        a.rays_o, a.rays_d = N.dptr(rays_o, torch.float32, "rays_o"), N.dptr(rays_d, torch.float32, "rays_d")
        a.mid_z, a.dists = N.dptr(mid, torch.float32, "mid_z"), N.dptr(dists, torch.float32, "dists")
        a.target_rgb = N.dptr(target, torch.float32, "target")
        if self.rgb is None or self.rgb.shape[0] != R:
            self.rgb = torch.empty((R, 3), device=mid.device, dtype=torch.float32)
        a.rgb = N.dptr(self.rgb)
        a.target_mask = N.dptr(target_mask, torch.float32, "target_mask")
        a.vr_bitmask = N.dptr(self.bitmask)
        # amplify_s_gradient (scheduler.py:14-33) on dL/dv in the reductions
        a.s_amp_k, a.amplified = float(cfg.s_amp_k), (N.dptr(self.amplified) if cfg.amplify else None)
        if self.scatter is not None:
            a.scatter_stream = self.scatter.handle
            a.scatter_events = ctypes.cast(self.scatter.events, ctypes.c_void_p)
        else:
            a.scatter_stream, a.scatter_events = None, None
        a.fork_event = None
        a.grid_wait_event = None
        a.kernel_events = (ctypes.cast(self.kernel_events.array, ctypes.c_void_p)
                           if self.kernel_events is not None else None)
        return a

    def _grid_binned(self) -> bool:
        b = self.cfg.binned_grid_gradient
        return self.cfg.mlp_mode == N.VXN_MLP_BF16_MFMA if b is None else bool(b)

    def sample(self, rays_o, rays_d, t_rand=None, u_rand=None, vr_bitmask=None, vr_cells=None):
        cfg = self.cfg
        # ray bounds from the unit sphere (trainer.py:188) are computed inside the sampler
        return Fn.upsample(rays_o, rays_d, None, None, self.grid, self.box, cfg.n_samples, cfg.n_importance,
                           cfg.up_sample_steps, cfg.perturb, self.occ_mask, self.occ_box, t_rand, u_rand,
                           bound_radius=1.0, vr_bitmask=vr_bitmask, vr_cells=vr_cells)

    def _cooldown(self) -> bool:
        return self.post_upsample_cooldown > 0 and self.cfg.w_eik > 0

    def _regularisers(self):
        # V_R (relaxed-sphere filtered) and both regularisers straight from the bitmask
        cfg = self.cfg
        if self._cooldown():
            # trainer.py:296-318: the eikonal gradient (weight 1) into its own buffer, the
            # curvature term into grad_grid; the clamped eikonal add follows the join
            self._g_eik.zero_()
            Fn.reg_eik_curv_vr(self.grid, self.eps3, self.box, cfg.relax_radius, self.bitmask, 1.0, 0.0, self._g_eik,
                               self.reg_ws, self.reg_losses, self.vr_counts)
            Fn.reg_eik_curv_vr(self.grid, self.eps3, self.box, cfg.relax_radius, self.bitmask, 0.0, cfg.w_curv,
                               self.grad_grid, self.reg_ws, self.reg_losses, self.vr_counts)
            return
        Fn.reg_eik_curv_vr(self.grid, self.eps3, self.box, cfg.relax_radius, self.bitmask, cfg.w_eik, cfg.w_curv,
                           self.grad_grid, self.reg_ws, self.reg_losses, self.vr_counts)

    def _eik_clamp(self):
        """trainer.py:310-316 on grad_grid, once the step's other writers of it are done."""
        Fn.eik_clamp_add(self.grad_grid, self._g_eik, self.cfg.w_eik, self.cfg.eik_grad_clamp, self._clamp_ws)
        self.post_upsample_cooldown -= 1

    def step(self, rays_o: torch.Tensor, rays_d: torch.Tensor, target: torch.Tensor, t_rand=None, u_rand=None,
             zs=None, target_mask: Optional[torch.Tensor] = None) -> Dict[str, torch.Tensor]:
        """One fwd+bwd step; gradients are overwritten.  Returns device tensors.
        ``target_mask`` ([R] or [R,1] foreground mask) is used when ``cfg.use_mask``."""
        cfg = self.cfg
        rays_o, rays_d, target = rays_o.contiguous(), rays_d.contiguous(), target.contiguous()
        if cfg.use_mask and target_mask is not None:
            target_mask = target_mask.reshape(-1).to(torch.float32).contiguous()
            if target_mask.numel() != rays_o.shape[0]:
                raise ValueError(f"target_mask has {target_mask.numel()} entries for {rays_o.shape[0]} rays")
        else:
            target_mask = None
        cooldown = self._cooldown()
        if cooldown and (self._g_eik is None or self._g_eik.shape != self.grid.shape):
            self._g_eik = torch.zeros_like(self.grid)
            self._clamp_ws = Fn.eik_clamp_workspace(self.grid.device)
        self.grad_grid.zero_()  # the render call overwrites grad_table
        # with the sampler in this step, it marks V_R and the regularisers run on the side
        # stream beside the render kernel and the scatters; both only add into grad_grid
        early = self.side is not None and zs is None
        # the sampler marks one bit per cell and then writes the bitmask whole from the cells
        # where the grid's columns hold whole words; otherwise the corners' bits
        cells = self.vr_cells if early else None
        (cells if cells is not None else self.bitmask).zero_()
        if zs is None:
            z, dists, mid = self.sample(rays_o, rays_d, t_rand, u_rand, self.bitmask if early else None, cells)
        else:
            z, dists, mid = zs
        main = Fn._stream(mid)
        a = self._render_args(rays_o, rays_d, mid.contiguous(), dists.contiguous(), target, target_mask)
        if early:
            a.vr_bitmask = None
            if cfg.fork_after_features:
                # recorded just before the render kernel (after the binned grid gradient,
                # whose plain read-add-writes of grad_grid the regularisers' atomics follow)
                a.fork_event = self.side.events[0]
            else:
                # right after the sampler, beside the hash-feature gather; the binned grid
                # gradient's read-add-writes of grad_grid wait for them
                self.side.fork(main)
                with torch.cuda.stream(self.side.stream):
                    self._regularisers()
                self.side.finish()
                a.grid_wait_event = self.side.events[1]
        N.check(N.lib().vxn_render_train(ctypes.byref(a), main), "render_train")
        if early:
            if cfg.fork_after_features:
                # queued after the render launch, so the render's kernels are dispatched first
                self.side.fork(None)
                with torch.cuda.stream(self.side.stream):
                    self._regularisers()
                self.side.finish()
        else:
            self._regularisers()
        if early:
            self.side.join(main)  # grad_grid, the losses and V_R counts are complete past this point
        if cooldown:
            self._eik_clamp()
        return {"loss": self.loss, "loss_parts": self.loss_parts, "reg_losses": self.reg_losses,
                "vr_counts": self.vr_counts, "rgb": self.rgb, "target": target, "target_mask": target_mask,
                "amplified": self.amplified, "z_vals": z, "dists": dists, "mid_z_vals": mid}

    def rescale_grid(self, new_resolution) -> None:
        """DenseSDFGrid.scale_volume_grid (sdf_grid.py:183-195) on this step's grid: the
        grid is replaced by its trilinear upsampling (aligned corners, on the device) and
        the gradient, V_R and regulariser buffers follow the new shape."""
        shape = (new_resolution,) * 3 if isinstance(new_resolution, int) else tuple(int(v) for v in new_resolution)
        self.adopt_grid(Fn.upsample_grid(self.grid, shape))

    def adopt_grid(self, grid: torch.Tensor) -> None:
        """Train ``grid`` ([X,Y,Z] f32 on the step's device) from now on; the gradient, V_R
        and regulariser buffers follow its shape."""
        dev = self.grid.device
        if grid.dim() != 3 or grid.dtype != torch.float32 or grid.device != dev or not grid.is_contiguous():
            raise ValueError("the grid must be a contiguous float32 [X,Y,Z] tensor on the step's device")
        self.grid = grid
        X, Y, Z = grid.shape
        self.grad_grid = torch.zeros_like(self.grid)
        self.bitmask = torch.zeros(Fn.vr_bitmask_words((X, Y, Z)), device=dev, dtype=torch.int32)
        self.vr_cells = torch.zeros_like(self.bitmask) if Z % 32 == 0 else None
        nb = ctypes.c_int64(0)
        N.check(N.lib().vxn_reg_vr_workspace_bytes(X, Y, Z, ctypes.byref(nb)))
        self.reg_ws = torch.empty(int(nb.value), device=dev, dtype=torch.uint8)
        self.eps3 = Fn.voxel_size_of(self.box, (X, Y, Z))
        self._g_eik = None
        self._args = None  # the render arguments hold the old grid

    def update_occupancy(self, subsample_ratio: float = 0.25, generator=None, full_sweep: bool = False,
                         decay: float = 0.95, threshold: float = 0.01):
        """The trainer's periodic occupancy update (trainer.py:171-173, every 16 steps;
        occupancy_grid.py:62-119) of this step's occupancy mask over its box, at the
        current s."""
        res = self.occ_mask.shape[0]
        idx = None
        if not full_sweep:
            n = max(int(res ** 3 * subsample_ratio), 1024)
            idx = tuple(torch.randint(0, res, (n,), device=self.grid.device, generator=generator)
                        for _ in range(3))
        s = float(torch.clamp(torch.exp(self.variance.detach() * 10.0), min=1e-4, max=self.s_max))
        Fn.occupancy_update(self.occ_density, self.occ_mask, self.occ_box, self.grid, self.box, s, decay, threshold,
                            self.occ_updates == 0, idx)
        self.occ_updates += 1

    def active_linear(self) -> torch.Tensor:
        """Sorted V_R linear indices of the last step after the relaxed-sphere filter
        (compacted from the bitmask on demand; one host read)."""
        lst, counts = Fn.vr_compact(self.bitmask, tuple(self.grid.shape), self.box, self.cfg.relax_radius)
        n_all, n_in = (int(v) for v in counts.tolist())
        lst = lst[:n_all]
        if n_in > 0:
            lst = lst[lst >= 0]
        return lst & 0x7FFFFFFF


class FusedOptimizer:
    """The reference's optimizer step (trainer.py:328-353, training/optimizer.py:15-154)
    on a ``FusedTrainStep``'s device buffers, with no host synchronisation:

    * SparseGridAdam (lr_sdf, betas 0.9/0.99, after reset_sdf 0.9/0.999, eps 1e-8) on the grid through the native
      masked kernel; ``n_updated`` counts the updated voxels on the device;
    * Adam (betas 0.9/0.999, eps 1e-8) on the hash table, the MLP and the variance in one
      launch, the hash and MLP rates under the cosine decay of trainer.py:329, the
      variance rate multiplied by k on the device when the step's s-gradient was
      amplified (trainer.py:332), and an optional variance floor (trainer.py:346-353).

    A bf16 hash table keeps an FP32 master here; each update writes its bf16 rounding
    back into the step's table."""

    def __init__(self, st: "FusedTrainStep", lr_sdf: float = 0.1, lr_hash: float = 0.01, lr_mlp: float = 1e-3,
                 lr_variance: float = 5e-4, total_steps: int = 40000):
        self.st = st
        self.lr_sdf, self.lr_hash, self.lr_mlp, self.lr_variance = lr_sdf, lr_hash, lr_mlp, lr_variance
        self.total_steps = total_steps
        z = lambda t: torch.zeros(t.shape, device=t.device, dtype=torch.float32)
        self.grid_m, self.grid_v = z(st.grid), z(st.grid)
        self.table_master = st.table.float().contiguous() if st.table.dtype == torch.bfloat16 else None
        self.table_m, self.table_v = z(st.table), z(st.table)
        self.mlp_m, self.mlp_v = z(st.mlp), z(st.mlp)
        self.var_m, self.var_v = z(st.variance), z(st.variance)
        self.n_updated = torch.zeros(1, device=st.grid.device, dtype=torch.int32)
        self.sdf_step = 0
        self.net_step = 0
        # create_optimizer builds the grid's SparseGridAdam with betas (0.9, 0.99)
        # (optimizer.py:139); the one rebuilt at a super-resolution milestone takes the
        # class default (0.9, 0.999) (trainer.py:157, optimizer.py:24)
        self.sdf_beta2 = 0.99

    def reset_sdf(self, lr: float) -> None:
        """A fresh SparseGridAdam on the step's (rescaled) grid at rate ``lr``
        (trainer.py:150-159): zero moments, step count from 0, and the class's default
        betas (0.9, 0.999), as ``SparseGridAdam(sdf_params, lr=new_lr)`` takes them."""
        z = torch.zeros(self.st.grid.shape, device=self.st.grid.device, dtype=torch.float32)
        self.grid_m, self.grid_v = z, torch.zeros_like(z)
        self.lr_sdf = float(lr)
        self.sdf_step = 0
        self.sdf_beta2 = 0.999

    # ---- checkpoint state in the reference's layout (checkpoint.py:30-31, 57-62) --------
    def _mlp_split(self, flat: torch.Tensor):
        """The flat MLP vector as the four tensors of radiance_mlp.net (W0, b0, W1, b1)."""
        n_in = (flat.numel() - 64 - 3 * 64 - 3) // 64
        sizes = (64 * n_in, 64, 3 * 64, 3)
        shapes = ((64, n_in), (64,), (3, 64), (3,))
        return [t.reshape(sh) for t, sh in zip(torch.split(flat, sizes), shapes)]

    def state_dict(self) -> Dict[str, Dict]:
        """Both optimizers' states as the reference saves them: ``sdf_optimizer`` is
        SparseGridAdam's ``state_dict()`` (one parameter, the grid: step, exp_avg,
        exp_avg_sq) and ``net_optimizer`` torch.optim.Adam's over the parameter groups of
        create_optimizer (optimizer.py:119-154): the hash table, the MLP's four tensors,
        the variance.  A bf16 table's FP32 master rides along as ``net_optimizer``'s
        ``fp32_master`` (torch.optim.Adam ignores it).  Tensors are copies on the device."""
        sdf = {"state": {0: {"step": self.sdf_step, "exp_avg": self.grid_m.clone(),
                             "exp_avg_sq": self.grid_v.clone()}},
               "param_groups": [{"lr": self.lr_sdf, "betas": (0.9, self.sdf_beta2), "eps": 1e-8,
                                 "weight_decay": 0.0, "params": [0]}]}
        step = lambda: torch.tensor(float(self.net_step))
        state = {0: {"step": step(), "exp_avg": self.table_m.clone(), "exp_avg_sq": self.table_v.clone()}}
        for i, (m, v) in enumerate(zip(self._mlp_split(self.mlp_m), self._mlp_split(self.mlp_v))):
            state[1 + i] = {"step": step(), "exp_avg": m.clone(), "exp_avg_sq": v.clone()}
        state[5] = {"step": step(), "exp_avg": self.var_m.clone(), "exp_avg_sq": self.var_v.clone()}
        common = {"betas": (0.9, 0.999), "eps": 1e-8, "weight_decay": 0.0, "amsgrad": False, "maximize": False,
                  "foreach": None, "capturable": False, "differentiable": False, "fused": None}
        groups = [dict(common, lr=self.lr_hash, params=[0]), dict(common, lr=self.lr_mlp, params=[1, 2, 3, 4]),
                  dict(common, lr=self.lr_variance, params=[5])]
        net = {"state": state, "param_groups": groups}
        if self.table_master is not None:
            net["fp32_master"] = self.table_master.clone()
        return {"sdf_optimizer": sdf, "net_optimizer": net}

    def load_state_dict(self, sd: Dict[str, Dict]) -> None:
        """Restore ``state_dict()``'s output (either key may be missing); the learning
        rates and the grid's betas come from the saved groups, the tensors are copied onto
        the step's device.  Without ``fp32_master`` a bf16 table's master is the table."""
        st = self.st
        dev = st.grid.device
        f32 = lambda t: t.to(device=dev, dtype=torch.float32).contiguous()
        if "sdf_optimizer" in sd:
            s0 = sd["sdf_optimizer"]["state"][0]
            g0 = sd["sdf_optimizer"]["param_groups"][0]
            if tuple(s0["exp_avg"].shape) != tuple(st.grid.shape):
                raise ValueError(f"saved SDF moments are {tuple(s0['exp_avg'].shape)}, the grid {tuple(st.grid.shape)}")
            self.grid_m, self.grid_v = f32(s0["exp_avg"]), f32(s0["exp_avg_sq"])
            self.sdf_step = int(s0["step"])
            self.lr_sdf = float(g0["lr"])
            self.sdf_beta2 = float(g0["betas"][1])
        if "net_optimizer" in sd:
            net = sd["net_optimizer"]
            s = net["state"]
            if tuple(s[0]["exp_avg"].shape) != tuple(st.table.shape):
                raise ValueError(f"saved hash moments are {tuple(s[0]['exp_avg'].shape)}, the table "
                                 f"{tuple(st.table.shape)}")
            self.table_m, self.table_v = f32(s[0]["exp_avg"]), f32(s[0]["exp_avg_sq"])
            self.mlp_m = torch.cat([f32(s[i]["exp_avg"]).reshape(-1) for i in range(1, 5)])
            self.mlp_v = torch.cat([f32(s[i]["exp_avg_sq"]).reshape(-1) for i in range(1, 5)])
            if self.mlp_m.numel() != st.mlp.numel():
                raise ValueError(f"saved MLP moments hold {self.mlp_m.numel()} values, the MLP {st.mlp.numel()}")
            self.var_m = f32(s[5]["exp_avg"]).reshape(st.variance.shape)
            self.var_v = f32(s[5]["exp_avg_sq"]).reshape(st.variance.shape)
            self.net_step = int(float(s[0]["step"]))
            groups = net["param_groups"]
            self.lr_hash, self.lr_mlp, self.lr_variance = (float(g["lr"]) for g in groups[:3])
            if self.table_master is not None:
                self.table_master = (f32(net["fp32_master"]) if "fp32_master" in net
                                     else st.table.float().contiguous())

    def as_optimizers(self) -> Dict[str, "_StateView"]:
        """``{"sdf_optimizer": ..., "net_optimizer": ...}`` objects with ``state_dict()`` and
        ``load_state_dict()``, for ``save_checkpoint`` / ``load_checkpoint``'s ``optimizers``."""
        return {"sdf_optimizer": _StateView(self, "sdf_optimizer"), "net_optimizer": _StateView(self, "net_optimizer")}

    def step(self, it: int, variance_floor: Optional[float] = None):
        """Apply both optimizers after ``FusedTrainStep.step`` of training iteration ``it``."""
        import math

        from .optim import cosine_lr_factor
        st = self.st
        L = N.lib()
        s = Fn._stream(st.grid)
        f = cosine_lr_factor(it, self.total_steps)
        if self.grid_m.shape != st.grid.shape:
            raise ValueError(f"the grid is {tuple(st.grid.shape)} but the SDF moments are "
                             f"{tuple(self.grid_m.shape)}: call reset_sdf after rescaling the grid")
        self.sdf_step += 1
        self.n_updated.zero_()
        N.check(L.vxn_masked_adam_upd(st.grid.data_ptr(), st.grad_grid.data_ptr(), self.grid_m.data_ptr(),
                                      self.grid_v.data_ptr(), st.grid.numel(), self.sdf_step, 0.9, self.sdf_beta2,
                                      self.lr_sdf, 1e-8, self.n_updated.data_ptr(), s), "masked_adam_upd")
        self.net_step += 1
        master = self.table_master
        ts = (N.AdamTensor * 3)()
        rows = (((master if master is not None else st.table), st.grad_table, self.table_m, self.table_v,
                 self.lr_hash * f, None, 1.0, -math.inf, st.table if master is not None else None),
                (st.mlp, st.grad_mlp, self.mlp_m, self.mlp_v, self.lr_mlp * f, None, 1.0, -math.inf, None),
                (st.variance, st.grad_v, self.var_m, self.var_v, self.lr_variance,
                 st.amplified if st.cfg.amplify else None, float(st.cfg.s_amp_k),
                 -math.inf if variance_floor is None else float(variance_floor), None))
        for d, (p, g, m, v, lr, flag, scale, pmin, bf) in zip(ts, rows):
            d.param, d.grad, d.exp_avg, d.exp_avg_sq = p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr()
            d.n = p.numel()
            d.lr, d.lr_flag, d.lr_flag_scale, d.param_min = lr, N.dptr(flag), scale, pmin
            d.param_bf16 = N.dptr(bf)
        N.check(L.vxn_adam_step(ctypes.cast(ts, ctypes.c_void_p), 3, self.net_step, 0.9, 0.999, 1e-8, s),
                "adam_step")



class _StateView:
    """One of FusedOptimizer's two optimizers, as save_checkpoint / load_checkpoint see it."""

    def __init__(self, opt: FusedOptimizer, name: str):
        self.opt, self.name = opt, name

    def state_dict(self) -> Dict:
        return self.opt.state_dict()[self.name]

    def load_state_dict(self, sd: Dict) -> None:
        self.opt.load_state_dict({self.name: sd})


@dataclass
class SuperResolutionSchedule:
    """Progressive grid super-resolution of the trainer (trainer.py:107-109, 131-164):
    at ``milestones[i]`` the grid goes to ``resolutions[i + 1]``, the SDF optimizer
    restarts at ``lr_sdf_stages[i + 1]``, and the eikonal gradient is clamped for
    ``cooldown_steps`` steps (the clamp itself is ``StepConfig.eik_grad_clamp``)."""

    milestones: tuple = (10000, 30000)
    resolutions: tuple = (96, 160, 320)
    lr_sdf_stages: tuple = (0.1, 0.05, 0.02)
    cooldown_steps: int = 500

    def apply(self, step: int, st: FusedTrainStep, opt: Optional[FusedOptimizer] = None) -> bool:
        """check_and_apply_upsampling: True if ``step`` is a milestone (and the grid was
        rescaled)."""
        ms = list(self.milestones)
        if step not in ms:
            return False
        idx = ms.index(step) + 1
        st.rescale_grid(int(self.resolutions[idx]))
        if opt is not None:
            opt.reset_sdf(float(self.lr_sdf_stages[min(idx, len(self.lr_sdf_stages) - 1)]))
        st.post_upsample_cooldown = int(self.cooldown_steps)
        return True
