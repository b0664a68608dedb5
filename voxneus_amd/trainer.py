"""``VoxNeuSTrainer.train_step`` (training/trainer.py:33-353) over the fused device step.

``FusedTrainer`` drives one ``FusedTrainStep`` and one ``FusedOptimizer`` whose buffers
are the model's own parameter storage, so the model renders, extracts meshes and
checkpoints the trained state with no copies.  Each ``train_step`` follows the
reference's order:

1. progressive super-resolution at the milestones (grid rescale, fresh SDF optimizer at
   the stage's rate, 500-step eikonal clamp);
2. the occupancy update every ``occupancy_update_interval`` steps;
3. the NeuS2 lambda of the step (progressive hash bandwidth);
4. ray batch, unit-sphere bounds, iter_cos ratio min(step / 10000, 1);
5. render, colour (and mask) loss, backward, V_R, eikonal and curvature gradients at the
   scheduler's weights;
6. the s-gradient amplification once ``step >= s_warmup_steps``;
7. the optimizer steps with the cosine rates and the progressive variance floor.

The losses stay on the device; ``read_metrics`` turns one step's metrics into the
reference's dictionary of floats with one synchronisation.  The dataset, monitor and
checkpoint I/O of the reference's loop are the caller's.
"""

from __future__ import annotations

import math
from typing import Any, Dict, Optional

import torch
import torch.nn as nn

from .optim import cosine_lr_factor
from .scheduler import VoxNeuSScheduler
from .training import FusedOptimizer, FusedTrainStep, StepConfig, SuperResolutionSchedule

# trainer.py:338-345: s >= 80 after the first milestone, s >= 150 after the second
_VARIANCE_FLOORS = (0.4382, 0.5010)


class FusedTrainer:
    def __init__(self, model, config: Optional[Dict[str, Any]] = None, dataset: Any = None,
                 device: Optional[torch.device] = None):
        config = config or {}
        self.config = config
        self.dataset = dataset
        self.device = device if device is not None else torch.device("cuda")
        self.model = model.to(self.device)
        train = config.get("training", {})
        samp = config.get("sampler", {})
        sched = config.get("scheduler", {})
        opt = config.get("optimizer", {})
        self.total_steps = int(train.get("total_steps", 40000))
        self.batch_size = int(train.get("batch_size", 2048))
        self.use_mask = bool(train.get("use_mask", False))
        self.mask_weight = float(train.get("mask_weight", 0.1))
        self.loss_type = train.get("loss_type", "Huber")
        self.occupancy_update_interval = int(train.get("occupancy_update_interval", 16))
        self.progressive_hash = bool(train.get("progressive_hash", True))
        self.scheduler = VoxNeuSScheduler(
            total_steps=self.total_steps, stage1_end=sched.get("stage1_end", 11000),
            stage2_end=sched.get("stage2_end", 21000), w_eik_init=sched.get("w_eik_init", 1e-2),
            w_eik_final=sched.get("w_eik_final", 1e-3), w_curv_init=sched.get("w_curv_init", 1e-8),
            w_curv_peak=sched.get("w_curv_peak", 5e-6), w_curv_final=sched.get("w_curv_final", 5e-7))
        self.lr_sdf_stages = [opt.get("lr_sdf", 0.1), opt.get("lr_sdf_stage2", 0.05), opt.get("lr_sdf_stage3", 0.02)]
        self.current_lr_sdf = self.lr_sdf_stages[0]
        self.lr_hash = float(opt.get("lr_hash", 0.01))
        self.lr_mlp = float(opt.get("lr_mlp", 0.001))
        self.lr_variance = float(opt.get("lr_variance", 0.0005))
        if not opt.get("use_sparse_sdf", True):
            raise NotImplementedError("the fused optimizer is SparseGridAdam on the grid (use_sparse_sdf=True)")
        if float(sched.get("w_tv_sdf", 0.0)) > 0:
            raise NotImplementedError("the SDF total-variation loss (w_tv_sdf > 0) is not on the fused path")
        self.sr_milestones = list(train.get("sr_milestones", [10000, 30000]))
        self.sr_resolutions = list(train.get("sr_resolutions", [96, 160, 320]))
        self.s_amplification_k = float(sched.get("s_amplification_k", 5.0))
        self.s_warmup_steps = int(sched.get("s_warmup_steps", 5000))
        self.eik_grad_clamp = float(sched.get("eik_grad_clamp", 0.1))
        self.use_iter_cos = bool(config.get("model", {}).get("use_iter_cos_annealing", False))
        self.sr = SuperResolutionSchedule(tuple(self.sr_milestones), tuple(self.sr_resolutions),
                                          tuple(self.lr_sdf_stages), cooldown_steps=500)

        m = self.model
        if getattr(m, "entangled", False) or getattr(m, "geom_feature_dim", 0):
            raise NotImplementedError("the fused step trains the disentangled model without geometry features")
        mlp = m.radiance_mlp
        flat = mlp.flat().detach().clone().contiguous()
        # the module's MLP parameters become views of the flat buffer the step trains
        o = 0
        for p in (mlp.net[0].weight, mlp.net[0].bias, mlp.net[2].weight, mlp.net[2].bias):
            n = p.numel()
            p.data = flat[o:o + n].view_as(p)
            o += n
        enc = m.radiance_hash.encoder
        g = m.sdf_grid.grid
        cfg = StepConfig(n_samples=samp.get("n_samples", 64), n_importance=samp.get("n_importance", 64),
                         up_sample_steps=samp.get("up_sample_steps", 4), perturb=bool(samp.get("perturb", True)),
                         loss_type=self.loss_type, use_mask=self.use_mask, mask_weight=self.mask_weight,
                         s_amp_k=self.s_amplification_k, mlp_mode=m.mlp_mode, gradient_mode=m.gradient_mode,
                         view_embed_type=m.view_embed_type, eik_grad_clamp=self.eik_grad_clamp,
                         use_iter_cos=self.use_iter_cos)
        self.step_engine = FusedTrainStep(g.data.view(g.shape[-3:]), enc.embeddings.data, flat,
                                          m.variance_net.variance.data, enc.level_res, box=m.sdf_grid.box6,
                                          s_max=m.variance_net._s_max, occ_mask=m.occupancy_grid.occupancy_mask,
                                          cfg=cfg, occ_box=m.occupancy_grid.box6)
        self.optimizer = FusedOptimizer(self.step_engine, lr_sdf=self.lr_sdf_stages[0], lr_hash=self.lr_hash,
                                        lr_mlp=self.lr_mlp, lr_variance=self.lr_variance,
                                        total_steps=self.total_steps)
        self.last_out = None

    # -------------------------------------------------------------------------------
    def check_and_apply_upsampling(self, step: int) -> bool:
        """trainer.py:131-164 on the shared grid: the model's grid parameter becomes the
        step's rescaled grid."""
        st = self.step_engine
        if not self.sr.apply(step, st, self.optimizer):
            return False
        idx = self.sr_milestones.index(step) + 1
        self.current_lr_sdf = self.lr_sdf_stages[min(idx, len(self.lr_sdf_stages) - 1)]
        X, Y, Z = st.grid.shape
        self.model.sdf_grid.grid = nn.Parameter(st.grid.view(1, 1, X, Y, Z))
        self.model.sdf_grid.res = torch.tensor([X, Y, Z], dtype=torch.float32)
        return True

    def variance_floor(self, step: int) -> Optional[float]:
        """trainer.py:338-345: the minimum of v after the optimizer step of ``step``."""
        ms = self.sr_milestones
        if len(ms) == 0 or step < ms[0]:
            return None
        return _VARIANCE_FLOORS[1] if len(ms) > 1 and step >= ms[1] else _VARIANCE_FLOORS[0]

    def train_step(self, step: int, batch: Optional[Dict[str, torch.Tensor]] = None) -> Dict[str, Any]:
        """One training iteration.  ``batch`` holds ``rays_o``, ``rays_d``, ``rgb`` and
        optionally ``mask`` (the dataset's ``sample_rays`` dictionary); without one the
        batch is drawn from the dataset.  Returns the step's metrics, the losses and
        counters as device tensors (see ``read_metrics``)."""
        st, cfg = self.step_engine, self.step_engine.cfg
        if batch is None and self.dataset is None:
            raise ValueError("train_step needs a ray batch or a dataset to draw one from")
        self.check_and_apply_upsampling(step)
        if step % self.occupancy_update_interval == 0:
            self.model.update_occupancy_grid(subsample_ratio=0.25)
            st.occ_mask = self.model.occupancy_grid.occupancy_mask
        active_lambda = self.scheduler.get_lambda(step) if self.progressive_hash else None
        st.set_active_levels(active_lambda)
        if batch is None:
            batch = self.dataset.sample_rays(self.batch_size)
        rays_o = batch["rays_o"].to(self.device, torch.float32)
        rays_d = batch["rays_d"].to(self.device, torch.float32)
        target = batch["rgb"].to(self.device, torch.float32)
        mask = batch.get("mask", None)
        if mask is not None:
            mask = mask.to(self.device)
        cfg.cos_anneal_ratio = min(float(step) / 10000.0, 1.0)
        w_eik, w_curv = self.scheduler.get_weights(step)
        cfg.w_eik, cfg.w_curv = float(w_eik), float(w_curv)
        cfg.amplify = step >= self.s_warmup_steps
        out = st.step(rays_o, rays_d, target, target_mask=mask)
        self.last_out = out  # the step's device outputs (losses, z values) until the next step
        self.optimizer.step(step, variance_floor=self.variance_floor(step))
        f = cosine_lr_factor(step, self.total_steps)
        return {"loss_total": out["loss"], "loss_rgb": out["loss_parts"][0:1],
                "loss_mask": out["loss_parts"][1:2] if (self.use_mask and mask is not None) else None,
                "loss_eik": out["reg_losses"][0:1], "w_eik": w_eik, "loss_curv": out["reg_losses"][1:2],
                "w_curv": w_curv, "variance": st.variance, "s_amplified": out["amplified"] if cfg.amplify else None,
                "grid_res": int(st.grid.shape[0]), "lr_sdf": self.current_lr_sdf, "lr_hash": self.lr_hash * f,
                "lr_mlp": self.lr_mlp * f, "sparse_updated": self.optimizer.n_updated,
                "active_lambda": active_lambda, "loss_name": self.loss_type, "vr_counts": out["vr_counts"]}

    def _psnr(self, loss_rgb: float) -> float:
        lo = self.last_out
        err = (lo["rgb"].double() - lo["target"].reshape(-1, 3).double()) ** 2
        m = lo.get("target_mask")
        if m is not None:
            fg = m.reshape(-1, 1).double().expand_as(err)
            fg_sum = float(fg.sum())
            mse = float((err * fg).sum()) / fg_sum if fg_sum > 0 else loss_rgb
        else:
            mse = float(err.mean())
        return -10.0 * math.log10(max(mse, 1e-8))

    def read_metrics(self, metrics: Dict[str, Any]) -> Dict[str, Any]:
        """The reference's metrics dictionary (trainer.py:361-382) as host values, read
        right after ``train_step`` (one synchronisation).  ``psnr`` is -10 log10 of the
        step's colour MSE for every loss type, as trainer.py:217-240 computes it: the
        foreground-weighted MSE with a mask (the colour loss itself when the mask is empty),
        else the MSE over all rays and channels, from the colours the step rendered."""
        out = {}
        for k, v in metrics.items():
            out[k] = float(v.reshape(-1)[0]) if torch.is_tensor(v) and k != "vr_counts" else v
        amp = metrics["s_amplified"]
        out["s_amplified"] = bool(int(amp.reshape(-1)[0])) if amp is not None else False
        out["sparse_updated"] = int(metrics["sparse_updated"].reshape(-1)[0])
        out["vr_counts"] = [int(x) for x in metrics["vr_counts"].tolist()]
        v = out.pop("variance")
        st = self.step_engine
        s = min(max(math.exp(v * 10.0), 1e-4), st.s_max)
        out["s"], out["std"] = s, 1.0 / s
        out["lr_s"] = self.lr_variance * (self.s_amplification_k if out["s_amplified"] else 1.0)
        out["psnr"] = self._psnr(out["loss_rgb"])
        out["occ_pct"] = self.model.occupancy_grid.occupied_percentage
        return out
