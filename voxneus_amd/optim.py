"""Optimizers of the VoxNeuS training step, updated on the device by the gfx950 kernels
(training/optimizer.py:15-154).

* ``SparseGridAdam`` mirrors the reference's class: Adam only where the grid gradient is
  nonzero, through the native masked kernel (``vxn_masked_adam_upd``, the reference's
  ``_backend.masked_adam_upd`` path) or, with weight decay, the masked PyTorch formula
  (``vxn_sparse_adam_decay``).  ``last_updated_count`` is counted on the device and read
  only when asked for.
* ``NetAdam`` is torch.optim.Adam's update (the reference's network optimizer) as one
  multi-tensor launch (``vxn_adam_step``), with two group options the training step
  uses to stay on the device: ``lr_flag`` (a device int32 flag; nonzero multiplies the
  group's lr by ``lr_flag_scale``: the s-amplified variance learning rate) and
  ``param_min`` (a floor applied after the update: the variance floor).  bfloat16
  parameters keep an FP32 master in the optimizer state and receive its rounding.
* ``create_optimizer`` builds both with the reference's groups, learning rates and betas.
"""

from __future__ import annotations

import ctypes
import math
from typing import Dict, Iterable, List, Optional

import torch
from torch.optim import Optimizer

from . import _native as N

_MAX_TENSORS = 8


def _check_tensor(t: torch.Tensor, name: str):
    if not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float32 HIP tensor (got {t.dtype}, {t.device})")
    if t.data_ptr() % 16 != 0:
        raise ValueError(f"{name} must be 16-byte aligned")


class SparseGridAdam(Optimizer):
    """Adam with sparse updates: only elements whose gradient is nonzero change
    (training/optimizer.py:15-116)."""

    def __init__(self, params, lr: float = 1e-1, betas: tuple = (0.9, 0.999), eps: float = 1e-8,
                 weight_decay: float = 0.0):
        defaults = dict(lr=lr, betas=betas, eps=eps, weight_decay=weight_decay)
        super().__init__(params, defaults)
        self._count_dev: Optional[torch.Tensor] = None
        self._count_host: Optional[int] = 0

    @property
    def last_updated_count(self) -> int:
        """Elements updated by the last step (one device read, on demand)."""
        if self._count_host is None:
            self._count_host = int(self._count_dev.item())
        return self._count_host

    @last_updated_count.setter
    def last_updated_count(self, v: int):
        self._count_host = int(v)

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        L = N.lib()
        counter = None
        for group in self.param_groups:
            lr = group["lr"]
            beta1, beta2 = group["betas"]
            eps = group["eps"]
            wd = group["weight_decay"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                grad = p.grad
                if grad.is_sparse:
                    raise RuntimeError("SparseGridAdam does not support torch.sparse tensors")
                _check_tensor(p, "param")
                _check_tensor(grad, "grad")
                state = self.state[p]
                if len(state) == 0:
                    state["step"] = 0
                    state["exp_avg"] = torch.zeros_like(p, memory_format=torch.preserve_format)
                    state["exp_avg_sq"] = torch.zeros_like(p, memory_format=torch.preserve_format)
                state["step"] += 1
                step = state["step"]
                if counter is None:
                    if self._count_dev is None or self._count_dev.device != p.device:
                        self._count_dev = torch.zeros(1, device=p.device, dtype=torch.int32)
                    counter = self._count_dev
                    counter.zero_()
                m, v = state["exp_avg"], state["exp_avg_sq"]
                s = N.stream_ptr(p.device)
                if wd == 0.0:
                    N.check(L.vxn_masked_adam_upd(p.data_ptr(), grad.data_ptr(), m.data_ptr(), v.data_ptr(),
                                                  p.numel(), int(step), float(beta1), float(beta2), float(lr),
                                                  float(eps), counter.data_ptr(), s), "masked_adam_upd")
                else:
                    N.check(L.vxn_sparse_adam_decay(p.data_ptr(), grad.data_ptr(), m.data_ptr(), v.data_ptr(),
                                                    p.numel(), int(step), float(beta1), float(beta2), float(lr),
                                                    float(eps), float(wd), counter.data_ptr(), s),
                            "sparse_adam_decay")
        if counter is not None:
            self._count_host = None
        return loss


class NetAdam(Optimizer):
    """torch.optim.Adam (no weight decay, no amsgrad) updated by one multi-tensor kernel.

    Extra per-group options: ``lr_flag`` (device int32 tensor or None), ``lr_flag_scale``
    (float), ``param_min`` (float floor after the update, default -inf)."""

    def __init__(self, params, lr: float = 1e-3, betas: tuple = (0.9, 0.999), eps: float = 1e-8,
                 weight_decay: float = 0.0, amsgrad: bool = False):
        if weight_decay != 0.0 or amsgrad:
            raise NotImplementedError("the network optimizer has no weight decay or amsgrad")
        defaults = dict(lr=lr, betas=betas, eps=eps, weight_decay=0.0, amsgrad=False, lr_flag=None,
                        lr_flag_scale=1.0, param_min=-math.inf)
        super().__init__(params, defaults)

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        batches: Dict[tuple, List[N.AdamTensor]] = {}
        keep = []  # temporaries referenced by the launch
        for group in self.param_groups:
            beta1, beta2 = group["betas"]
            flag = group["lr_flag"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                if p.grad.is_sparse:
                    raise RuntimeError("Adam does not support sparse gradients")
                state = self.state[p]
                if len(state) == 0:
                    state["step"] = torch.tensor(0.0, dtype=torch.float32)
                    shape_src = p if p.dtype == torch.float32 else p.float()
                    state["exp_avg"] = torch.zeros_like(shape_src, memory_format=torch.preserve_format)
                    state["exp_avg_sq"] = torch.zeros_like(shape_src, memory_format=torch.preserve_format)
                    if p.dtype == torch.bfloat16:
                        state["master"] = p.detach().float().contiguous()
                    elif p.dtype != torch.float32:
                        raise ValueError(f"parameters must be float32 or bfloat16 (got {p.dtype})")
                state["step"] += 1
                master = state.get("master")
                target = master if master is not None else p
                grad = p.grad if p.grad.dtype == torch.float32 else p.grad.float()
                grad = grad.contiguous()
                keep.append(grad)
                for t, nm in ((target, "param"), (grad, "grad"), (state["exp_avg"], "exp_avg"),
                              (state["exp_avg_sq"], "exp_avg_sq")):
                    _check_tensor(t, nm)
                if master is not None and (not p.is_contiguous() or p.data_ptr() % 2):
                    raise ValueError("bfloat16 parameters must be contiguous")
                d = N.AdamTensor()
                d.param, d.grad = target.data_ptr(), grad.data_ptr()
                d.exp_avg, d.exp_avg_sq = state["exp_avg"].data_ptr(), state["exp_avg_sq"].data_ptr()
                d.n = target.numel()
                d.lr = float(group["lr"])
                d.lr_flag = None if flag is None else N.dptr(flag, torch.int32, "lr_flag")
                d.lr_flag_scale = float(group["lr_flag_scale"])
                d.param_min = float(group["param_min"])
                d.param_bf16 = p.data_ptr() if master is not None else None
                key = (int(state["step"].item()), float(beta1), float(beta2), float(group["eps"]), p.device)
                batches.setdefault(key, []).append(d)
        L = N.lib()
        for (step, beta1, beta2, eps, dev), ts in batches.items():
            for i in range(0, len(ts), _MAX_TENSORS):
                chunk = ts[i:i + _MAX_TENSORS]
                arr = (N.AdamTensor * len(chunk))(*chunk)
                N.check(L.vxn_adam_step(ctypes.cast(arr, ctypes.c_void_p), len(chunk), step, beta1, beta2, eps,
                                        N.stream_ptr(dev)), "adam_step")
        del keep
        return loss


def create_optimizer(model, lr_sdf: float = 0.1, lr_hash: float = 1e-2, lr_mlp: float = 1e-3,
                     lr_variance: float = 1e-2, use_sparse_sdf: bool = True) -> Dict[str, Optimizer]:
    """training/optimizer.py:119-154: sparse Adam (betas 0.9, 0.99) on the SDF grid, and
    Adam (betas 0.9, 0.999, eps 1e-8) on the hash table, MLP and variance, one group each."""
    sdf_params = [model.sdf_grid.grid]
    if getattr(model.sdf_grid, "feat_grid", None) is not None:
        sdf_params.append(model.sdf_grid.feat_grid)
    if use_sparse_sdf:
        sdf_opt = SparseGridAdam(sdf_params, lr=lr_sdf, betas=(0.9, 0.99))
    else:
        sdf_opt = NetAdam(sdf_params, lr=lr_sdf, betas=(0.9, 0.99))
    net_params = [
        {"params": model.radiance_hash.parameters(), "lr": lr_hash},
        {"params": model.radiance_mlp.parameters(), "lr": lr_mlp},
        {"params": model.variance_net.parameters(), "lr": lr_variance},
    ]
    net_opt = NetAdam(net_params, betas=(0.9, 0.999), eps=1e-8)
    return {"sdf_optimizer": sdf_opt, "net_optimizer": net_opt}


def cosine_lr_factor(step: int, total_steps: int) -> float:
    """trainer.py:329: cosine decay to 5 % for the hash and MLP learning rates."""
    return (math.cos(math.pi * step / max(total_steps, 1)) + 1.0) * 0.5 * (1.0 - 0.05) + 0.05
