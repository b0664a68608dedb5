"""Active vertex set and closed-form regularisers, drop-in for regularization/
{active_vertices,eikonal,curvature}.py, backed by the gfx950 V_R and regulariser kernels.

Losses are reduced in FP64 on the device (means over interior vertices, as the
reference's pure-PyTorch branch defines them); gradients use the reference's 2/|V_R|
scale with |V_R| counting every active vertex.
"""

from __future__ import annotations

from typing import Optional, Tuple

import torch

from . import functional as Fn


def extract_active_vertices(corner_indices: torch.Tensor, grid_res: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """active_vertices.py:12-50: sorted unique vertices (V,3) int64 and linear indices (V,).
    Bitmask marking + ordered compaction; one device read for V (as torch.unique)."""
    shape = tuple(int(v) for v in grid_res.tolist())
    X, Y, Z = shape
    dev = corner_indices.device
    bm = torch.zeros(Fn.vr_bitmask_words(shape), device=dev, dtype=torch.int32)
    Fn.vr_mark_corners(corner_indices, shape, bm)
    lst, counts = Fn.vr_compact(bm, shape, (0.0, 0.0, 0.0, float(X - 1), float(Y - 1), float(Z - 1)),
                                relax_radius=float("inf"))
    n = int(counts[0].item())
    lin = (lst[:n] & 0x7FFFFFFF).to(torch.int64)
    iz = lin % Z
    rem = lin // Z
    return torch.stack([rem // Y, rem % Y, iz], -1), lin


def _vr_from_vertices(active_vertices: torch.Tensor, grid_res):
    X, Y, Z = (int(v) for v in grid_res)
    a = active_vertices.to(torch.int64)
    lin = (a[:, 0] * (Y * Z) + a[:, 1] * Z + a[:, 2]).to(torch.int32).contiguous()
    V = lin.shape[0]
    counts = torch.tensor([V, V], device=a.device, dtype=torch.int32)
    return lin, counts


def _reg(grid_data, active_vertices, voxel_size, grid_res, accumulate_grad, weight, which):
    data = grid_data.reshape(grid_data.shape[-3:]) if grid_data.ndim == 5 else grid_data
    data = data.detach().contiguous()
    X, Y, Z = (int(v) for v in grid_res.tolist()) if torch.is_tensor(grid_res) else grid_res
    if active_vertices.shape[0] == 0:
        zero = None if accumulate_grad is not None else torch.zeros([1, 1, X, Y, Z], dtype=data.dtype,
                                                                     device=data.device)
        return torch.tensor(0.0, device=data.device), zero
    lin, counts = _vr_from_vertices(active_vertices, (X, Y, Z))
    eps3 = [float(v) for v in voxel_size.detach().cpu().tolist()]
    if accumulate_grad is not None:
        target = accumulate_grad.reshape(accumulate_grad.shape[-3:]) if accumulate_grad.ndim == 5 else accumulate_grad
        w = float(weight)
    else:
        target = torch.zeros((X, Y, Z), device=data.device, dtype=torch.float32)
        w = 1.0
    we, wc = (w, 0.0) if which == 0 else (0.0, w)
    losses = Fn.reg_eik_curv(data, eps3, lin, counts, we, wc, target)
    loss = losses[which].to(torch.float32)
    return loss, (None if accumulate_grad is not None else target.reshape(1, 1, X, Y, Z))


def compute_eikonal_loss_and_gradient(grid_data, active_vertices, voxel_size, grid_res, eps: float = 1e-8,
                                      accumulate_grad: Optional[torch.Tensor] = None, weight: float = 1.0):
    """eikonal.py:13-171."""
    return _reg(grid_data, active_vertices, voxel_size, grid_res, accumulate_grad, weight, 0)


def compute_curvature_loss_and_gradient(grid_data, active_vertices, voxel_size, grid_res,
                                        accumulate_grad: Optional[torch.Tensor] = None, weight: float = 1.0):
    """curvature.py:13-165."""
    return _reg(grid_data, active_vertices, voxel_size, grid_res, accumulate_grad, weight, 1)
