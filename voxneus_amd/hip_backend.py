"""Mirror of the reference's native dispatch shim (cuda/__init__.py:1-106) with the same
names and contracts, so code written against ``voxneus.cuda`` can import this module
instead.  ``cuda_*`` entry points return None / False only when the tensors are not on
a HIP device (the reference's fallback signal); on a GPU they always run the gfx950
kernels and raise on any error.
"""

from __future__ import annotations

from typing import Optional

import torch

from . import _native as N
from . import functional as Fn

GradientInterpFunction = Fn.GradientInterpFunction


def is_hip_available() -> bool:
    try:
        N.lib()
    except Exception:
        return False
    return torch.cuda.is_available()


is_cuda_available = is_hip_available
HAS_CUDA_BACKEND = True


def cuda_gradient_interp(grid: torch.Tensor, coords: torch.Tensor, voxel_size: torch.Tensor) -> Optional[torch.Tensor]:
    """cuda/__init__.py:62-70."""
    if not grid.is_cuda:
        return None
    return GradientInterpFunction.apply(grid, coords, voxel_size)


def cuda_eikonal_grad(grid, active_verts, grad_grid, voxel_size, weight: float = 1.0) -> bool:
    """cuda/__init__.py:73-88 (accumulates into grad_grid)."""
    if not grid.is_cuda:
        return False
    g3 = grid.reshape(grid.shape[-3:]).contiguous()
    t3 = grad_grid.reshape(grad_grid.shape[-3:])
    Fn.eikonal_grad_ref(g3, active_verts, t3, voxel_size.detach().cpu().tolist(), weight)
    return True


def cuda_curvature_grad(grid, active_verts, grad_grid, voxel_size, weight: float = 1.0) -> bool:
    """cuda/__init__.py:91-106 (accumulates into grad_grid)."""
    if not grid.is_cuda:
        return False
    g3 = grid.reshape(grid.shape[-3:]).contiguous()
    t3 = grad_grid.reshape(grad_grid.shape[-3:])
    Fn.curvature_grad_ref(g3, active_verts, t3, voxel_size.detach().cpu().tolist(), weight)
    return True
