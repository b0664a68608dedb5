"""Dense SDF voxel grid, drop-in for DenseSDFGrid (models/sdf_grid.py:24-199) and the
geometry helpers it uses (geometry/coordinates.py, geometry/gradients.py), backed by the
fused gfx950 grid-query kernels.

Layout: ``grid`` is an ``nn.Parameter`` of shape [1, 1, X, Y, Z] float32, z fastest, with
vertices at ``linspace(xyz_min, xyz_max, res)`` - the reference's layout, so state dicts
load unchanged.
"""

from __future__ import annotations

from typing import List, Optional, Tuple, Union

import numpy as np
import torch
import torch.nn as nn

from . import _native as N
from . import functional as Fn


def world_to_grid_index(xyz, xyz_min, xyz_max, grid_res):
    """coordinates.py:10-29 (elementwise)."""
    return (xyz - xyz_min) / (xyz_max - xyz_min + 1e-8) * (grid_res - 1.0)


class DenseSDFGrid(nn.Module):
    def __init__(self, resolution: Union[int, List[int], Tuple[int, int, int]] = 96, xyz_min=(-1.0, -1.0, -1.0),
                 xyz_max=(1.0, 1.0, 1.0), init_mode: str = "sphere", sphere_radius: float = 0.5, feat_dim: int = 0):
        super().__init__()
        if feat_dim:
            raise NotImplementedError("geometry feature grids (NeuS2 feat_dim > 0) are outside this path")
        if isinstance(resolution, int):
            self.res = torch.tensor([resolution] * 3, dtype=torch.float32)
        else:
            self.res = torch.tensor(list(resolution), dtype=torch.float32)
        self.register_buffer("xyz_min", torch.tensor(xyz_min, dtype=torch.float32))
        self.register_buffer("xyz_max", torch.tensor(xyz_max, dtype=torch.float32))
        self.init_mode = init_mode
        self.sphere_radius = sphere_radius
        self.feat_dim = 0
        self.feat_grid = None
        rx, ry, rz = (int(v) for v in self.res)
        self.grid = nn.Parameter(torch.zeros([1, 1, rx, ry, rz], dtype=torch.float32))
        self.reset_parameters()

    def reset_parameters(self) -> None:
        """sdf_grid.py:59-81."""
        rx, ry, rz = (int(v) for v in self.res)
        with torch.no_grad():
            if self.init_mode == "sphere":
                axes = [torch.linspace(self.xyz_min[a].item(), self.xyz_max[a].item(), r)
                        for a, r in enumerate((rx, ry, rz))]
                gx, gy, gz = torch.meshgrid(*axes, indexing="ij")
                dist = torch.linalg.norm(torch.stack([gx, gy, gz], -1), dim=-1) - self.sphere_radius
                self.grid.data.copy_(dist[None, None].to(self.grid.device))
            elif self.init_mode == "zero":
                self.grid.data.zero_()
            else:
                self.grid.data.uniform_(-0.05, 0.05)

    @property
    def box(self) -> Tuple[Tuple[float, float, float], Tuple[float, float, float]]:
        return tuple(self.xyz_min.tolist()), tuple(self.xyz_max.tolist())

    @property
    def box6(self) -> Tuple[float, ...]:
        return tuple(self.xyz_min.tolist() + self.xyz_max.tolist())

    @property
    def voxel_size(self) -> torch.Tensor:
        """sdf_grid.py:83-88."""
        return (self.xyz_max - self.xyz_min) / (self.res.to(self.xyz_min.device) - 1.0)

    def query_sdf(self, xyz: torch.Tensor) -> torch.Tensor:
        """Trilinear SDF (sdf_grid.py:90-120) -> (..., 1)."""
        sdf, _ = Fn.GridQueryFunction.apply(self.grid, xyz, self.box)
        return sdf.reshape(*xyz.shape[:-1], 1)

    def query_normals(self, xyz: torch.Tensor, gradient_mode: str = "interpolated"):
        """Gradient (Eq. 7 "interpolated" or Eq. 4 "analytical"), unit normal, corner indices
        (..., 8, 3) int64 and weights (..., 8) (sdf_grid.py:122-160); any other mode is a
        ValueError."""
        mode = N.gradient_mode_code(gradient_mode)
        shape = xyz.shape[:-1]
        raw = Fn.grid_query_raw(self.grid.detach().reshape(self.grid.shape[-3:]), self.box, xyz, want_sdf=False,
                                want_grad=False, want_corners=True, want_weights=True)
        _, grad = Fn.GridQueryFunction.apply(self.grid, xyz, self.box, mode)
        grad = grad.reshape(*shape, 3)
        normal = grad / (torch.linalg.norm(grad, dim=-1, keepdim=True) + 1e-8)
        return grad, normal, raw["corners"].reshape(*shape, 8, 3), raw["weights"].reshape(*shape, 8)

    def query_features(self, xyz: torch.Tensor) -> Optional[torch.Tensor]:
        return None

    def query_lattice(self, NX: int, NY: int, NZ: int, lattice_min=None, lattice_max=None,
                      gradient_mode: str = "interpolated"):
        """SDF and unit normal on a full NX*NY*NZ lattice (mesh-extraction query)."""
        lo = self.xyz_min.tolist() if lattice_min is None else list(lattice_min)
        hi = self.xyz_max.tolist() if lattice_max is None else list(lattice_max)
        return Fn.lattice_query(self.grid.detach().reshape(self.grid.shape[-3:]), self.box, NX, NY, NZ,
                                tuple(lo + hi), N.gradient_mode_code(gradient_mode))

    def scale_volume_grid(self, new_resolution: Union[int, List[int], Tuple[int, int, int]]) -> None:
        """sdf_grid.py:183-195: replace the grid by its trilinear upsampling (aligned corners)
        to ``new_resolution``, on the device."""
        self.grid = nn.Parameter(upsample_sdf_grid(self.grid, new_resolution))
        if isinstance(new_resolution, int):
            self.res = torch.tensor([new_resolution] * 3, dtype=torch.float32)
        else:
            self.res = torch.tensor(list(new_resolution), dtype=torch.float32)

    def get_numpy_grid(self) -> np.ndarray:
        return self.grid.detach().squeeze().cpu().numpy()


def get_voxel_corners_and_weights(continuous_coords: torch.Tensor, grid_res: torch.Tensor):
    """coordinates.py:55-124 on continuous grid coordinates."""
    r = [int(v) for v in grid_res.tolist()]
    corners, weights = Fn.voxel_corners(continuous_coords, r)
    shape = continuous_coords.shape[:-1]
    return corners.reshape(*shape, 8, 3), weights.reshape(*shape, 8)


def interpolate_gradients(grid, continuous_coords, voxel_size, grid_res):
    """gradients.py:103-139 on continuous coordinates (the compiled-backend entry)."""
    corners, weights = get_voxel_corners_and_weights(continuous_coords, grid_res)
    g = Fn.GradientInterpFunction.apply(grid, continuous_coords, voxel_size)
    g = g.reshape(*continuous_coords.shape[:-1], 3)
    return g, g / (torch.linalg.norm(g, dim=-1, keepdim=True) + 1e-8), corners, weights


def compute_analytical_gradient(grid, continuous_coords, voxel_size, grid_res):
    """gradients.py:142-220: the Eq. 4 gradient and its unit normal on continuous grid
    coordinates; the gradient is differentiable in the grid."""
    return Fn.compute_analytical_gradient(grid, continuous_coords, voxel_size, grid_res)


def upsample_sdf_grid(grid_tensor: torch.Tensor, new_res) -> torch.Tensor:
    """upsample.py:12-39: (1, 1, X, Y, Z) -> (1, 1, NX, NY, NZ), trilinear with aligned
    corners (``vxn_upsample_grid``)."""
    shape = (new_res,) * 3 if isinstance(new_res, int) else tuple(int(v) for v in new_res)
    return Fn.upsample_grid(grid_tensor.detach(), shape).reshape(1, 1, *shape)


def get_resolution_for_step(step: int, milestones=(10000, 30000), resolutions=(96, 160, 320)) -> int:
    """upsample.py:42-63: the grid resolution of a training step (96^3 until 10k, 160^3 until
    30k, then 320^3)."""
    if step < milestones[0]:
        return resolutions[0]
    if step < milestones[1]:
        return resolutions[1]
    return resolutions[2]
