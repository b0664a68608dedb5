"""Mesh extraction on the device (geometry/mesh.py:34-77): marching cubes over the dense
SDF grid by the gfx950 kernels (``vxn_mc_count`` / ``vxn_mc_emit``), in place of the
reference's host scikit-image call.  One vertex per grid edge whose end values straddle
the level, at the linear-interpolation crossing; triangles from a face-consistent case
table (closed where the surface is, normals towards increasing values).  The optional
Gaussian pre-filter (mesh.py:17-31) runs on the device too (``vxn_gaussian_filter``)."""

from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np
import torch

from . import _native as N
from . import functional as Fn


def marching_cubes(grid: torch.Tensor, box6, level: float = 0.0,
                   workspace: Optional[torch.Tensor] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """grid [X,Y,Z] (or [1,1,X,Y,Z]) float32 on a HIP device -> (verts [V,3] float32 in
    world space, faces [F,3] int32), both on the device.  One host read (the counts)."""
    g3 = Fn._grid3(grid).contiguous()
    X, Y, Z = g3.shape
    L = N.lib()
    nb = ctypes.c_int64(0)
    N.check(L.vxn_mc_workspace_bytes(X, Y, Z, ctypes.byref(nb)))
    ws = workspace
    if ws is None or ws.numel() * ws.element_size() < nb.value:
        ws = torch.empty(int(nb.value), device=g3.device, dtype=torch.uint8)
    counts = torch.zeros(2, device=g3.device, dtype=torch.int32)
    s = Fn._stream(g3)
    wsb = ws.numel() * ws.element_size()
    N.check(L.vxn_mc_count(N.dptr(g3, torch.float32, "grid"), X, Y, Z, float(level), ws.data_ptr(), wsb,
                           counts.data_ptr(), s), "mc_count")
    nv, nf = (int(v) for v in counts.tolist())
    verts = torch.empty(nv, 3, device=g3.device, dtype=torch.float32)
    faces = torch.empty(nf, 3, device=g3.device, dtype=torch.int32)
    N.check(L.vxn_mc_emit(g3.data_ptr(), X, Y, Z, N.box6(*box6), float(level), ws.data_ptr(), wsb,
                          N.dptr(verts), nv, N.dptr(faces), nf, s), "mc_emit")
    return verts, faces


def _device_grid(sdf_grid) -> torch.Tensor:
    g = sdf_grid if torch.is_tensor(sdf_grid) else torch.from_numpy(np.ascontiguousarray(sdf_grid, np.float32))
    if not g.is_cuda:
        g = g.to("cuda")
    return g.to(torch.float32)


def apply_gaussian_filter_to_grid(sdf_grid, sigma: float = 1.0):
    """mesh.py:17-31: the post-training 3-D Gaussian smoothing of the SDF volume
    (scipy.ndimage.gaussian_filter with mode="nearest") on the device.  A numpy input
    gives a numpy result, a tensor a device tensor."""
    out = Fn.gaussian_filter_grid(_device_grid(sdf_grid), sigma)
    return out if torch.is_tensor(sdf_grid) else out.cpu().numpy()


def extract_mesh_marching_cubes(sdf_grid, xyz_min, xyz_max, level: float = 0.0,
                                post_filter_sigma: Optional[float] = None) -> Tuple[np.ndarray, np.ndarray]:
    """mesh.py:34-77: world-space vertices (float32) and faces (int32) as numpy arrays.
    ``sdf_grid`` may be a numpy array (copied to the current HIP device) or a device
    tensor; ``post_filter_sigma`` > 0 smooths it first (mesh.py:56-57)."""
    g = _device_grid(sdf_grid)
    if post_filter_sigma is not None and post_filter_sigma > 0:
        g = Fn.gaussian_filter_grid(g, post_filter_sigma)
    box6 = tuple(float(v) for v in list(np.asarray(xyz_min, np.float32)) + list(np.asarray(xyz_max, np.float32)))
    verts, faces = marching_cubes(g.to(torch.float32), box6, level)
    return verts.cpu().numpy(), faces.cpu().numpy()
