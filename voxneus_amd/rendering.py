"""Ray sampling and volume rendering, drop-in for rendering/{sampler,neus_alpha,
volume_render,occupancy_grid,rays}.py.

The multi-step upsampler runs in the gfx950 sampler kernel; ``VoxNeuS.render`` composes
it with the fused render kernel.  The standalone ``compute_neus_alpha`` /
``composite_weights`` / ``volume_render_from_samples`` entry points keep the reference's
signatures for callers that already hold per-sample colours, SDF and normals; they run
the alpha and compositing kernels (csrc/composite.hip), each with its hand-written
backward, in the fused render kernel's arithmetic.
"""

from __future__ import annotations

from typing import Any, Dict, Optional, Tuple

import torch
import torch.nn as nn

from . import functional as Fn
from .training import near_far_from_sphere  # noqa: F401  (re-exported, rays.py:102-128)


def logistic_density(sdf: torch.Tensor, s: torch.Tensor) -> torch.Tensor:
    """occupancy_grid.py:15-22: phi_s(x) = s * sigmoid(s x) * (1 - sigmoid(s x)) (elementwise,
    on the tensors' device; the occupancy update evaluates it inside its kernel)."""
    sig = torch.sigmoid(s * sdf)
    return s * sig * (1.0 - sig)


class OccupancyGrid(nn.Module):
    """Occupancy mask consumed by the upsampler (occupancy_grid.py:33-147).  A fresh grid
    is all True; ``update_from_sdf`` is the EMA update from the SDF grid's logistic density
    (occupancy_grid.py:62-119) on the device."""

    def __init__(self, resolution: int = 128, xyz_min=(-1.0, -1.0, -1.0), xyz_max=(1.0, 1.0, 1.0),
                 decay: float = 0.95, threshold: float = 0.01):
        super().__init__()
        self.resolution = resolution
        self.register_buffer("xyz_min", torch.tensor(xyz_min, dtype=torch.float32))
        self.register_buffer("xyz_max", torch.tensor(xyz_max, dtype=torch.float32))
        self.decay = decay
        self.threshold = threshold
        self.register_buffer("density_grid", torch.zeros((resolution,) * 3, dtype=torch.float32))
        self.register_buffer("occupancy_mask", torch.ones((resolution,) * 3, dtype=torch.bool))
        self.total_updates = 0

    @property
    def box6(self):
        return tuple(self.xyz_min.tolist() + self.xyz_max.tolist())

    @property
    def occupied_percentage(self) -> float:
        return (self.occupancy_mask.sum().item() / self.occupancy_mask.numel()) * 100.0

    @torch.no_grad()
    def update_from_sdf(self, query_sdf_fn, s_val, subsample_ratio: float = 0.25, generator=None,
                        full_sweep: bool = False) -> None:
        """occupancy_grid.py:62-119.  ``query_sdf_fn`` must be a DenseSDFGrid's ``query_sdf``
        (the kernel evaluates the grid itself).  The cells are drawn as the reference draws
        them (three torch.randint vectors of max(res^3 * ratio, 1024) cells, with
        replacement, on the grid's device); ``full_sweep`` updates every cell once instead."""
        grid = getattr(query_sdf_fn, "__self__", None)
        if grid is None or not hasattr(grid, "grid"):
            raise NotImplementedError("the occupancy update evaluates a DenseSDFGrid directly")
        dev = self.density_grid.device
        res = self.resolution
        idx = None
        if not full_sweep:
            n = max(int(res * res * res * subsample_ratio), 1024)
            idx = tuple(torch.randint(0, res, (n,), device=dev, generator=generator) for _ in range(3))
        s = float(s_val.detach()) if torch.is_tensor(s_val) else float(s_val)
        Fn.occupancy_update(self.density_grid, self.occupancy_mask, self.box6,
                            grid.grid.detach().reshape(grid.grid.shape[-3:]), grid.box6, s, self.decay,
                            self.threshold, self.total_updates == 0, idx)
        self.total_updates += 1


def sdf_grid_of(sdf_fn):
    """The DenseSDFGrid that ``sdf_fn`` evaluates: a bound ``query_sdf``, or a closure over
    a grid (or over a model with an ``sdf_grid``) whose values on probe points equal that
    grid's ``query_sdf`` bit for bit.  Anything else raises: the sampler kernel cannot call
    back into Python."""
    from .sdf_grid import DenseSDFGrid

    owner = getattr(sdf_fn, "__self__", None)
    if isinstance(owner, DenseSDFGrid):
        return owner
    found = None
    for cell in getattr(sdf_fn, "__closure__", None) or ():
        try:
            c = cell.cell_contents
        except ValueError:
            continue
        g = c if isinstance(c, DenseSDFGrid) else getattr(c, "sdf_grid", None)
        if isinstance(g, DenseSDFGrid):
            found = g
            break
    if found is None:
        raise NotImplementedError("the gfx950 sampler evaluates a DenseSDFGrid directly: pass its query_sdf, a "
                                  "closure over the grid or over a model holding it as sdf_grid, or grid=")
    dev = found.grid.device
    lo = torch.tensor(found.box6[:3], device=dev)
    hi = torch.tensor(found.box6[3:], device=dev)
    gen = torch.Generator(device="cpu").manual_seed(0)
    probe = (lo + (hi - lo) * torch.rand(64, 3, generator=gen).to(dev)).contiguous()
    with torch.no_grad():
        got = sdf_fn(probe).reshape(-1)
        ref = found.query_sdf(probe).reshape(-1)
    if got.shape != ref.shape or not torch.equal(got, ref):
        raise NotImplementedError("sdf_fn does not evaluate the DenseSDFGrid it refers to (its values differ on "
                                  "probe points); the gfx950 sampler only samples the grid itself")
    return found


class RaySampler(nn.Module):
    """sampler.py:69-204 with the upsampler on the GPU."""

    def __init__(self, n_samples: int = 64, n_importance: int = 64, perturb: bool = True):
        super().__init__()
        self.n_samples = n_samples
        self.n_importance = n_importance
        self.perturb = perturb

    def sample_multistep_sdf(self, rays_o, rays_d, near, far, sdf_fn=None, n_samples: Optional[int] = None,
                             n_importance: Optional[int] = None, up_sample_steps: int = 4,
                             perturb: Optional[bool] = None, occupancy_grid: Optional[Any] = None,
                             grid=None, t_rand=None, u_rand=None):
        """sampler.py:127-204.  The gfx950 sampler evaluates the dense grid itself, so
        ``sdf_fn`` must evaluate one of this package's DenseSDFGrids: its ``query_sdf``, or a
        closure over the grid or over a model holding it as ``sdf_grid`` (what
        ``VoxNeuS.render`` passes, voxneus.py:210-224).  The closure's values are checked
        against the grid's own on probe points; ``grid=`` names the grid directly."""
        if grid is None:
            grid = sdf_grid_of(sdf_fn)
        n_samples = n_samples or self.n_samples
        n_importance = self.n_importance if n_importance is None else n_importance
        perturb = self.perturb if perturb is None else perturb
        occ_mask = occ_box = None
        if occupancy_grid is not None:
            occ_mask, occ_box = occupancy_grid.occupancy_mask, occupancy_grid.box6
        g3 = grid.grid.detach().reshape(grid.grid.shape[-3:])
        return Fn.upsample(rays_o, rays_d, near, far, g3, grid.box6, n_samples, n_importance, up_sample_steps,
                           bool(perturb), occ_mask, occ_box, t_rand, u_rand)


def compute_neus_alpha(sdf, normals, ray_dirs, dists, s, cos_anneal_ratio=0.0, use_iter_cos_annealing=False,
                       eps=1e-5):
    """neus_alpha.py:12-81 on the device (vxn_neus_alpha_fwd / _bwd, one thread per sample):
    sdf and dists (..., 1) or (...), normals and ray_dirs (..., 3), broadcast together; s a
    number or a one-element tensor.  Returns alpha (...)."""
    if eps != 1e-5:
        raise NotImplementedError("the alpha kernel uses the reference's eps = 1e-5")
    if sdf.ndim < normals.ndim:
        sdf = sdf.unsqueeze(-1)
    if dists.ndim < normals.ndim:
        dists = dists.unsqueeze(-1)
    lead = torch.broadcast_shapes(sdf.shape[:-1], dists.shape[:-1], normals.shape[:-1], ray_dirs.shape[:-1])
    sdf = sdf.expand(*lead, 1)
    dists = dists.expand(*lead, 1)
    normals = normals.expand(*lead, 3)
    ray_dirs = ray_dirs.expand(*lead, 3)
    return Fn.NeusAlphaFunction.apply(sdf.reshape(lead), normals, ray_dirs, dists.reshape(lead), s,
                                      cos_anneal_ratio, use_iter_cos_annealing)


def composite_weights(alpha, eps=1e-7):
    """volume_render.py:12-40 on the device (vxn_composite_* with the alpha as input):
    alpha [B, S] (S <= 128) -> (weights, transmittance) [B, S]."""
    if eps != 1e-7:
        raise NotImplementedError("the compositing kernel uses the reference's eps = 1e-7")
    out = Fn.CompositeFunction.apply(None, None, None, None, None, None, None, alpha, None, None, 0.0, False)
    return out[4], out[6]


def compute_huber_color_loss(pred_rgb, target_rgb, delta=0.1):
    """volume_render.py:43-54."""
    diff = torch.abs(pred_rgb - target_rgb)
    return torch.mean(torch.where(diff < delta, 0.5 * (diff ** 2) / delta, diff - 0.5 * delta))


def volume_render_from_samples(colors, sdf, normals, ray_dirs, z_vals, dists, s, pts=None, sphere_radius=1.0,
                               background_rgb=None, cos_anneal_ratio=0.0, use_iter_cos_annealing=False):
    """volume_render.py:57-153 on the device (vxn_composite_fwd / _bwd: one ray per work-group,
    FP64 transmittance, the fused render kernel's arithmetic): colors / normals (R, S, 3),
    sdf and dists (R, S) or (R, S, 1), ray_dirs (R, 3) or (R, S, 3), z_vals (R, S), pts
    (R, S, 3) for the unit-sphere mask; S <= 128.  Gradients reach colors, sdf, normals and
    a tensor s."""
    R, S = colors.shape[0], colors.shape[1]
    sdf = sdf.reshape(R, S)
    dists = dists.reshape(R, S)
    if ray_dirs.ndim == 3:
        ray_dirs = ray_dirs.expand(R, S, 3)
    rgb, depth, nrm, opacity, w, alpha, T = Fn.CompositeFunction.apply(
        colors, sdf, normals, ray_dirs, z_vals, dists, s, None, pts,
        sphere_radius if (pts is not None and sphere_radius is not None and sphere_radius > 0) else None,
        cos_anneal_ratio, use_iter_cos_annealing)
    rgb = apply_background(rgb, opacity, background_rgb)
    return {"rgb": rgb, "depth": depth, "normal": nrm, "opacity": opacity, "weights": w, "alpha": alpha,
            "transmittance": T}


def apply_background(rgb, opacity, background_rgb):
    """volume_render.py:136-143."""
    if background_rgb is None:
        return rgb
    if isinstance(background_rgb, (float, int)):
        bg = torch.full_like(rgb, float(background_rgb))
    elif background_rgb.ndim == 1:
        bg = background_rgb.unsqueeze(0).expand_as(rgb)
    else:
        bg = background_rgb
    return rgb + (1.0 - opacity) * bg
