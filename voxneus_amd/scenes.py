"""Seeded synthetic scenes for the VoxNeuS voxel-SDF path.

Every array here is generated from a numpy PCG64 seed so that the same scene can be
rebuilt bit-for-bit on any host.  The reference trains on DTU / BlendedMVS captures
(``Code/voxneus/datasets/dtu.py:215-256``); these scenes replace the dataset with
closed-form geometry whose true surface, normals and colours are known, while keeping
the reference's ``sample_rays`` contract (``rays_o``, ``rays_d``, ``rgb``).

Geometry conventions follow the reference:
  * the SDF grid is ``[X, Y, Z]`` float32 with z fastest, vertices at
    ``linspace(xyz_min, xyz_max, res)`` (``Code/voxneus/models/sdf_grid.py:59-75``);
  * hash tables are ``[L, T, F]`` float32 drawn from U(-1e-4, 1e-4)
    (``Code/voxneus/models/radiance_hash.py:56-63``);
  * the radiance MLP is ``Linear(51, 64) -> ReLU -> Linear(64, 3) -> Sigmoid`` with a
    kaiming-uniform (ReLU gain) weight init and zero biases
    (``Code/voxneus/models/radiance_mlp.py:36-79``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np


# ----------------------------------------------------------------------------------------
# Analytic signed distance functions (float64)
# ----------------------------------------------------------------------------------------

def smooth_min(a: np.ndarray, b: np.ndarray, k: float) -> np.ndarray:
    """Polynomial smooth minimum (Quilez): min(a,b) - h^2 k / 4 with h = max(k-|a-b|,0)/k."""
    h = np.maximum(k - np.abs(a - b), 0.0) / k
    return np.minimum(a, b) - h * h * k * 0.25


@dataclass
class MultiPrimitive:
    """Smooth-min (k) union of spheres and one z-axis torus centred at the origin."""

    centres: np.ndarray  # (n, 3)
    radii: np.ndarray  # (n,)
    torus_R: float = 0.4
    torus_r: float = 0.1
    k: float = 0.05

    @staticmethod
    def seeded(seed: int, n_spheres: int = 3) -> "MultiPrimitive":
        rng = np.random.default_rng(seed)
        centres = rng.uniform(-0.4, 0.4, size=(n_spheres, 3))
        radii = rng.uniform(0.15, 0.35, size=(n_spheres,))
        return MultiPrimitive(centres=centres, radii=radii)

    def sdf(self, p: np.ndarray) -> np.ndarray:
        p = np.asarray(p, dtype=np.float64)
        d = None
        for c, r in zip(self.centres, self.radii):
            ds = np.linalg.norm(p - c, axis=-1) - r
            d = ds if d is None else smooth_min(d, ds, self.k)
        q = np.stack([np.hypot(p[..., 0], p[..., 1]) - self.torus_R, p[..., 2]], axis=-1)
        dt = np.linalg.norm(q, axis=-1) - self.torus_r
        return smooth_min(d, dt, self.k)

    def normal(self, p: np.ndarray, h: float = 1e-6) -> np.ndarray:
        g = np.zeros(p.shape, dtype=np.float64)
        for a in range(3):
            e = np.zeros(3)
            e[a] = h
            g[..., a] = (self.sdf(p + e) - self.sdf(p - e)) / (2 * h)
        return g / np.maximum(np.linalg.norm(g, axis=-1, keepdims=True), 1e-12)


@dataclass
class Sphere:
    radius: float = 0.5

    def sdf(self, p: np.ndarray) -> np.ndarray:
        return np.linalg.norm(np.asarray(p, dtype=np.float64), axis=-1) - self.radius

    def normal(self, p: np.ndarray) -> np.ndarray:
        p = np.asarray(p, dtype=np.float64)
        return p / np.maximum(np.linalg.norm(p, axis=-1, keepdims=True), 1e-12)


# ----------------------------------------------------------------------------------------
# Grids and parameters
# ----------------------------------------------------------------------------------------

def vertex_axes(res: Tuple[int, int, int], xyz_min, xyz_max) -> List[np.ndarray]:
    """Vertex coordinates per axis, computed like ``torch.linspace(min, max, res)`` in float32."""
    import torch

    return [
        torch.linspace(float(xyz_min[a]), float(xyz_max[a]), int(res[a]), dtype=torch.float32).numpy()
        for a in range(3)
    ]


def sdf_grid_from(shape_fn, res: int, xyz_min=(-1.0, -1.0, -1.0), xyz_max=(1.0, 1.0, 1.0),
                  noise_std: float = 0.0, seed: int = 0) -> np.ndarray:
    """Sample an analytic SDF at every grid vertex; optionally add N(0, noise_std) noise."""
    r3 = (res, res, res)
    ax = vertex_axes(r3, xyz_min, xyz_max)
    out = np.empty(r3, dtype=np.float32)
    gy, gz = np.meshgrid(ax[1].astype(np.float64), ax[2].astype(np.float64), indexing="ij")
    for i, x in enumerate(ax[0].astype(np.float64)):
        p = np.stack([np.full_like(gy, x), gy, gz], axis=-1)
        out[i] = shape_fn(p).astype(np.float32)
    if noise_std > 0:
        rng = np.random.default_rng(seed)
        out += rng.normal(0.0, noise_std, size=r3).astype(np.float32)
    return out


def hash_embeddings(n_levels: int, log2_size: int, n_features: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(-1e-4, 1e-4, size=(n_levels, 1 << log2_size, n_features)).astype(np.float32)


def mlp_params(in_dim: int = 51, hidden: int = 64, seed: int = 0) -> Dict[str, np.ndarray]:
    """Kaiming-uniform (ReLU gain: bound = sqrt(6 / fan_in)) weights, zero biases."""
    rng = np.random.default_rng(seed)
    b0 = math.sqrt(6.0 / in_dim)
    b1 = math.sqrt(6.0 / hidden)
    return {
        "W0": rng.uniform(-b0, b0, size=(hidden, in_dim)).astype(np.float32),
        "b0": np.zeros((hidden,), np.float32),
        "W1": rng.uniform(-b1, b1, size=(3, hidden)).astype(np.float32),
        "b1": np.zeros((3,), np.float32),
    }


def hash_max_resolution(base: int, per_level_scale: float, n_levels: int) -> float:
    """``max_resolution`` that makes the reference's geometric factor equal ``per_level_scale``."""
    return float(base) * per_level_scale ** (n_levels - 1)


# ----------------------------------------------------------------------------------------
# Rays
# ----------------------------------------------------------------------------------------

def normalize(v: np.ndarray) -> np.ndarray:
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def sphere_rays(n_rays: int, radius: float = 3.0, jitter: float = 0.2, seed: int = 0):
    """Origins uniform on a sphere of ``radius``, aimed at the origin plus U(+-jitter)^3."""
    rng = np.random.default_rng(seed)
    o = normalize(rng.normal(size=(n_rays, 3))) * radius
    target = rng.uniform(-jitter, jitter, size=(n_rays, 3))
    d = normalize(target - o)
    return o.astype(np.float32), d.astype(np.float32)


def look_at_c2w(eye: np.ndarray, target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """OpenCV camera-to-world (x right, y down, z forward), as ``get_rays`` expects
    (``Code/voxneus/rendering/rays.py:40-56``)."""
    eye = np.asarray(eye, np.float64)
    fwd = normalize(np.asarray(target, np.float64) - eye)
    right = np.cross(fwd, np.asarray(up, np.float64))
    if np.linalg.norm(right) < 1e-8:
        right = np.cross(fwd, np.array([0.0, 1.0, 0.0]))
    right = normalize(right)
    down = np.cross(fwd, right)
    c2w = np.eye(4)
    c2w[:3, 0], c2w[:3, 1], c2w[:3, 2], c2w[:3, 3] = right, down, fwd, eye
    return c2w


def hemisphere_cameras(n_views: int = 49, radius: float = 3.0) -> np.ndarray:
    """Deterministic Fibonacci spiral of ``n_views`` positions on the z>0 hemisphere."""
    i = np.arange(n_views) + 0.5
    z = 1.0 - i / n_views * 0.9  # elevation from ~0.1 to ~1.0 (cos of polar angle)
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    rxy = np.sqrt(1.0 - z * z)
    pos = np.stack([rxy * np.cos(phi), rxy * np.sin(phi), z], axis=-1) * radius
    return np.stack([look_at_c2w(p) for p in pos])


def pinhole_rays(c2w: np.ndarray, W: int, H: int, fx: float, fy: float, cx: float, cy: float,
                 px: np.ndarray, py: np.ndarray):
    """Rays through pixel centres (+0.5) in the OpenCV convention, normalised."""
    dirs = np.stack([(px + 0.5 - cx) / fx, (py + 0.5 - cy) / fy, np.ones_like(px, dtype=np.float64)], -1)
    d = normalize(dirs @ c2w[:3, :3].T)
    o = np.broadcast_to(c2w[:3, 3], d.shape)
    return o.astype(np.float32), d.astype(np.float32)


def hemisphere_ray_batch(n_rays: int, seed: int, n_views: int = 49, W: int = 400, H: int = 300,
                         f: float = 600.0):
    """``n_rays`` rays drawn uniformly over (view, pixel) of the hemisphere camera rig."""
    rng = np.random.default_rng(seed)
    cams = hemisphere_cameras(n_views)
    v = rng.integers(0, n_views, size=n_rays)
    px = rng.integers(0, W, size=n_rays).astype(np.float64)
    py = rng.integers(0, H, size=n_rays).astype(np.float64)
    o = np.empty((n_rays, 3), np.float32)
    d = np.empty((n_rays, 3), np.float32)
    for k in range(n_views):
        m = v == k
        if m.any():
            o[m], d[m] = pinhole_rays(cams[k], W, H, f, f, W / 2, H / 2, px[m], py[m])
    return o, d


def full_view_rays(view: int, n_views: int = 49, W: int = 800, H: int = 600, f: float = 1200.0):
    cams = hemisphere_cameras(n_views)
    py, px = np.meshgrid(np.arange(H, dtype=np.float64), np.arange(W, dtype=np.float64), indexing="ij")
    return pinhole_rays(cams[view], W, H, f, f, W / 2, H / 2, px.reshape(-1), py.reshape(-1))


def sphere_trace(shape, o: np.ndarray, d: np.ndarray, t_max: float = 6.0, iters: int = 256,
                 tol: float = 1e-7):
    """FP64 sphere tracing of an analytic SDF; returns (hit mask, hit points)."""
    o = o.astype(np.float64)
    d = d.astype(np.float64)
    t = np.zeros(o.shape[0])
    alive = np.ones(o.shape[0], bool)
    hit = np.zeros(o.shape[0], bool)
    for _ in range(iters):
        p = o + d * t[:, None]
        s = shape.sdf(p)
        newly = alive & (np.abs(s) < tol)
        hit |= newly
        alive &= ~newly
        t = np.where(alive, t + np.maximum(s, 0.0) * 0.9 + np.where(s < 0, s, 0.0), t)
        alive &= t < t_max
    return hit, o + d * t[:, None]


def sphere_hit_rgb(o: np.ndarray, d: np.ndarray, radius: float = 0.5) -> np.ndarray:
    """Config-1 target: 0.5 + 0.5 n at the analytic sphere hit, 0 on background."""
    o = o.astype(np.float64)
    d = d.astype(np.float64)
    b = np.sum(o * d, -1)
    c = np.sum(o * o, -1) - radius * radius
    disc = b * b - c
    hit = disc > 0
    t = -b - np.sqrt(np.maximum(disc, 0.0))
    p = o + d * t[:, None]
    n = p / radius
    rgb = np.where(hit[:, None], 0.5 + 0.5 * n, 0.0)
    return rgb.astype(np.float32)


def procedural_rgb(shape, o: np.ndarray, d: np.ndarray) -> np.ndarray:
    """Config-3 target: 0.5 + 0.5 sin(6 p) at the FP64 sphere-traced hit, 0 on background."""
    hit, p = sphere_trace(shape, o, d)
    return np.where(hit[:, None], 0.5 + 0.5 * np.sin(6.0 * p), 0.0).astype(np.float32)


# ----------------------------------------------------------------------------------------
# Complete scene bundles for the measured configurations
# ----------------------------------------------------------------------------------------

@dataclass
class Scene:
    """All arrays the VoxNeuS path reads, in the reference's layouts."""

    grid: np.ndarray  # (X, Y, Z) float32
    embeddings: np.ndarray  # (L, T, F) float32
    mlp: Dict[str, np.ndarray]
    xyz_min: Tuple[float, float, float] = (-1.0, -1.0, -1.0)
    xyz_max: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    hash_base_resolution: int = 16
    hash_max_resolution: float = 2048.0
    init_s_val: float = 0.3
    s_max: float = 300.0
    occupancy_res: int = 128
    shape: object = None
    meta: Dict[str, object] = field(default_factory=dict)

    @property
    def grid_res(self) -> int:
        return int(self.grid.shape[0])


def make_scene(kind: str, grid_res: int, log2_T: int, seed: int = 0, n_levels: int = 16,
               per_level_scale: float = 1.5, noise_std: Optional[float] = None) -> Scene:
    """``kind`` is "sphere" (config-1 geometry) or "multi" (seeded multi-primitive)."""
    if kind == "sphere":
        shape = Sphere(0.5)
        noise = 1e-3 if noise_std is None else noise_std
    elif kind == "multi":
        shape = MultiPrimitive.seeded(seed)
        noise = 0.0 if noise_std is None else noise_std
    else:
        raise ValueError(f"unknown scene kind {kind!r}")
    grid = sdf_grid_from(shape.sdf, grid_res, noise_std=noise, seed=seed + 1)
    emb = hash_embeddings(n_levels, log2_T, 2, seed + 2)
    mlp = mlp_params(51, 64, seed + 3)
    return Scene(grid=grid, embeddings=emb, mlp=mlp, shape=shape,
                 hash_max_resolution=hash_max_resolution(16, per_level_scale, n_levels),
                 meta={"kind": kind, "seed": seed, "grid_res": grid_res, "log2_T": log2_T,
                       "per_level_scale": per_level_scale, "noise_std": noise})


def config1_scene(seed: int = 0) -> Scene:
    return make_scene("sphere", 64, 15, seed)


def config_rays(kind: str, n_rays: int, seed: int, shape=None):
    """(rays_o, rays_d, target_rgb) for config-1 style ("sphere") or hemisphere rigs."""
    if kind == "sphere":
        o, d = sphere_rays(n_rays, seed=seed)
        return o, d, sphere_hit_rgb(o, d, 0.5)
    o, d = hemisphere_ray_batch(n_rays, seed)
    return o, d, procedural_rgb(shape, o, d)
