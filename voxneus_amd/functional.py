"""torch.autograd.Functions and thin wrappers over the HIP C ABI.

Every function here runs hand-written gfx950 kernels through ``_native``; torch only
allocates device memory and supplies the stream.  Inputs must live on a HIP device.
"""

from __future__ import annotations

import ctypes
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as N


def _stream(t: torch.Tensor) -> int:
    return N.stream_ptr(t.device)


def _grid3(grid: torch.Tensor) -> torch.Tensor:
    g = grid.reshape(grid.shape[-3:]) if grid.ndim == 5 else grid
    if g.ndim != 3:
        raise ValueError(f"grid must be [X,Y,Z] or [1,1,X,Y,Z] (got {tuple(grid.shape)})")
    return g


# ----------------------------------------------------------------------------------------
# Dense SDF grid
# ----------------------------------------------------------------------------------------

def grid_query_raw(grid3: torch.Tensor, box, xyz: torch.Tensor, want_sdf=True, want_grad=True,
                   want_normal=False, want_corners=False, want_weights=False,
                   gradient_mode: int = N.VXN_GRADIENT_INTERPOLATED):
    """Run the fused SDF + gradient query (Eq. 7, or Eq. 4 with ``VXN_GRADIENT_ANALYTICAL``)
    on flattened points; returns a dict of tensors."""
    grid3 = grid3.contiguous()
    xyz = xyz.reshape(-1, 3).contiguous()
    n = xyz.shape[0]
    dev = xyz.device
    out = {}
    if want_sdf:
        out["sdf"] = torch.empty(n, device=dev, dtype=torch.float32)
    if want_grad:
        out["grad"] = torch.empty(n, 3, device=dev, dtype=torch.float32)
    if want_normal:
        out["normal"] = torch.empty(n, 3, device=dev, dtype=torch.float32)
    if want_corners:
        out["corners"] = torch.empty(n, 8, 3, device=dev, dtype=torch.int64)
    if want_weights:
        out["weights"] = torch.empty(n, 8, device=dev, dtype=torch.float32)
    X, Y, Z = grid3.shape
    N.check(N.lib().vxn_grid_query_fwd(
        N.dptr(grid3, torch.float32, "grid"), X, Y, Z, N.box6(*box), N.dptr(xyz, torch.float32, "xyz"), n,
        N.dptr(out.get("sdf")), N.dptr(out.get("grad")), N.dptr(out.get("normal")), N.dptr(out.get("corners")),
        N.dptr(out.get("weights")), int(gradient_mode), _stream(xyz)), "grid_query_fwd")
    return out


class GridQueryFunction(torch.autograd.Function):
    """(grid, xyz) -> (sdf [n], grad [n,3]); backward scatters into the grid."""

    @staticmethod
    def forward(ctx, grid, xyz, box, gradient_mode=N.VXN_GRADIENT_INTERPOLATED):
        g3 = _grid3(grid)
        o = grid_query_raw(g3, box, xyz, want_sdf=True, want_grad=True, gradient_mode=gradient_mode)
        ctx.save_for_backward(g3, xyz)
        ctx.box = box
        ctx.mode = int(gradient_mode)
        ctx.grid_shape = grid.shape
        return o["sdf"], o["grad"]

    @staticmethod
    def backward(ctx, d_sdf, d_grad):
        g3, xyz = ctx.saved_tensors
        gg = None
        if ctx.needs_input_grad[0]:
            gg = torch.zeros_like(g3)
            xf = xyz.reshape(-1, 3).contiguous()
            X, Y, Z = g3.shape
            ds = d_sdf.contiguous() if d_sdf is not None else None
            dg = d_grad.contiguous() if d_grad is not None else None
            N.check(N.lib().vxn_grid_query_bwd(
                N.dptr(g3), X, Y, Z, N.box6(*ctx.box), N.dptr(xf), xf.shape[0], N.dptr(ds), N.dptr(dg),
                N.dptr(gg), ctx.mode, _stream(gg)), "grid_query_bwd")
            gg = gg.reshape(ctx.grid_shape)
        return gg, None, None, None


class GradientInterpFunction(torch.autograd.Function):
    """Reference-shaped Eq. 7 on continuous grid coordinates (cuda/__init__.py:29-59)."""

    @staticmethod
    def forward(ctx, grid, coords, voxel_size):
        g3 = _grid3(grid).contiguous()
        c = coords.reshape(-1, 3).contiguous()
        eps = [float(v) for v in voxel_size.detach().cpu().tolist()]
        out = torch.empty(c.shape[0], 3, device=c.device, dtype=torch.float32)
        X, Y, Z = g3.shape
        N.check(N.lib().vxn_gradient_interp_fwd(N.dptr(g3, torch.float32, "grid"), X, Y, Z, N.dptr(c), c.shape[0],
                                                N.f3(eps), N.dptr(out), _stream(c)), "gradient_interp_fwd")
        ctx.save_for_backward(c)
        ctx.eps, ctx.shape3, ctx.grid_shape = eps, g3.shape, grid.shape
        return out

    @staticmethod
    def backward(ctx, gout):
        (c,) = ctx.saved_tensors
        gg = None
        if ctx.needs_input_grad[0]:
            X, Y, Z = ctx.shape3
            gg = torch.zeros(ctx.shape3, device=c.device, dtype=torch.float32)
            go = gout.contiguous()
            N.check(N.lib().vxn_gradient_interp_bwd(N.dptr(go), N.dptr(c), c.shape[0], N.dptr(gg), X, Y, Z,
                                                    N.f3(ctx.eps), _stream(gg)), "gradient_interp_bwd")
            gg = gg.reshape(ctx.grid_shape)
        return gg, None, None


class AnalyticalGradientFunction(torch.autograd.Function):
    """Eq. 4 on continuous grid coordinates: (grid, coords, voxel_size) -> grad [n,3]
    (geometry/gradients.py:142-220); backward scatters dL/dgrad into the 8 cell corners."""

    @staticmethod
    def forward(ctx, grid, coords, voxel_size):
        g3 = _grid3(grid).contiguous()
        c = coords.reshape(-1, 3).contiguous()
        eps = [float(v) for v in voxel_size.detach().cpu().tolist()]
        out = torch.empty(c.shape[0], 3, device=c.device, dtype=torch.float32)
        X, Y, Z = g3.shape
        N.check(N.lib().vxn_analytical_gradient_fwd(N.dptr(g3, torch.float32, "grid"), X, Y, Z,
                                                    N.dptr(c, torch.float32, "coords"), c.shape[0], N.f3(eps),
                                                    N.dptr(out), None, _stream(c)), "analytical_gradient_fwd")
        ctx.save_for_backward(c)
        ctx.eps, ctx.shape3, ctx.grid_shape = eps, g3.shape, grid.shape
        return out

    @staticmethod
    def backward(ctx, gout):
        (c,) = ctx.saved_tensors
        gg = None
        if ctx.needs_input_grad[0]:
            X, Y, Z = ctx.shape3
            gg = torch.zeros(ctx.shape3, device=c.device, dtype=torch.float32)
            go = gout.contiguous()
            N.check(N.lib().vxn_analytical_gradient_bwd(N.dptr(go), N.dptr(c), c.shape[0], N.dptr(gg), X, Y, Z,
                                                        N.f3(ctx.eps), _stream(gg)), "analytical_gradient_bwd")
            gg = gg.reshape(ctx.grid_shape)
        return gg, None, None


def compute_analytical_gradient(grid, continuous_coords, voxel_size, grid_res=None):
    """geometry/gradients.py:142-220: (grad, normal) of Eq. 4; grad is differentiable in grid."""
    shape = continuous_coords.shape[:-1]
    grad = AnalyticalGradientFunction.apply(grid, continuous_coords, voxel_size).reshape(*shape, 3)
    return grad, grad / (torch.linalg.norm(grad, dim=-1, keepdim=True) + 1e-8)


def voxel_corners(coords: torch.Tensor, grid_shape):
    """-> (corners [n,8,3] int64, weights [n,8]) of continuous grid coordinates."""
    c = coords.reshape(-1, 3).contiguous()
    n = c.shape[0]
    corners = torch.empty(n, 8, 3, device=c.device, dtype=torch.int64)
    weights = torch.empty(n, 8, device=c.device, dtype=torch.float32)
    X, Y, Z = grid_shape
    N.check(N.lib().vxn_voxel_corners(N.dptr(c, torch.float32, "coords"), n, X, Y, Z, N.dptr(corners),
                                      N.dptr(weights), _stream(c)), "voxel_corners")
    return corners, weights


def lattice_query(grid3, box, NX, NY, NZ, lattice_box, gradient_mode: int = N.VXN_GRADIENT_INTERPOLATED):
    g3 = grid3.contiguous()
    sdf = torch.empty(NX, NY, NZ, device=g3.device, dtype=torch.float32)
    nrm = torch.empty(NX, NY, NZ, 3, device=g3.device, dtype=torch.float32)
    X, Y, Z = g3.shape
    N.check(N.lib().vxn_lattice_query(N.dptr(g3, torch.float32, "grid"), X, Y, Z, N.box6(*box), NX, NY, NZ,
                                      N.box6(*lattice_box), N.dptr(sdf), N.dptr(nrm), int(gradient_mode),
                                      _stream(g3)), "lattice_query")
    return sdf, nrm


# ----------------------------------------------------------------------------------------
# Hash encoding
# ----------------------------------------------------------------------------------------

def _levels(level_res: Sequence[int], level_w: Optional[Sequence[float]]):
    L = len(level_res)
    res = (ctypes.c_int32 * L)(*[int(r) for r in level_res])
    w = None if level_w is None else (ctypes.c_float * L)(*[float(v) for v in level_w])
    return res, w


def _table_dtype(table: torch.Tensor) -> int:
    if table.dtype == torch.float32:
        return N.VXN_DTYPE_F32
    if table.dtype == torch.bfloat16:
        return N.VXN_DTYPE_BF16
    raise TypeError(f"hash table must be float32 or bfloat16 (got {table.dtype})")


class HashEncodeFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, table, x01, level_res, level_w):
        L, T, F = table.shape
        log2T = int(round(math.log2(T)))
        if 1 << log2T != T:
            raise ValueError("hash table size must be a power of two")
        tc = table.contiguous()
        xf = x01.reshape(-1, 3).contiguous()
        feats = torch.empty(xf.shape[0], L * F, device=xf.device, dtype=torch.float32)
        res, w = _levels(level_res, level_w)
        N.check(N.lib().vxn_hash_fwd(N.dptr(tc), _table_dtype(tc), L, log2T, F, res, w, N.dptr(xf, torch.float32, "x"),
                                     xf.shape[0], N.dptr(feats), _stream(xf)), "hash_fwd")
        ctx.save_for_backward(xf)
        ctx.meta = (L, T, F, log2T, list(level_res), None if level_w is None else list(level_w), table.dtype)
        return feats.reshape(*x01.shape[:-1], L * F)

    @staticmethod
    def backward(ctx, d_feats):
        (xf,) = ctx.saved_tensors
        L, T, F, log2T, lres, lw, tdt = ctx.meta
        g = None
        if ctx.needs_input_grad[0]:
            g = torch.zeros(L, T, F, device=xf.device, dtype=torch.float32)
            df = d_feats.reshape(-1, L * F).contiguous()
            res, w = _levels(lres, lw)
            N.check(N.lib().vxn_hash_bwd(L, log2T, F, res, w, N.dptr(xf), xf.shape[0], N.dptr(df), N.dptr(g),
                                         _stream(g)), "hash_bwd")
            g = g.to(tdt)
        return g, None, None, None


# ----------------------------------------------------------------------------------------
# NeuS alpha and compositing of per-sample values (volume_render.py, neus_alpha.py)
# ----------------------------------------------------------------------------------------

def _s_tensor(s, device) -> torch.Tensor:
    """The sharpness s as a one-element float32 device tensor (a Python number or a tensor
    holding one value)."""
    if torch.is_tensor(s):
        if s.numel() != 1:
            raise NotImplementedError("the compositing kernels take one s for every sample")
        return s.reshape(1).to(device=device, dtype=torch.float32)
    return torch.tensor([float(s)], device=device, dtype=torch.float32)


def _f32c(t):
    return None if t is None else t.to(torch.float32).contiguous()


class CompositeFunction(torch.autograd.Function):
    """volume_render_from_samples (volume_render.py:57-153) on [R,S] samples (S <= 128):
    returns rgb [R,3], depth [R,1], normal [R,3], opacity [R,1], weights, alpha and
    transmittance [R,S].  With ``alpha_in`` the alpha is an input (composite_weights)."""

    @staticmethod
    def forward(ctx, colors, sdf, normals, ray_dirs, z_vals, dists, s, alpha_in, pts, sphere_radius,
                cos_anneal_ratio, use_iter_cos):
        ref = alpha_in if alpha_in is not None else sdf
        R, S = ref.shape[0], ref.shape[1]
        dev = ref.device
        a = N.CompositeArgs()
        a.n_rays, a.n_samples = R, S
        keep = [_f32c(colors), _f32c(sdf), _f32c(normals), _f32c(ray_dirs), _f32c(z_vals), _f32c(dists),
                None if alpha_in is not None else _s_tensor(s, dev), _f32c(pts), _f32c(alpha_in)]
        c, f, n, d, z, dt, sv, p, ai = keep
        a.colors, a.sdf, a.normals, a.ray_dirs = N.dptr(c), N.dptr(f), N.dptr(n), N.dptr(d)
        a.dirs_per_sample = int(d is not None and d.dim() == 3)
        a.z_vals, a.dists, a.s, a.pts, a.alpha_in = N.dptr(z), N.dptr(dt), N.dptr(sv), N.dptr(p), N.dptr(ai)
        a.sphere_radius = float(sphere_radius) if (sphere_radius is not None and p is not None) else 0.0
        a.cos_anneal_ratio, a.use_iter_cos = float(cos_anneal_ratio), int(bool(use_iter_cos))
        e = lambda *sh: torch.empty(*sh, device=dev, dtype=torch.float32)
        outs = [e(R, 3), e(R, 1), e(R, 3), e(R, 1), e(R, S), e(R, S), e(R, S)]
        a.rgb, a.depth, a.normal, a.opacity, a.weights, a.alpha, a.transmittance = (N.dptr(o) for o in outs)
        N.check(N.lib().vxn_composite_fwd(ctypes.byref(a), _stream(ref)), "composite_fwd")
        ctx.save_for_backward(*[t if t is not None else torch.empty(0) for t in keep])
        ctx.present = [t is not None for t in keep]
        ctx.meta = (R, S, a.dirs_per_sample, a.sphere_radius, a.cos_anneal_ratio, a.use_iter_cos)
        ctx.s_is_tensor = torch.is_tensor(s)
        return tuple(outs)

    @staticmethod
    def backward(ctx, d_rgb, d_depth, d_normal, d_opacity, d_weights, d_alpha, d_trans):
        saved = [t if pr else None for t, pr in zip(ctx.saved_tensors, ctx.present)]
        c, f, n, d, z, dt, sv, p, ai = saved
        R, S, dps, sr, car, uic = ctx.meta
        ref = ai if ai is not None else f
        dev = ref.device
        a = N.CompositeArgs()
        a.n_rays, a.n_samples = R, S
        a.colors, a.sdf, a.normals, a.ray_dirs = N.dptr(c), N.dptr(f), N.dptr(n), N.dptr(d)
        a.dirs_per_sample = dps
        a.z_vals, a.dists, a.s, a.pts, a.alpha_in = N.dptr(z), N.dptr(dt), N.dptr(sv), N.dptr(p), N.dptr(ai)
        a.sphere_radius, a.cos_anneal_ratio, a.use_iter_cos = sr, car, uic
        g = [_f32c(t) for t in (d_rgb, d_depth, d_normal, d_opacity, d_weights, d_alpha, d_trans)]
        (a.d_rgb, a.d_depth, a.d_normal, a.d_opacity, a.d_weights, a.d_alpha,
         a.d_transmittance) = (N.dptr(t) for t in g)
        need = ctx.needs_input_grad
        dc = torch.empty(R, S, 3, device=dev) if (need[0] and c is not None) else None
        df = torch.empty(R, S, device=dev) if (need[1] and f is not None) else None
        dn = torch.empty(R, S, 3, device=dev) if (need[2] and n is not None) else None
        da = torch.empty(R, S, device=dev) if (need[7] and ai is not None) else None
        ds = torch.empty(1, device=dev) if (need[6] and sv is not None) else None
        a.d_colors, a.d_sdf, a.d_normals, a.d_alpha_in, a.d_s = (N.dptr(t) for t in (dc, df, dn, da, ds))
        nb = ctypes.c_int64(0)
        N.check(N.lib().vxn_composite_bwd_workspace_bytes(R, ctypes.byref(nb)), "composite_bwd_workspace_bytes")
        ws = torch.empty(max(int(nb.value), 8), device=dev, dtype=torch.uint8)
        N.check(N.lib().vxn_composite_bwd(ctypes.byref(a), ws.data_ptr(), ws.numel(), _stream(ref)), "composite_bwd")
        d_s = None if ds is None else (ds.reshape(()) if ctx.s_is_tensor else None)
        return dc, df, dn, None, None, None, d_s, da, None, None, None, None


class NeusAlphaFunction(torch.autograd.Function):
    """compute_neus_alpha (neus_alpha.py:12-81) elementwise on flattened samples."""

    @staticmethod
    def forward(ctx, sdf, normals, dirs, dists, s, cos_anneal_ratio, use_iter_cos):
        f, n, d, dt = _f32c(sdf.reshape(-1)), _f32c(normals.reshape(-1, 3)), _f32c(dirs.reshape(-1, 3)), \
            _f32c(dists.reshape(-1))
        if not (f.shape[0] == n.shape[0] == d.shape[0] == dt.shape[0]):
            raise ValueError("sdf, normals, ray_dirs and dists must hold the same number of samples")
        sv = _s_tensor(s, f.device)
        alpha = torch.empty_like(f)
        N.check(N.lib().vxn_neus_alpha_fwd(N.dptr(f), N.dptr(n), N.dptr(d), N.dptr(dt), N.dptr(sv), f.shape[0],
                                           float(cos_anneal_ratio), int(bool(use_iter_cos)), N.dptr(alpha),
                                           _stream(f)), "neus_alpha_fwd")
        ctx.save_for_backward(f, n, d, dt, sv)
        ctx.meta = (sdf.shape, normals.shape, float(cos_anneal_ratio), int(bool(use_iter_cos)), torch.is_tensor(s))
        return alpha.reshape(sdf.shape)

    @staticmethod
    def backward(ctx, d_alpha):
        f, n, d, dt, sv = ctx.saved_tensors
        sshape, nshape, car, uic, s_tensor = ctx.meta
        cnt = f.shape[0]
        need = ctx.needs_input_grad
        df = torch.empty_like(f) if need[0] else None
        dn = torch.empty_like(n) if need[1] else None
        ds = torch.empty(1, device=f.device) if (need[4] and s_tensor) else None
        nb = ctypes.c_int64(0)
        N.check(N.lib().vxn_neus_alpha_workspace_bytes(cnt, ctypes.byref(nb)), "neus_alpha_workspace_bytes")
        ws = torch.empty(max(int(nb.value), 8), device=f.device, dtype=torch.uint8)
        da = _f32c(d_alpha.reshape(-1))
        N.check(N.lib().vxn_neus_alpha_bwd(N.dptr(f), N.dptr(n), N.dptr(d), N.dptr(dt), N.dptr(sv), cnt, car, uic,
                                           N.dptr(da), N.dptr(df), N.dptr(dn), N.dptr(ds), ws.data_ptr(), ws.numel(),
                                           _stream(f)), "neus_alpha_bwd")
        return (None if df is None else df.reshape(sshape), None if dn is None else dn.reshape(nshape), None, None,
                None if ds is None else ds.reshape(()), None, None)


# ----------------------------------------------------------------------------------------
# Radiance MLP, one point at a time (forward_points)
# ----------------------------------------------------------------------------------------

class PointMlpFunction(torch.autograd.Function):
    """RadianceMLP.forward (radiance_mlp.py:81-99) on [..., 32] hash features, [..., V]
    view embeddings (V = 16 or 27) and [..., 3] normals with the flat MLP ``mlp`` (W0, b0,
    W1, b1); ``mlp_mode`` VXN_MLP_BF16_MFMA takes the BF16 form's rounding points."""

    @staticmethod
    def forward(ctx, feats, view, normals, mlp, mlp_mode):
        lead = feats.shape[:-1]
        V = view.shape[-1]
        f = feats.reshape(-1, 32).to(torch.float32).contiguous()
        v = view.reshape(-1, V).to(torch.float32).contiguous()
        n = normals.reshape(-1, 3).to(torch.float32).contiguous()
        w = mlp.to(torch.float32).contiguous()
        if not (f.shape[0] == v.shape[0] == n.shape[0]):
            raise ValueError("feats, view and normals must have the same leading shape")
        rgb = torch.empty(f.shape[0], 3, device=f.device, dtype=torch.float32)
        N.check(N.lib().vxn_point_mlp_fwd(N.dptr(f), N.dptr(v), V, N.dptr(n), N.dptr(w), f.shape[0], int(mlp_mode),
                                          N.dptr(rgb), _stream(f)), "point_mlp_fwd")
        ctx.save_for_backward(f, v, n, w)
        ctx.meta = (lead, V, int(mlp_mode))
        return rgb.reshape(*lead, 3)

    @staticmethod
    def backward(ctx, d_rgb):
        f, v, n, w = ctx.saved_tensors
        lead, V, mode = ctx.meta
        cnt = f.shape[0]
        dr = d_rgb.reshape(-1, 3).to(torch.float32).contiguous()
        df = torch.empty_like(f) if ctx.needs_input_grad[0] else None
        dv = torch.empty_like(v) if ctx.needs_input_grad[1] else None
        dn = torch.empty_like(n) if ctx.needs_input_grad[2] else None
        dw = torch.empty_like(w)
        nb = ctypes.c_int64(0)
        N.check(N.lib().vxn_point_mlp_bwd_workspace_bytes(cnt, V, ctypes.byref(nb)), "point_mlp_bwd_workspace_bytes")
        ws = torch.empty(max(int(nb.value), 1), device=f.device, dtype=torch.uint8)
        N.check(N.lib().vxn_point_mlp_bwd(N.dptr(f), N.dptr(v), V, N.dptr(n), N.dptr(w), cnt, mode, N.dptr(dr),
                                          N.dptr(df), N.dptr(dv), N.dptr(dn), N.dptr(dw), ws.data_ptr(), ws.numel(),
                                          _stream(f)), "point_mlp_bwd")
        r = lambda t, k: None if t is None else t.reshape(*lead, k)
        return r(df, 32), r(dv, V), r(dn, 3), (dw if ctx.needs_input_grad[3] else None), None


# ----------------------------------------------------------------------------------------
# Progressive upsampler
# ----------------------------------------------------------------------------------------

_table_cache: Dict[Tuple, torch.Tensor] = {}


def _linspace_dev(a: float, b: float, n: int, device) -> torch.Tensor:
    key = (a, b, n, str(device))
    t = _table_cache.get(key)
    if t is None:
        # built by torch on the CPU exactly as the reference builds it, then copied over
        t = torch.linspace(a, b, steps=n).to(device)
        _table_cache[key] = t
    return t


def upsample(rays_o, rays_d, near, far, grid3, box, n_samples=64, n_importance=64, up_sample_steps=4,
             perturb=False, occ_mask=None, occ_box=None, t_rand=None, u_rand=None, return_inds=False,
             bound_radius=1.0, vr_bitmask=None, vr_cells=None):
    """Multi-step SDF upsampling on the GPU; returns (z_vals, dists, mid_z[, inds]).
    With ``near = far = None`` the kernel derives them from the ``bound_radius`` sphere.
    ``vr_bitmask`` (optional int32 words) receives the V_R marks of the final samples.
    ``vr_cells`` (optional, as many words, all zero; used when the grid's Z is a multiple
    of 32) takes one mark per sample cell instead, and ``vr_bitmask`` is then overwritten
    with their corners (the same bits with far fewer atomics); the cells stay marked."""
    dev = rays_o.device
    R = rays_o.shape[0]
    steps = up_sample_steps if (n_importance > 0 and up_sample_steps > 0) else 0
    per = max(n_importance // steps, 1) if steps else 0
    S = n_samples + steps * per
    z = torch.empty(R, S, device=dev, dtype=torch.float32)
    dists = torch.empty_like(z)
    mid = torch.empty_like(z)
    inds = torch.empty(steps, R, per, device=dev, dtype=torch.int32) if (return_inds and steps) else None
    ro, rd = rays_o.contiguous(), rays_d.contiguous()
    nr = None if near is None else near.reshape(-1).contiguous()
    fr = None if far is None else far.reshape(-1).contiguous()
    if perturb:
        if t_rand is None:
            t_rand = torch.rand(R, n_samples, device=dev)
        if u_rand is None and steps:
            u_rand = torch.rand(steps, R, per, device=dev)
        t_rand = t_rand.contiguous()
        u_rand = u_rand.contiguous() if u_rand is not None else None
    a = N.SamplerArgs()
    a.rays_o, a.rays_d, a.near, a.far = N.dptr(ro, torch.float32), N.dptr(rd, torch.float32), N.dptr(nr), N.dptr(fr)
    a.n_rays = R
    g3 = grid3.contiguous()
    a.grid = N.dptr(g3, torch.float32, "grid")
    a.X, a.Y, a.Z = g3.shape
    a.box = N.box6(*box)
    if occ_mask is not None:
        om = occ_mask.contiguous()
        if om.dtype != torch.bool:
            raise TypeError("occupancy mask must be bool")
        a.occ_mask = om.data_ptr()
        a.occ_res = om.shape[0]
        a.occ_box = N.box6(*(occ_box if occ_box is not None else box))
    a.n_samples, a.n_importance, a.up_sample_steps, a.perturb = n_samples, n_importance, up_sample_steps, int(perturb)
    a.t_lin = N.dptr(_linspace_dev(0.0, 1.0, n_samples, dev))
    if steps:
        a.u_det = N.dptr(_linspace_dev(0.5 / per, 1.0 - 0.5 / per, per, dev))
    a.t_rand = N.dptr(t_rand) if perturb else None
    a.u_rand = N.dptr(u_rand) if (perturb and steps) else None
    a.z_vals, a.dists, a.mid_z = N.dptr(z), N.dptr(dists), N.dptr(mid)
    a.inds = N.dptr(inds)
    a.vr_bitmask = N.dptr(vr_bitmask)
    if vr_bitmask is not None and vr_cells is not None:
        if vr_cells.numel() < vr_bitmask.numel() or vr_cells.dtype != vr_bitmask.dtype:
            raise ValueError("vr_cells must match vr_bitmask's words")
        a.vr_cells = N.dptr(vr_cells)
    a.bound_radius = float(bound_radius)
    N.check(N.lib().vxn_upsample(ctypes.byref(a), _stream(ro)), "upsample")
    if return_inds:
        return z, dists, mid, inds
    return z, dists, mid


# ----------------------------------------------------------------------------------------
# Fused render
# ----------------------------------------------------------------------------------------

class RenderConfig:
    """Static settings of one fused render call."""

    def __init__(self, box, level_res, level_w=None, sphere_radius=1.0, cos_anneal_ratio=0.0,
                 use_iter_cos=False, mlp_mode=N.VXN_MLP_FP32, table_grad_out=None,
                 gradient_mode=N.VXN_GRADIENT_INTERPOLATED, view_embed=N.VXN_VIEW_SH, grid_binned=False):
        # table_grad_out: optional FP32 [L,T,F] tensor that the backward accumulates the
        # hash-table gradient into (for bf16 tables with an FP32 master copy); the table
        # input then receives no gradient through autograd
        self.table_grad_out = table_grad_out
        self.box = tuple(float(v) for v in box)
        self.level_res = list(level_res)
        self.level_w = None if level_w is None else list(level_w)
        self.sphere_radius = float(sphere_radius) if sphere_radius is not None else 0.0
        self.cos_anneal_ratio = float(cos_anneal_ratio)
        self.use_iter_cos = bool(use_iter_cos)
        self.mlp_mode = mlp_mode
        self.gradient_mode = int(gradient_mode)
        self.view_embed = int(view_embed)  # VXN_VIEW_SH or VXN_VIEW_PE
        # the backward's grid gradient voxel-binned (vxn_render_args.grid_binned) instead of
        # the atomic scatter
        self.grid_binned = bool(grid_binned)


def fill_render_args(a: "N.RenderArgs", cfg: RenderConfig, grid3, table, mlp_flat, s: float, rays_o, rays_d,
                     mid, dists):
    R, S = mid.shape
    a.rays_o, a.rays_d = N.dptr(rays_o, torch.float32, "rays_o"), N.dptr(rays_d, torch.float32, "rays_d")
    a.mid_z, a.dists = N.dptr(mid, torch.float32, "mid_z"), N.dptr(dists, torch.float32, "dists")
    a.n_rays, a.n_samples = R, S
    a.grid = N.dptr(grid3, torch.float32, "grid")
    a.X, a.Y, a.Z = grid3.shape
    a.box = N.box6(cfg.box[:3], cfg.box[3:])
    L, T, F = table.shape
    a.table, a.table_dtype = N.dptr(table), _table_dtype(table)
    a.n_levels, a.log2_T, a.n_features = L, int(round(math.log2(T))), F
    for i, r in enumerate(cfg.level_res):
        a.level_res[i] = int(r)
        a.level_w[i] = 1.0 if cfg.level_w is None else float(cfg.level_w[i])
    if mlp_flat.numel() != N.mlp_params(cfg.view_embed):
        raise ValueError(f"the MLP has {mlp_flat.numel()} parameters; the view embedding needs "
                         f"{N.mlp_params(cfg.view_embed)}")
    a.mlp = N.dptr(mlp_flat, torch.float32, "mlp")
    a.mlp_mode = cfg.mlp_mode
    a.view_embed = cfg.view_embed
    a.s = float(s)
    a.sphere_radius = cfg.sphere_radius
    a.cos_anneal_ratio = cfg.cos_anneal_ratio
    a.use_iter_cos = int(cfg.use_iter_cos)
    a.gradient_mode = cfg.gradient_mode
    a.grid_binned = int(cfg.grid_binned)


def render_workspace(a: "N.RenderArgs", device) -> torch.Tensor:
    nb = ctypes.c_int64(0)
    N.check(N.lib().vxn_render_workspace_bytes(ctypes.byref(a), ctypes.byref(nb)), "render_workspace_bytes")
    return torch.empty(max(int(nb.value), 8), device=device, dtype=torch.uint8)


class RenderFunction(torch.autograd.Function):
    """Fused render of passes 2-3.  Differentiable outputs: rgb, depth, normal, opacity,
    weights; the per-sample diagnostics (alpha, transmittance, sdf/grad/normal samples)
    are returned without gradient."""

    @staticmethod
    def forward(ctx, grid, table, mlp_flat, s, rays_o, rays_d, mid, dists, cfg: RenderConfig, want_corners,
                vr_bitmask):
        g3 = _grid3(grid).contiguous()
        tb = table.contiguous()
        mf = mlp_flat.contiguous()
        ro, rd, md, ds = rays_o.contiguous(), rays_d.contiguous(), mid.contiguous(), dists.contiguous()
        R, S = md.shape
        dev = md.device
        f = lambda *sh: torch.empty(*sh, device=dev, dtype=torch.float32)
        out = {"rgb": f(R, 3), "depth": f(R, 1), "normal": f(R, 3), "opacity": f(R, 1), "weights": f(R, S),
               "alpha": f(R, S), "transmittance": f(R, S), "sdf_samples": f(R, S, 1), "grad_samples": f(R, S, 3),
               "normal_samples": f(R, S, 3)}
        corners = torch.empty(R, S, 8, 3, device=dev, dtype=torch.int64) if want_corners else None
        a = N.RenderArgs()
        s_val = float(s.detach()) if torch.is_tensor(s) else float(s)
        fill_render_args(a, cfg, g3, tb, mf, s_val, ro, rd, md, ds)
        for k in ("rgb", "depth", "normal", "opacity", "weights", "alpha", "transmittance"):
            setattr(a, k, N.dptr(out[k]))
        a.sdf_samples, a.grad_samples, a.normal_samples = (N.dptr(out["sdf_samples"]), N.dptr(out["grad_samples"]),
                                                           N.dptr(out["normal_samples"]))
        a.corner_indices = N.dptr(corners)
        a.vr_bitmask = N.dptr(vr_bitmask)
        N.check(N.lib().vxn_render_forward(ctypes.byref(a), _stream(md)), "render_forward")
        ctx.save_for_backward(g3, tb, mf, ro, rd, md, ds)
        ctx.cfg, ctx.s_val, ctx.grid_shape = cfg, s_val, grid.shape
        ctx.mark_non_differentiable(out["alpha"], out["transmittance"], out["sdf_samples"], out["grad_samples"],
                                    out["normal_samples"])
        if corners is not None:
            ctx.mark_non_differentiable(corners)
        return (out["rgb"], out["depth"], out["normal"], out["opacity"], out["weights"], out["alpha"],
                out["transmittance"], out["sdf_samples"], out["grad_samples"], out["normal_samples"], corners)

    @staticmethod
    def backward(ctx, d_rgb, d_depth, d_normal, d_opacity, d_weights, *unused):
        g3, tb, mf, ro, rd, md, ds = ctx.saved_tensors
        dev = md.device
        a = N.RenderArgs()
        fill_render_args(a, ctx.cfg, g3, tb, mf, ctx.s_val, ro, rd, md, ds)
        c = lambda t: None if t is None else t.contiguous()
        d_rgb, d_depth, d_normal, d_opacity, d_weights = map(c, (d_rgb, d_depth, d_normal, d_opacity, d_weights))
        a.d_rgb, a.d_depth, a.d_normal, a.d_opacity, a.d_weights = (N.dptr(d_rgb), N.dptr(d_depth), N.dptr(d_normal),
                                                                     N.dptr(d_opacity), N.dptr(d_weights))
        gg = torch.zeros_like(g3)
        gt = ctx.cfg.table_grad_out
        if gt is None:
            gt = torch.zeros(tb.shape, device=dev, dtype=torch.float32)
        elif gt.shape != tb.shape or gt.dtype != torch.float32 or not gt.is_contiguous():
            raise ValueError("table_grad_out must be a contiguous float32 tensor of the table's shape")
        gm = torch.empty(mf.numel(), device=dev, dtype=torch.float32)
        gs = torch.empty(1, device=dev, dtype=torch.float32)
        a.grad_grid, a.grad_table, a.grad_mlp, a.grad_s = N.dptr(gg), N.dptr(gt), N.dptr(gm), N.dptr(gs)
        ws = render_workspace(a, dev)
        a.workspace, a.workspace_bytes = ws.data_ptr(), ws.numel()
        N.check(N.lib().vxn_render_backward(ctypes.byref(a), _stream(md)), "render_backward")
        gtab = None if ctx.cfg.table_grad_out is not None else gt.to(tb.dtype)
        return (gg.reshape(ctx.grid_shape), gtab, gm, gs.reshape(()), None, None, None, None, None, None, None)


# ----------------------------------------------------------------------------------------
# V_R and regularisers
# ----------------------------------------------------------------------------------------

def vr_bitmask_words(grid_shape) -> int:
    X, Y, Z = grid_shape
    return (X * Y * Z + 31) // 32


def vr_mark_corners(corners: torch.Tensor, grid_shape, bitmask: torch.Tensor):
    X, Y, Z = grid_shape
    c = corners.reshape(-1, 3).to(torch.int64).contiguous()
    N.check(N.lib().vxn_vr_mark_corners(N.dptr(c), c.shape[0], X, Y, Z, N.dptr(bitmask, torch.int32, "bitmask"),
                                        _stream(c)), "vr_mark_corners")


def vr_compact(bitmask: torch.Tensor, grid_shape, box, relax_radius: float = 1.2):
    """-> (list int32 [X*Y*Z] with bit 31 = outside the relaxed sphere, counts int32 [2])."""
    X, Y, Z = grid_shape
    dev = bitmask.device
    lst = torch.empty(X * Y * Z, device=dev, dtype=torch.int32)
    counts = torch.zeros(2, device=dev, dtype=torch.int32)
    nb = ctypes.c_int64(0)
    N.check(N.lib().vxn_vr_workspace_bytes(X, Y, Z, ctypes.byref(nb)), "vr_workspace_bytes")
    ws = torch.empty(max(int(nb.value), 8), device=dev, dtype=torch.uint8)
    N.check(N.lib().vxn_vr_compact(N.dptr(bitmask, torch.int32, "bitmask"), X, Y, Z, N.box6(*box), float(relax_radius),
                                   N.dptr(lst), N.dptr(counts), ws.data_ptr(), ws.numel(), _stream(bitmask)),
            "vr_compact")
    return lst, counts


def voxel_size_of(box6, grid_shape):
    """(xyz_max - xyz_min) / (res - 1) in float32, as DenseSDFGrid.voxel_size computes it."""
    mn = torch.tensor(box6[:3], dtype=torch.float32)
    mx = torch.tensor(box6[3:], dtype=torch.float32)
    return ((mx - mn) / (torch.tensor(list(grid_shape), dtype=torch.float32) - 1.0)).tolist()


def reg_eik_curv(grid3, eps3, vr_list, counts, w_eik, w_curv, grad_grid):
    """Fused eikonal + curvature over V_R; returns FP64 losses [2] (device)."""
    X, Y, Z = grid3.shape
    dev = grid3.device
    losses = torch.empty(2, device=dev, dtype=torch.float64)
    nb = ctypes.c_int64(0)
    N.check(N.lib().vxn_reg_workspace_bytes(X, Y, Z, ctypes.byref(nb)), "reg_workspace_bytes")
    ws = torch.empty(int(nb.value), device=dev, dtype=torch.uint8)
    N.check(N.lib().vxn_reg_eik_curv(N.dptr(grid3, torch.float32, "grid"), X, Y, Z, N.f3(eps3), N.dptr(vr_list),
                                     N.dptr(counts), float(w_eik), float(w_curv),
                                     N.dptr(grad_grid, torch.float32, "grad_grid"), N.dptr(losses), ws.data_ptr(),
                                     ws.numel(), _stream(grid3)), "reg_eik_curv")
    return losses


def reg_eik_curv_vr(grid3, eps3, box, relax_radius, bitmask, w_eik, w_curv, grad_grid, workspace=None,
                    losses=None, counts=None):
    """Eikonal + curvature straight from the V_R bitmask (atomic-free gather); returns
    (FP64 losses [2], int32 counts [2] = {|V_R|, |V_R inside the relaxed sphere|})."""
    X, Y, Z = grid3.shape
    dev = grid3.device
    if losses is None:
        losses = torch.empty(2, device=dev, dtype=torch.float64)
    if counts is None:
        counts = torch.empty(2, device=dev, dtype=torch.int32)
    if workspace is None:
        nb = ctypes.c_int64(0)
        N.check(N.lib().vxn_reg_vr_workspace_bytes(X, Y, Z, ctypes.byref(nb)), "reg_vr_workspace_bytes")
        workspace = torch.empty(int(nb.value), device=dev, dtype=torch.uint8)
    N.check(N.lib().vxn_reg_eik_curv_vr(N.dptr(grid3, torch.float32, "grid"), X, Y, Z, N.f3(eps3), N.box6(*box),
                                        float(relax_radius), N.dptr(bitmask, torch.int32, "bitmask"), float(w_eik),
                                        float(w_curv), N.dptr(grad_grid, torch.float32, "grad_grid"), N.dptr(losses),
                                        N.dptr(counts), workspace.data_ptr(), workspace.numel(), _stream(grid3)),
            "reg_eik_curv_vr")
    return losses, counts


def eikonal_grad_ref(grid3, active, grad3, eps3, weight):
    X, Y, Z = grid3.shape
    a = active.to(torch.int64).contiguous()
    N.check(N.lib().vxn_eikonal_grad(N.dptr(grid3.contiguous()), X, Y, Z, N.dptr(a), a.shape[0], N.dptr(grad3),
                                     N.f3(eps3), float(weight), _stream(grid3)), "eikonal_grad")


def curvature_grad_ref(grid3, active, grad3, eps3, weight):
    X, Y, Z = grid3.shape
    a = active.to(torch.int64).contiguous()
    N.check(N.lib().vxn_curvature_grad(N.dptr(grid3.contiguous()), X, Y, Z, N.dptr(a), a.shape[0], N.dptr(grad3),
                                       N.f3(eps3), float(weight), _stream(grid3)), "curvature_grad")


def s_amplify(grad_v: torch.Tensor, k: float, amplified: Optional[torch.Tensor] = None):
    N.check(N.lib().vxn_s_amplify(N.dptr(grad_v, torch.float32, "grad_v"), float(k), N.dptr(amplified),
                                  _stream(grad_v)), "s_amplify")


def eik_clamp_workspace(device) -> torch.Tensor:
    nb = ctypes.c_int64(0)
    N.check(N.lib().vxn_eik_clamp_add_workspace_bytes(ctypes.byref(nb)), "eik_clamp_add_workspace_bytes")
    return torch.empty(int(nb.value), device=device, dtype=torch.uint8)


def eik_clamp_add(dst: torch.Tensor, src: torch.Tensor, weight: float, clamp: float,
                  workspace: Optional[torch.Tensor] = None) -> None:
    """dst += src * weight, scaled down to max-abs ``clamp`` when it exceeds it
    (trainer.py:309-316, the post-upsample cooldown); on dst's current stream."""
    if dst.shape != src.shape:
        raise ValueError(f"dst {tuple(dst.shape)} and src {tuple(src.shape)} differ")
    if workspace is None:
        workspace = eik_clamp_workspace(dst.device)
    N.check(N.lib().vxn_eik_clamp_add(N.dptr(dst, torch.float32, "dst"), N.dptr(src, torch.float32, "src"),
                                      dst.numel(), float(weight), float(clamp), workspace.data_ptr(),
                                      workspace.numel(), _stream(dst)), "eik_clamp_add")


# ----------------------------------------------------------------------------------------
# Occupancy grid EMA update
# ----------------------------------------------------------------------------------------

def occupancy_update(density: torch.Tensor, mask: torch.Tensor, occ_box, grid3: torch.Tensor, box, s: float,
                     decay: float, threshold: float, first_update: bool, idx=None,
                     workspace: Optional[torch.Tensor] = None):
    """OccupancyGrid.update_from_sdf (occupancy_grid.py:62-119) on the device.  ``idx`` =
    (idx_x, idx_y, idx_z) int64 device tensors of the cells to update, or None for every
    cell.  ``mask`` is the bool occupancy mask, updated in place with ``density``."""
    res = density.shape[0]
    if tuple(density.shape) != (res, res, res) or tuple(mask.shape) != (res, res, res):
        raise ValueError("density and mask must be [res, res, res]")
    if mask.dtype != torch.bool or not mask.is_contiguous() or not density.is_contiguous():
        raise ValueError("mask must be a contiguous bool tensor, density contiguous float32")
    g3 = grid3.contiguous()
    X, Y, Z = g3.shape
    ix = iy = iz = None
    n = 0
    ws = None
    if idx is not None:
        ix, iy, iz = (t.contiguous() for t in idx)
        n = ix.numel()
        if iy.numel() != n or iz.numel() != n:
            raise ValueError("the three index arrays must have the same length")
        nb = ctypes.c_int64(0)
        N.check(N.lib().vxn_occupancy_update_workspace_bytes(n, ctypes.byref(nb)))
        ws = workspace
        if ws is None or ws.numel() * ws.element_size() < nb.value:
            ws = torch.empty(max(int(nb.value), 1), device=density.device, dtype=torch.uint8)
    N.check(N.lib().vxn_occupancy_update(
        N.dptr(density, torch.float32, "density"), mask.data_ptr(), res, N.box6(*occ_box),
        N.dptr(g3, torch.float32, "grid"), X, Y, Z, N.box6(*box), float(s), float(decay), float(threshold),
        int(bool(first_update)), N.dptr(ix, torch.int64, "idx_x"), N.dptr(iy, torch.int64, "idx_y"),
        N.dptr(iz, torch.int64, "idx_z"), n, N.dptr(ws), 0 if ws is None else ws.numel() * ws.element_size(),
        _stream(density)), "occupancy_update")


def upsample_grid(grid: torch.Tensor, new_shape) -> torch.Tensor:
    """Trilinear, aligned-corner resampling of an [X,Y,Z] (or [1,1,X,Y,Z]) grid to
    ``new_shape`` (upsample.py:12-39); returns a new [NX,NY,NZ] tensor."""
    g3 = _grid3(grid).contiguous()
    NX, NY, NZ = (int(v) for v in new_shape)
    out = torch.empty(NX, NY, NZ, device=g3.device, dtype=torch.float32)
    X, Y, Z = g3.shape
    N.check(N.lib().vxn_upsample_grid(N.dptr(g3, torch.float32, "grid"), X, Y, Z, N.dptr(out), NX, NY, NZ,
                                      _stream(g3)), "upsample_grid")
    return out


def gaussian_weights(sigma: float, truncate: float = 4.0):
    """The correlation weights scipy.ndimage.gaussian_filter1d uses (order 0): radius
    int(truncate * sigma + 0.5), exp(-0.5 / sigma^2 * x^2) normalised to sum 1, in
    float64 -> (weights [2r+1] numpy float64, r)."""
    sd = float(sigma)
    r = int(truncate * sd + 0.5)
    x = np.arange(-r, r + 1)
    phi = np.exp(-0.5 / (sd * sd) * x ** 2)
    return np.ascontiguousarray((phi / phi.sum())[::-1]), r


def gaussian_filter_grid(grid: torch.Tensor, sigma: float, truncate: float = 4.0) -> torch.Tensor:
    """Separable Gaussian smoothing of an [X,Y,Z] (or [1,1,X,Y,Z]) grid with edge
    replication (mesh.py:17-31, scipy's mode="nearest"); returns a new [X,Y,Z] tensor."""
    g3 = _grid3(grid).contiguous()
    X, Y, Z = g3.shape
    if float(sigma) <= 1e-15:  # scipy skips every axis
        return g3.clone()
    w, r = gaussian_weights(sigma, truncate)
    out = torch.empty_like(g3)
    nb = ctypes.c_int64(0)
    N.check(N.lib().vxn_gaussian_filter_workspace_bytes(X, Y, Z, ctypes.byref(nb)), "gaussian_filter_workspace")
    ws = torch.empty(int(nb.value), device=g3.device, dtype=torch.uint8)
    N.check(N.lib().vxn_gaussian_filter(N.dptr(g3, torch.float32, "grid"), X, Y, Z,
                                        w.ctypes.data_as(ctypes.c_void_p), r, N.dptr(out), ws.data_ptr(), ws.numel(),
                                        _stream(g3)), "gaussian_filter")
    return out
