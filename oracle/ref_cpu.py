"""CPU oracle for the VoxNeuS voxel-SDF rendering and regularisation path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the checker or
the timed CPU baseline.  The product path (``voxneus_amd``) never calls it.

This is an op-for-op restatement, in PyTorch on the CPU, of the reference's pure-PyTorch
path (``force_pytorch=True``, no compiled backend).  Every function cites the reference
``file:line`` it follows (paths relative to the reference checkout).  The same code runs
in float32 (the reference's own arithmetic: it is pinned bit-for-bit against golden
vectors written by the reference) and in float64 (the error yardstick used to state the
FP32 / BF16 device error).

Pinning (what checks each function, see ``tests/test_oracle.py``):
  * corners/weights, grid_sample SDF, Eq.7 normals, hash encoding, SH, MLP, alpha,
    compositing, sampler, V_R, eikonal, curvature, render and the train step:
    golden vectors written by the reference itself (``tests/golden/``);
  * additionally by closed forms: linear SDF -> exact constant normal, one-sided
    borders, sphere SDF zero set, alpha peak/bounds, transmittance cutoff, hash
    brute force on tiny inputs, eikonal/curvature vs float64 autograd, searchsorted
    semantics, sortedness of z.
"""

from __future__ import annotations

import math
from typing import Dict, List, Optional, Tuple

import torch
import torch.nn.functional as F


# ----------------------------------------------------------------------------------------
# Coordinates (Code/voxneus/geometry/coordinates.py)
# ----------------------------------------------------------------------------------------

def world_to_grid_index(xyz, xyz_min, xyz_max, grid_res):
    """coordinates.py:10-29."""
    norm = (xyz - xyz_min) / (xyz_max - xyz_min + 1e-8)
    return norm * (grid_res - 1.0)


def get_voxel_corners_and_weights(coords, grid_res):
    """coordinates.py:55-124: clamp to [0, res-1.0001], floor, 8 corners in binary
    order k = 4*ox + 2*oy + oz, trilinear weights."""
    rx, ry, rz = int(grid_res[0]), int(grid_res[1]), int(grid_res[2])
    max_bound = torch.tensor([rx - 1.0001, ry - 1.0001, rz - 1.0001], device=coords.device)
    min_bound = torch.zeros_like(max_bound)
    cc = torch.clamp(coords, min=min_bound, max=max_bound)
    c0 = torch.floor(cc).to(torch.int64)
    c1 = c0 + 1
    for a, r in enumerate((rx, ry, rz)):
        c0[..., a] = torch.clamp(c0[..., a], 0, r - 1)
        c1[..., a] = torch.clamp(c1[..., a], 0, r - 1)
    d = cc - c0.to(cc.dtype)
    dx, dy, dz = d[..., 0:1], d[..., 1:2], d[..., 2:3]
    ox, oy, oz = 1.0 - dx, 1.0 - dy, 1.0 - dz
    corners = []
    for k in range(8):
        sx, sy, sz = (k >> 2) & 1, (k >> 1) & 1, k & 1
        corners.append(torch.stack([(c1 if sx else c0)[..., 0], (c1 if sy else c0)[..., 1],
                                    (c1 if sz else c0)[..., 2]], dim=-1))
    corner_indices = torch.stack(corners, dim=-2)
    w = [ox * oy * oz, ox * oy * dz, ox * dy * oz, ox * dy * dz,
         dx * oy * oz, dx * oy * dz, dx * dy * oz, dx * dy * dz]
    return corner_indices, torch.cat(w, dim=-1)


# ----------------------------------------------------------------------------------------
# Dense SDF grid (Code/voxneus/models/sdf_grid.py, Code/voxneus/geometry/gradients.py)
# ----------------------------------------------------------------------------------------

class GridBox:
    """The grid's box and derived constants, built exactly as the reference builds them."""

    def __init__(self, res, xyz_min=(-1.0, -1.0, -1.0), xyz_max=(1.0, 1.0, 1.0), dtype=torch.float32):
        if isinstance(res, int):
            res = (res, res, res)
        self.res = torch.tensor(res, dtype=torch.float32)  # sdf_grid.py:39-42 (always float32)
        self.xyz_min = torch.tensor(xyz_min, dtype=dtype)  # sdf_grid.py:44-45
        self.xyz_max = torch.tensor(xyz_max, dtype=dtype)

    @property
    def voxel_size(self):
        """sdf_grid.py:83-88."""
        return (self.xyz_max - self.xyz_min) / (self.res.to(self.xyz_min.dtype) - 1.0)


def query_sdf(grid5, box: GridBox, xyz):
    """sdf_grid.py:90-120: F.grid_sample on (z,y,x)-flipped [-1,1] coordinates."""
    orig = xyz.shape[:-1]
    flat = xyz.reshape(1, 1, 1, -1, 3)
    norm = (flat - box.xyz_min) / (box.xyz_max - box.xyz_min + 1e-8)
    norm = norm * 2.0 - 1.0
    out = F.grid_sample(grid5, norm.flip(-1), mode="bilinear", padding_mode="border", align_corners=True)
    return out.reshape(1, -1).T.reshape(*orig, 1)


def query_vertex_normals_at_corners(grid5, corner_indices, voxel_size):
    """gradients.py:59-100: per-corner central difference, one-sided at borders."""
    data = grid5.squeeze(0).squeeze(0)
    rx, ry, rz = data.shape
    ex, ey, ez = voxel_size[0], voxel_size[1], voxel_size[2]
    ix, iy, iz = corner_indices[..., 0], corner_indices[..., 1], corner_indices[..., 2]
    ixp, ixm = torch.clamp(ix + 1, max=rx - 1), torch.clamp(ix - 1, min=0)
    dxn = torch.clamp((ixp - ixm).to(data.dtype) * ex, min=ex)
    gx = (data[ixp, iy, iz] - data[ixm, iy, iz]) / dxn
    iyp, iym = torch.clamp(iy + 1, max=ry - 1), torch.clamp(iy - 1, min=0)
    dyn = torch.clamp((iyp - iym).to(data.dtype) * ey, min=ey)
    gy = (data[ix, iyp, iz] - data[ix, iym, iz]) / dyn
    izp, izm = torch.clamp(iz + 1, max=rz - 1), torch.clamp(iz - 1, min=0)
    dzn = torch.clamp((izp - izm).to(data.dtype) * ez, min=ez)
    gz = (data[ix, iy, izp] - data[ix, iy, izm]) / dzn
    return torch.stack([gx, gy, gz], dim=-1)


def interpolate_gradients(grid5, coords, voxel_size, grid_res):
    """gradients.py:103-139 (pure-PyTorch branch): grad = sum_k w_k n_k, normal = grad/(|grad|+1e-8)."""
    corners, weights = get_voxel_corners_and_weights(coords, grid_res)
    cn = query_vertex_normals_at_corners(grid5, corners, voxel_size)
    grad = torch.sum(weights.unsqueeze(-1) * cn, dim=-2)
    norm = torch.linalg.norm(grad, dim=-1, keepdim=True)
    return grad, grad / (norm + 1e-8), corners, weights


def query_normals(grid5, box: GridBox, xyz, gradient_mode="interpolated"):
    """sdf_grid.py:122-160: Eq. 7 ("interpolated") or Eq. 4 ("analytical"), with the
    corner indices and trilinear weights of either; any other mode is a ValueError."""
    coords = world_to_grid_index(xyz, box.xyz_min, box.xyz_max, box.res.to(xyz.dtype))
    mode = gradient_mode.lower()
    if mode == "interpolated":
        return interpolate_gradients(grid5, coords, box.voxel_size, box.res)
    if mode == "analytical":
        grad, normal = compute_analytical_gradient(grid5, coords, box.voxel_size, box.res)
        corners, weights = get_voxel_corners_and_weights(coords, box.res)
        return grad, normal, corners, weights
    raise ValueError(f"Unknown gradient mode: {gradient_mode}")


def upsample_sdf_grid(grid5, new_res):
    """upsample.py:12-39: F.interpolate(trilinear, align_corners=True) to the new shape."""
    shape = (new_res,) * 3 if isinstance(new_res, int) else tuple(new_res)
    return F.interpolate(grid5.detach(), size=shape, mode="trilinear", align_corners=True)


def compute_analytical_gradient(grid5, coords, voxel_size, grid_res):
    """gradients.py:142-220 (Eq. 4): d/dx of the trilinear interpolant of the cell's 8
    corner values, the ablation baseline of VoxNeuS(gradient_mode="analytical")."""
    corners, w = get_voxel_corners_and_weights(coords, grid_res)
    data = grid5.squeeze(0).squeeze(0)
    rx, ry, rz = int(grid_res[0]), int(grid_res[1]), int(grid_res[2])
    mb = torch.tensor([rx - 1.0001, ry - 1.0001, rz - 1.0001])
    cc = torch.clamp(coords, min=torch.zeros_like(mb), max=mb)
    c0 = torch.floor(cc)
    d = cc - c0
    dx, dy, dz = d[..., 0:1], d[..., 1:2], d[..., 2:3]
    ox, oy, oz = 1.0 - dx, 1.0 - dy, 1.0 - dz
    ex, ey, ez = voxel_size[0], voxel_size[1], voxel_size[2]
    dwx = torch.cat([-oy * oz, -oy * dz, -dy * oz, -dy * dz, oy * oz, oy * dz, dy * oz, dy * dz], -1) / ex
    dwy = torch.cat([-ox * oz, -ox * dz, ox * oz, ox * dz, -dx * oz, -dx * dz, dx * oz, dx * dz], -1) / ey
    dwz = torch.cat([-ox * oy, ox * oy, -ox * dy, ox * dy, -dx * oy, dx * oy, -dx * dy, dx * dy], -1) / ez
    vals = data[corners[..., 0], corners[..., 1], corners[..., 2]]
    g = torch.cat([torch.sum(dwx * vals, -1, keepdim=True), torch.sum(dwy * vals, -1, keepdim=True),
                   torch.sum(dwz * vals, -1, keepdim=True)], -1)
    return g, g / (torch.linalg.norm(g, dim=-1, keepdim=True) + 1e-8)


# ----------------------------------------------------------------------------------------
# Hash grid (Code/voxneus/models/radiance_hash.py)
# ----------------------------------------------------------------------------------------

PRIMES = (1, 2654435761, 805459861)  # radiance_hash.py:53


def hash_level_resolutions(n_levels=16, base_resolution=16, max_resolution=2048.0) -> List[int]:
    """radiance_hash.py:49,108: b = exp((ln max - ln base)/(L-1)), N_l = floor(base * b**l), in double."""
    b = math.exp((math.log(max_resolution) - math.log(base_resolution)) / max(n_levels - 1, 1))
    return [math.floor(base_resolution * (b ** lvl)) for lvl in range(n_levels)]


def level_weight(level: int, active_levels: Optional[float]) -> float:
    """radiance_hash.py:90-102 (NeuS2 lambda mask)."""
    if active_levels is None:
        return 1.0
    li = level + 1.0
    if li <= active_levels:
        return 1.0
    if li - 1.0 < active_levels < li:
        return active_levels - (li - 1.0)
    return 0.0


def hash_index(c, T):
    """radiance_hash.py:65-69: (x*1) ^ (y*2654435761) ^ (z*805459861) mod T."""
    h = (c[..., 0] * PRIMES[0]) ^ (c[..., 1] * PRIMES[1]) ^ (c[..., 2] * PRIMES[2])
    return torch.remainder(h, T)


def hash_encode(x01, embeddings, level_res: List[int], active_levels=None):
    """radiance_hash.py:71-166 on already-clamped [0,1] coordinates (clamp is :213)."""
    L, T, Fd = embeddings.shape
    orig = x01.shape[:-1]
    xf = x01.reshape(-1, 3)
    outs = []
    for lvl in range(L):
        lw = level_weight(lvl, active_levels)
        if lw == 0.0:
            outs.append(torch.zeros((xf.shape[0], Fd), dtype=embeddings.dtype))
            continue
        res_t = torch.tensor(level_res[lvl], dtype=xf.dtype)
        s = xf * res_t
        c0 = torch.floor(s).to(torch.int64)
        c1 = c0 + 1
        d = s - c0.to(s.dtype)
        dx, dy, dz = d[:, 0:1], d[:, 1:2], d[:, 2:3]
        ox, oy, oz = 1.0 - dx, 1.0 - dy, 1.0 - dz
        emb = embeddings[lvl]
        ws = [ox * oy * oz, ox * oy * dz, ox * dy * oz, ox * dy * dz,
              dx * oy * oz, dx * oy * dz, dx * dy * oz, dx * dy * dz]
        feat = None
        for k in range(8):
            sx, sy, sz = (k >> 2) & 1, (k >> 1) & 1, k & 1
            c = torch.stack([(c1 if sx else c0)[:, 0], (c1 if sy else c0)[:, 1], (c1 if sz else c0)[:, 2]], -1)
            term = ws[k] * emb[hash_index(c, T)]
            feat = term if feat is None else feat + term
        if lw < 1.0:
            feat = feat * lw
        outs.append(feat)
    return torch.cat(outs, -1).reshape(*orig, L * Fd)


# ----------------------------------------------------------------------------------------
# View embedding, MLP, variance (Code/voxneus/models/{embedder,radiance_mlp,variance}.py)
# ----------------------------------------------------------------------------------------

SH_C = (0.28209479177387814, 0.4886025119029199, 1.0925484305920792, 0.31539156525252005,
        0.5462742152960396, 0.5900435899266435, 2.890611442640554, 0.4570457994644658,
        0.3731763325901154, 1.445305721320277)


def sh_embed(dirs):
    """embedder.py:55-87, degree 4 (16 outputs)."""
    x, y, z = dirs[..., 0:1], dirs[..., 1:2], dirs[..., 2:3]
    C = SH_C
    r = [torch.full_like(x, C[0]),
         -C[1] * y, C[1] * z, -C[1] * x,
         C[2] * x * y, -C[2] * y * z, C[3] * (2.0 * z * z - x * x - y * y), -C[2] * x * z,
         C[4] * (x * x - y * y),
         -C[5] * y * (3.0 * x * x - y * y), C[6] * x * y * z, -C[7] * y * (4.0 * z * z - x * x - y * y),
         C[8] * z * (2.0 * z * z - 3.0 * x * x - 3.0 * y * y), -C[7] * x * (4.0 * z * z - x * x - y * y),
         C[9] * z * (x * x - y * y), -C[5] * x * (x * x - 3.0 * y * y)]
    return torch.cat(r, -1)


def pe_embed(dirs):
    """embedder.py:12-44 with num_freqs 4, include_input, log sampling (voxneus.py:81):
    [d, sin(d f), cos(d f) for f in 2 ** linspace(0, 3, 4)], 27 outputs."""
    freq_bands = 2.0 ** torch.linspace(0.0, 3.0, 4)
    outputs = [dirs]
    for freq in freq_bands:
        outputs.append(torch.sin(dirs * freq))
        outputs.append(torch.cos(dirs * freq))
    return torch.cat(outputs, dim=-1)


def view_embed(dirs, view_embed_type="sh"):
    """voxneus.py:78-82: the model's view embedder."""
    t = view_embed_type.lower()
    if t == "sh":
        return sh_embed(dirs)
    if t == "pe":
        return pe_embed(dirs)
    raise ValueError(f"unknown view_embed_type {view_embed_type!r}")


def radiance_mlp(hash_feats, view_embed, normals, W0, b0, W1, b1):
    """radiance_mlp.py:81-99 with the 2-layer net of :36-53."""
    x = torch.cat([hash_feats, view_embed, normals], -1)
    return torch.sigmoid(F.linear(torch.relu(F.linear(x, W0, b0)), W1, b1))


def _bf16(t):
    return t.to(torch.bfloat16).to(t.dtype)


class _RadianceMlpBF16(torch.autograd.Function):
    """The BF16 matrix-core form of the radiance MLP (this build's definition, not the
    reference's): bf16 operands and FP32 accumulation at the points the device kernel
    rounds.  Forward: the 35 per-sample inputs (hash features, normal) and W0's columns
    for them are rounded to bf16; the view-embedding part b0 + W0[:, 32:32+V] . view stays
    FP32 (V = 16 SH or 27 positional-encoding values); relu(h) and W1 are rounded to bf16.
    Backward: the output gradient and dL/dh are rounded to bf16; dW0's view columns use
    bf16 view values; db1 sums the unrounded output gradient."""

    @staticmethod
    def forward(ctx, feats, sh, nrm, W0, b0, W1, b1, rounding=True):
        rnd = _bf16 if rounding else (lambda t: t)
        V = sh.shape[-1]
        x = rnd(torch.cat([feats, nrm], -1))
        W0x = rnd(torch.cat([W0[:, :32], W0[:, 32 + V:35 + V]], 1))
        hid = (b0 + sh @ W0[:, 32:32 + V].T) + x @ W0x.T
        hb = rnd(torch.relu(hid))
        W1b = rnd(W1)
        rgb = torch.sigmoid(b1 + hb @ W1b.T)
        ctx.save_for_backward(x, W0x, hid > 0, hb, W1b, sh, rgb)
        ctx.rounding = rounding
        return rgb

    @staticmethod
    def backward(ctx, g):
        x, W0x, mask, hb, W1b, sh, rgb = ctx.saved_tensors
        rnd = _bf16 if ctx.rounding else (lambda t: t)
        shp = x.shape[:-1]
        dout = (g * (1.0 - rgb) * rgb).reshape(-1, 3)
        doutb = rnd(dout)
        dh = rnd((doutb @ W1b) * mask.reshape(-1, 64))
        dx = dh @ W0x
        V = sh.shape[-1]
        xf, shf = x.reshape(-1, 35), rnd(sh.reshape(-1, V))
        dW0 = torch.zeros(64, 35 + V, dtype=g.dtype)
        dW0[:, :32] = dh.T @ xf[:, :32]
        dW0[:, 32:32 + V] = dh.T @ shf
        dW0[:, 32 + V:35 + V] = dh.T @ xf[:, 32:35]
        return (dx[:, :32].reshape(*shp, 32), None, dx[:, 32:35].reshape(*shp, 3), dW0, dh.sum(0), doutb.T @ hb.reshape(-1, 64),
                dout.sum(0), None)


def radiance_mlp_bf16(hash_feats, view_embed, normals, W0, b0, W1, b1, rounding=True):
    """The BF16 form of ``radiance_mlp`` (see ``_RadianceMlpBF16``).  ``rounding=False``
    keeps every operand at the input precision: the same contractions as the BF16 form
    with the bf16 rounding points removed, which must equal ``radiance_mlp``."""
    return _RadianceMlpBF16.apply(hash_feats, view_embed, normals, W0, b0, W1, b1, rounding)


def variance_s(v, s_max=300.0):
    """variance.py:38-41: s = clamp(exp(10 v), 1e-4, s_max)."""
    return torch.clamp(torch.exp(v * 10.0), min=1e-4, max=float(s_max))


# ----------------------------------------------------------------------------------------
# Sampler (Code/voxneus/rendering/sampler.py) and occupancy (occupancy_grid.py)
# ----------------------------------------------------------------------------------------

def is_occupied(mask, xyz_min, xyz_max, pts):
    """occupancy_grid.py:121-147."""
    res = mask.shape[0]
    shape = pts.shape[:-1]
    fp = pts.reshape(-1, 3)
    norm = (fp - xyz_min) / (xyz_max - xyz_min + 1e-8)
    inb = ((norm >= 0.0) & (norm <= 1.0)).all(dim=-1)
    c = torch.clamp(norm * float(res), 0.0, float(res - 1.0001)).to(torch.int64)
    occ = mask[c[:, 0], c[:, 1], c[:, 2]] & inb
    return occ.reshape(shape)


def logistic_density(sdf, s):
    """occupancy_grid.py:15-22: phi_s(d) = s * sigmoid(s d) * (1 - sigmoid(s d))."""
    sig = torch.sigmoid(s * sdf)
    return s * sig * (1.0 - sig)


def occupancy_update(density, mask, xyz_min, xyz_max, grid5, box: "GridBox", s, idx_x, idx_y, idx_z,
                     first_update, decay=0.95, threshold=0.01):
    """occupancy_grid.py:62-119 with the cells given (the reference draws them with
    torch.randint): density and mask ([res]^3, float32 / bool) are updated in place."""
    res = density.shape[0]
    step = (xyz_max - xyz_min) / float(res)
    nx = (idx_x.float() + 0.5) / float(res)
    ny = (idx_y.float() + 0.5) / float(res)
    nz = (idx_z.float() + 0.5) / float(res)
    pts = torch.stack([xyz_min[0] + nx * (xyz_max[0] - xyz_min[0]), xyz_min[1] + ny * (xyz_max[1] - xyz_min[1]),
                       xyz_min[2] + nz * (xyz_max[2] - xyz_min[2])], dim=-1)
    sdf = query_sdf(grid5, box, pts).squeeze(-1)
    dens = logistic_density(sdf, s)
    voxel_diag = torch.linalg.norm(step).item()
    near = (torch.abs(sdf) < 2.0 * voxel_diag).float()
    eff = torch.maximum(dens, near * (threshold * 2.0))
    prev = density[idx_x, idx_y, idx_z]
    upd = eff if first_update else torch.maximum(prev * decay, eff)
    density[idx_x, idx_y, idx_z] = upd
    mask[idx_x, idx_y, idx_z] = upd > threshold


def sample_pdf(bins, weights, n_samples, det=False, eps=1e-5, u=None, record=None):
    """sampler.py:11-66.  ``u`` injects the uniform draws of the stochastic branch."""
    weights = weights + eps
    pdf = weights / torch.sum(weights, dim=-1, keepdim=True)
    cdf = torch.cumsum(pdf, dim=-1)
    cdf = torch.cat([torch.zeros_like(cdf[..., :1]), cdf], dim=-1)
    bs = list(cdf.shape[:-1])
    if u is None:
        if det:
            u = torch.linspace(0.5 / n_samples, 1.0 - 0.5 / n_samples, steps=n_samples).to(cdf.dtype)
            u = u.expand(bs + [n_samples])
        else:
            u = torch.rand(bs + [n_samples]).to(cdf.dtype)
    u = u.contiguous()
    inds = torch.searchsorted(cdf, u, right=True)
    below = torch.clamp(inds - 1, min=0)
    above = torch.clamp(inds, max=cdf.shape[-1] - 1)
    ig = torch.stack([below, above], -1)
    ms = [ig.shape[0], ig.shape[1], cdf.shape[-1]]
    cdf_g = torch.gather(cdf.unsqueeze(1).expand(ms), 2, ig)
    bins_g = torch.gather(bins.unsqueeze(1).expand(ms), 2, ig)
    denom = cdf_g[..., 1] - cdf_g[..., 0]
    denom = torch.where(denom < eps, torch.ones_like(denom), denom)
    t = (u - cdf_g[..., 0]) / denom
    if record is not None:
        record.append({"inds": inds, "cdf": cdf, "u": u})
    return bins_g[..., 0] + t * (bins_g[..., 1] - bins_g[..., 0])


def up_sample(rays_o, rays_d, z_vals, sdf, n_importance, inv_s, det, occ=None, u=None, record=None):
    """sampler.py:243-292.  ``occ`` is (mask, xyz_min, xyz_max) or None."""
    B, n = z_vals.shape
    pts = rays_o[:, None, :] + rays_d[:, None, :] * z_vals[..., :, None]
    radius = torch.linalg.norm(pts, ord=2, dim=-1, keepdim=False)
    inside = ((radius[:, :-1] < 1.0) | (radius[:, 1:] < 1.0)).to(z_vals.dtype)
    sdf = sdf.reshape(B, n)
    ps, ns = sdf[:, :-1], sdf[:, 1:]
    pz, nz = z_vals[:, :-1], z_vals[:, 1:]
    mid = (ps + ns) * 0.5
    cos = (ns - ps) / (nz - pz + 1e-5)
    pc = torch.cat([torch.zeros([B, 1], dtype=z_vals.dtype), cos[:, :-1]], -1)
    cos, _ = torch.min(torch.stack([pc, cos], -1), -1, keepdim=False)
    cos = cos.clip(-1e3, 0.0) * inside
    dist = nz - pz
    pe = mid - cos * dist * 0.5
    ne = mid + cos * dist * 0.5
    pcdf = torch.sigmoid(pe * inv_s)
    ncdf = torch.sigmoid(ne * inv_s)
    alpha = (pcdf - ncdf + 1e-5) / (pcdf + 1e-5)
    w = alpha * torch.cumprod(torch.cat([torch.ones([B, 1], dtype=z_vals.dtype), 1.0 - alpha + 1e-7], -1), -1)[:, :-1]
    if occ is not None:
        mask, mn, mx = occ
        mz = (pz + nz) * 0.5
        mp = rays_o[:, None, :] + rays_d[:, None, :] * mz[..., :, None]
        om = is_occupied(mask, mn, mx, mp)
        w = w * (om.to(z_vals.dtype) * 0.90 + 0.10)
    if record is not None:
        record.append({"weights": w})
    return sample_pdf(z_vals, w, n_importance, det=det, u=u, record=record).detach()


def cat_z_vals(rays_o, rays_d, z_vals, new_z, sdf, sdf_fn, last):
    """sampler.py:295-320."""
    B, n = z_vals.shape
    _, m = new_z.shape
    z = torch.cat([z_vals, new_z], -1)
    z, index = torch.sort(z, dim=-1)
    if not last:
        pts = rays_o[:, None, :] + rays_d[:, None, :] * new_z[..., :, None]
        nsdf = sdf_fn(pts.reshape(-1, 3)).reshape(B, m)
        sdf = torch.cat([sdf, nsdf], -1)
        xx = torch.arange(B)[:, None].expand(B, n + m).reshape(-1)
        sdf = sdf[(xx, index.reshape(-1))].reshape(B, n + m)
    return z, sdf


def sample_stratified(near, far, n_samples, perturb, t_rand=None):
    """sampler.py:85-125.  ``t_rand`` injects the jitter draws."""
    t = torch.linspace(0.0, 1.0, steps=n_samples).to(near.dtype)
    z = near * (1.0 - t) + far * t
    if perturb:
        mids = 0.5 * (z[..., 1:] + z[..., :-1])
        upper = torch.cat([mids, z[..., -1:]], -1)
        lower = torch.cat([z[..., :1], mids], -1)
        if t_rand is None:
            t_rand = torch.rand_like(z)
        z = lower + (upper - lower) * t_rand
    dists = z[..., 1:] - z[..., :-1]
    dists = torch.cat([dists, dists[..., -1:]], -1)
    return z, dists


def sample_multistep_sdf(rays_o, rays_d, near, far, sdf_fn, n_samples=64, n_importance=64,
                         up_sample_steps=4, perturb=False, occ=None, t_rand=None, u_steps=None,
                         record=None):
    """sampler.py:127-204."""
    z, _ = sample_stratified(near, far, n_samples, perturb, t_rand)
    B = z.shape[0]
    if n_importance > 0 and up_sample_steps > 0:
        with torch.no_grad():
            pts = rays_o[:, None, :] + rays_d[:, None, :] * z[..., :, None]
            sdf = sdf_fn(pts.reshape(-1, 3)).reshape(B, n_samples)
            per = max(n_importance // up_sample_steps, 1)
            for i in range(up_sample_steps):
                inv_s = 64.0 * (2.0 ** i)
                u = None if u_steps is None else u_steps[i]
                nz = up_sample(rays_o, rays_d, z, sdf, per, inv_s, det=not perturb, occ=occ, u=u, record=record)
                z, sdf = cat_z_vals(rays_o, rays_d, z, nz, sdf, sdf_fn, last=(i + 1 == up_sample_steps))
    dists = z[..., 1:] - z[..., :-1]
    dists = torch.cat([dists, dists[..., -1:]], -1)
    return z, dists, z + dists * 0.5


def near_far_from_sphere(rays_o, rays_d, radius=1.0, default_near=0.05):
    """rays.py:102-128."""
    a = torch.sum(rays_d ** 2, dim=-1, keepdim=True)
    b = 2.0 * torch.sum(rays_o * rays_d, dim=-1, keepdim=True)
    mid = 0.5 * (-b) / a
    near = torch.clamp(mid - radius, min=default_near)
    far = torch.clamp(mid + radius, min=near + 0.1)
    return near, far


# ----------------------------------------------------------------------------------------
# Alpha and compositing (Code/voxneus/rendering/{neus_alpha,volume_render}.py)
# ----------------------------------------------------------------------------------------

def compute_neus_alpha(sdf, normals, ray_dirs, dists, s, cos_anneal_ratio=0.0, use_iter_cos=False, eps=1e-5):
    """neus_alpha.py:12-81."""
    if sdf.ndim < normals.ndim:
        sdf = sdf.unsqueeze(-1)
    if dists.ndim < normals.ndim:
        dists = dists.unsqueeze(-1)
    true_cos = torch.sum(normals * ray_dirs, dim=-1, keepdim=True)
    if use_iter_cos:
        cos_val = -(F.relu(-true_cos * 0.5 + 0.5) * (1.0 - cos_anneal_ratio) + F.relu(-true_cos) * cos_anneal_ratio)
    else:
        cos_val = true_cos
    off = 0.5 * dists * cos_val
    pp = torch.sigmoid(s * (sdf - off))
    pn = torch.sigmoid(s * (sdf + off))
    alpha = torch.clamp((pp - pn) / torch.clamp(pp, min=eps), min=0.0, max=1.0)
    if alpha.shape[-1] == 1 and sdf.ndim == normals.ndim:
        alpha = alpha.squeeze(-1)
    return alpha


def composite_weights(alpha, eps=1e-7):
    """volume_render.py:12-40 (cumprod transmittance, T<1e-4 cutoff)."""
    B = alpha.shape[0]
    oma = torch.clamp(1.0 - alpha + eps, min=0.0, max=1.0)
    cp = torch.cumprod(torch.cat([torch.ones([B, 1], dtype=alpha.dtype), oma], -1), -1)
    T = cp[:, :-1]
    w = alpha * T
    w = torch.where(T < 1e-4, torch.zeros_like(w), w)
    return w, T


def huber(pred, target, delta=0.1):
    """volume_render.py:43-54."""
    diff = torch.abs(pred - target)
    return torch.mean(torch.where(diff < delta, 0.5 * (diff ** 2) / delta, diff - 0.5 * delta))


def color_loss(pred_rgb, target_rgb, loss_type="Huber", use_mask=False, target_mask=None, pred_opacity=None,
               mask_weight=0.1):
    """trainer.py:213-245: (loss_total, loss_rgb, loss_mask or None).  Huber (delta 0.1) or
    MSE; with use_mask and a mask, the foreground-weighted colour loss (plain mean when the
    mask sums to zero) plus mask_weight * BCE(clamp(opacity, 1e-6, 1 - 1e-6), mask)."""
    if use_mask and target_mask is not None:
        if loss_type.lower() == "huber":
            diff = torch.abs(pred_rgb - target_rgb)
            is_small = diff < 0.1
            per_ray_loss = torch.where(is_small, 0.5 * (diff ** 2) / 0.1, diff - 0.05)
        else:
            per_ray_loss = (pred_rgb - target_rgb) ** 2
        fg_weight = target_mask.expand_as(pred_rgb)
        fg_sum = fg_weight.sum()
        if fg_sum > 0:
            loss_rgb = (per_ray_loss * fg_weight).sum() / fg_sum
        else:
            loss_rgb = per_ray_loss.mean()
    else:
        if loss_type.lower() == "huber":
            loss_rgb = huber(pred_rgb, target_rgb, delta=0.1)
        else:
            loss_rgb = F.mse_loss(pred_rgb, target_rgb)
    loss_total = loss_rgb
    loss_mask = None
    if use_mask and target_mask is not None:
        loss_mask = F.binary_cross_entropy(torch.clamp(pred_opacity, 1e-6, 1.0 - 1e-6), target_mask)
        loss_total = loss_total + mask_weight * loss_mask
    return loss_total, loss_rgb, loss_mask


def volume_render_from_samples(colors, sdf, normals, ray_dirs, z_vals, dists, s, pts=None, sphere_radius=1.0,
                               background_rgb=None, cos_anneal_ratio=0.0, use_iter_cos=False):
    """volume_render.py:57-153."""
    rde = ray_dirs.unsqueeze(1).expand(-1, colors.shape[1], -1) if ray_dirs.ndim == 2 else ray_dirs
    alpha = compute_neus_alpha(sdf, normals, rde, dists, s, cos_anneal_ratio, use_iter_cos)
    if pts is not None and sphere_radius is not None and sphere_radius > 0:
        alpha = alpha * (torch.linalg.norm(pts, ord=2, dim=-1) < sphere_radius).to(alpha.dtype).detach()
    w, T = composite_weights(alpha)
    rgb = torch.sum(w.unsqueeze(-1) * colors, dim=1)
    depth = torch.sum(w * z_vals, dim=-1, keepdim=True)
    nrm = torch.sum(w.unsqueeze(-1) * normals, dim=1)
    nrm = nrm / (torch.linalg.norm(nrm, dim=-1, keepdim=True) + 1e-8)
    opacity = torch.sum(w, dim=-1, keepdim=True)
    if background_rgb is not None:
        bg = background_rgb if torch.is_tensor(background_rgb) else torch.full_like(rgb, float(background_rgb))
        if bg.ndim == 1:
            bg = bg.unsqueeze(0).expand_as(rgb)
        rgb = rgb + (1.0 - opacity) * bg
    return {"rgb": rgb, "depth": depth, "normal": nrm, "opacity": opacity, "weights": w, "alpha": alpha,
            "transmittance": T}


# ----------------------------------------------------------------------------------------
# Regularisers (Code/voxneus/regularization/*.py) and s-amplification (training/scheduler.py)
# ----------------------------------------------------------------------------------------

def extract_active_vertices(corner_indices, grid_res):
    """active_vertices.py:12-50: sorted unique linear indices ix*Y*Z+iy*Z+iz."""
    rx, ry, rz = int(grid_res[0]), int(grid_res[1]), int(grid_res[2])
    fc = corner_indices.reshape(-1, 3).to(torch.int64)
    fc[:, 0] = torch.clamp(fc[:, 0], 0, rx - 1)
    fc[:, 1] = torch.clamp(fc[:, 1], 0, ry - 1)
    fc[:, 2] = torch.clamp(fc[:, 2], 0, rz - 1)
    lin = torch.unique(fc[:, 0] * (ry * rz) + fc[:, 1] * rz + fc[:, 2])
    iz = lin % rz
    rem = lin // rz
    return torch.stack([rem // ry, rem % ry, iz], -1), lin


def _interior(active, rx, ry, rz):
    ix, iy, iz = active[:, 0], active[:, 1], active[:, 2]
    return (ix > 0) & (ix < rx - 1) & (iy > 0) & (iy < ry - 1) & (iz > 0) & (iz < rz - 1)


def eikonal_loss_and_grad(data, active, voxel_size, eps=1e-8):
    """eikonal.py:13-171 (pure-PyTorch branch).  Loss = mean over interior vertices;
    gradient scale 2/|V_R| counts every active vertex (:44,:152)."""
    rx, ry, rz = data.shape
    g = torch.zeros_like(data)
    nact = active.shape[0]
    if nact == 0:
        return torch.tensor(0.0, dtype=data.dtype), g
    m = _interior(active, rx, ry, rz)
    if not m.any():
        return torch.tensor(0.0, dtype=data.dtype), g
    ix, iy, iz = active[m, 0], active[m, 1], active[m, 2]
    ex, ey, ez = voxel_size[0], voxel_size[1], voxel_size[2]
    dx, dy, dz = 2.0 * ex, 2.0 * ey, 2.0 * ez
    nx = (data[ix + 1, iy, iz] - data[ix - 1, iy, iz]) / dx
    ny = (data[ix, iy + 1, iz] - data[ix, iy - 1, iz]) / dy
    nz = (data[ix, iy, iz + 1] - data[ix, iy, iz - 1]) / dz
    n = torch.stack([nx, ny, nz], -1)
    nn_ = torch.linalg.norm(n, dim=-1, keepdim=True)
    loss = torch.mean((nn_ - 1.0) ** 2)
    v = (2.0 / nact) * (1.0 - 1.0 / (nn_ + eps)) * n
    g.index_put_((ix + 1, iy, iz), v[:, 0] / dx, accumulate=True)
    g.index_put_((ix - 1, iy, iz), -v[:, 0] / dx, accumulate=True)
    g.index_put_((ix, iy + 1, iz), v[:, 1] / dy, accumulate=True)
    g.index_put_((ix, iy - 1, iz), -v[:, 1] / dy, accumulate=True)
    g.index_put_((ix, iy, iz + 1), v[:, 2] / dz, accumulate=True)
    g.index_put_((ix, iy, iz - 1), -v[:, 2] / dz, accumulate=True)
    return loss, g


def curvature_loss_and_grad(data, active, voxel_size):
    """curvature.py:13-165 (pure-PyTorch branch)."""
    rx, ry, rz = data.shape
    g = torch.zeros_like(data)
    nact = active.shape[0]
    if nact == 0:
        return torch.tensor(0.0, dtype=data.dtype), g
    m = _interior(active, rx, ry, rz)
    if not m.any():
        return torch.tensor(0.0, dtype=data.dtype), g
    ix, iy, iz = active[m, 0], active[m, 1], active[m, 2]
    ex, ey, ez = voxel_size[0], voxel_size[1], voxel_size[2]
    ex2, ey2, ez2 = ex * ex, ey * ey, ez * ez
    fc = data[ix, iy, iz]
    Lx = (data[ix + 1, iy, iz] + data[ix - 1, iy, iz] - 2.0 * fc) / ex2
    Ly = (data[ix, iy + 1, iz] + data[ix, iy - 1, iz] - 2.0 * fc) / ey2
    Lz = (data[ix, iy, iz + 1] + data[ix, iy, iz - 1] - 2.0 * fc) / ez2
    loss = torch.mean(Lx ** 2 + Ly ** 2 + Lz ** 2)
    sc = 2.0 / nact
    g.index_put_((ix, iy, iz), sc * (-2.0 * Lx / ex2 - 2.0 * Ly / ey2 - 2.0 * Lz / ez2), accumulate=True)
    g.index_put_((ix + 1, iy, iz), sc * (Lx / ex2), accumulate=True)
    g.index_put_((ix - 1, iy, iz), sc * (Lx / ex2), accumulate=True)
    g.index_put_((ix, iy + 1, iz), sc * (Ly / ey2), accumulate=True)
    g.index_put_((ix, iy - 1, iz), sc * (Ly / ey2), accumulate=True)
    g.index_put_((ix, iy, iz + 1), sc * (Lz / ez2), accumulate=True)
    g.index_put_((ix, iy, iz - 1), sc * (Lz / ez2), accumulate=True)
    return loss, g


def amplify_s_gradient(grad_v, k=5.0):
    """scheduler.py:14-33: grad <- k*grad if grad < 0."""
    if float(grad_v) < 0.0:
        return grad_v * k, True
    return grad_v, False


def eik_clamp_add(grad, g_eik, w_eik, clamp, cooldown=True):
    """trainer.py:309-316: grid.grad + w_eik * g_eik, the weighted eikonal gradient scaled
    to max-abs ``clamp`` while the post-upsample cooldown runs.  Float32 tensor ops in the
    reference's order; ``clamp / norm`` is a Python float over a 0-d tensor, which PyTorch
    evaluates as ``norm.reciprocal() * clamp``."""
    g = g_eik * w_eik
    if cooldown:
        norm = g.abs().max()
        if norm > clamp:
            g = g * (norm.reciprocal() * clamp)
    return grad + g


# ----------------------------------------------------------------------------------------
# Optimizers (Code/voxneus/training/optimizer.py, cuda/adam_upd_kernel.cu, torch.optim.Adam)
# Every elementwise step is its own PyTorch op, so each rounds to float32 once.  Square
# roots are correctly rounded (IEEE sqrt, as sqrtf on the GPU): PyTorch's vectorised CPU
# float32 sqrt is not, so it is taken in float64 and rounded once to float32, which for
# a square root gives the correctly rounded float32 result.
# ----------------------------------------------------------------------------------------

def _sqrt32(x):
    return torch.sqrt(x.double()).to(torch.float32)


def adam_native_step(p, g, m, v, step, beta1, beta2, lr, eps, masked):
    """adam_upd_kernel.cu:6-38 per element, with the host step size of :50 / :70 formed in
    float: step_size = lr * sqrtf(1 - powf(b2, t)) / (1 - powf(b1, t)).  Updates p, m, v
    (float32 tensors) in place; ``masked`` leaves elements with g == 0 untouched."""
    import ctypes
    import ctypes.util

    import numpy as np
    f = np.float32
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    libm.powf.restype, libm.powf.argtypes = ctypes.c_float, [ctypes.c_float, ctypes.c_float]
    powf = lambda a, b: f(libm.powf(float(a), float(b)))  # the C library's powf, as the host code calls
    b1, b2, lr32, eps32 = f(beta1), f(beta2), f(lr), f(eps)
    one = f(1.0)
    step_size = lr32 * np.sqrt(one - powf(b2, f(step))) / (one - powf(b1, f(step)))
    sel = (g != 0.0) if masked else torch.ones_like(g, dtype=torch.bool)
    omb1, omb2 = one - b1, one - b2
    mn = torch.mul(m, float(b1)) + torch.mul(g, float(omb1))
    vn = torch.mul(v, float(b2)) + torch.mul(torch.mul(g, float(omb2)), g)
    upd = torch.div(torch.mul(mn, float(step_size)), _sqrt32(vn) + float(eps32))
    pn = p - upd
    m.copy_(torch.where(sel, mn, m))
    v.copy_(torch.where(sel, vn, v))
    p.copy_(torch.where(sel, pn, p))
    return int(sel.sum())


def sparse_adam_masked_step(p, g, m, v, step, lr, beta1, beta2, eps, weight_decay):
    """optimizer.py:92-113 (the masked PyTorch formula, the path taken with weight decay):
    only elements with g != 0 change; eps is added after dividing by sqrt(1 - b2^t)."""
    mask = g != 0.0
    n = int(mask.sum())
    if n > 0:
        ga = g[mask]
        if weight_decay != 0.0:
            ga = ga + weight_decay * p[mask]
        ma = m[mask] * beta1 + (1.0 - beta1) * ga
        m[mask] = ma
        va = v[mask] * beta2 + (1.0 - beta2) * (ga * ga)
        v[mask] = va
        step_size = lr / (1.0 - beta1 ** step)
        denom = (_sqrt32(va) / math.sqrt(1.0 - beta2 ** step)) + eps
        p[mask] -= step_size * (ma / denom)
    return n


def torch_adam_step(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """torch.optim.Adam (no weight decay, no amsgrad) in the order of its multi-tensor
    update, the one used for device tensors (torch/optim/adam.py, _multi_tensor_adam):
    m = lerp(m, g, 1 - b1); v = v * b2 + (1 - b2) * (g * g);
    p = p + (-lr / bc1) * (m / (sqrt(v) / sqrt(bc2) + eps)), scalars from Python floats.
    PyTorch evaluates lerp with weight < 0.5 as one fused multiply-add, m + w * (g - m)
    rounded once; it is emulated here in float64 (w * (g - m) is exact there, so only a
    float64 sum within 2^-29 ulp of a float32 rounding midpoint could round twice)."""
    w = torch.tensor(1.0 - beta1, dtype=torch.float32).double()
    m.copy_((m.double() + w * (g - m).double()).to(torch.float32))
    v.copy_(torch.mul(v, beta2))
    v.copy_(v + torch.mul(torch.mul(g, g), 1.0 - beta2))
    bc1 = 1.0 - beta1 ** step
    bc2 = 1.0 - beta2 ** step
    denom = torch.div(_sqrt32(v), bc2 ** 0.5) + eps
    p.copy_(p + torch.mul(torch.div(m, denom), (lr / bc1) * -1))


# ----------------------------------------------------------------------------------------
# Assembled model: forward_points, render, train step (voxneus.py, trainer.py)
# ----------------------------------------------------------------------------------------

class RefModel:
    """The reference's VoxNeuS parameters as plain tensors (``voxneus.py:26-92``)."""

    def __init__(self, grid, embeddings, W0, b0, W1, b1, variance=0.3, s_max=300.0,
                 xyz_min=(-1.0, -1.0, -1.0), xyz_max=(1.0, 1.0, 1.0), hash_base_resolution=16,
                 hash_max_resolution=2048.0, occupancy_mask=None, occupancy_res=128, dtype=torch.float32,
                 mlp_form="fp32", gradient_mode="interpolated", view_embed_type="sh"):
        t = lambda a: torch.as_tensor(a).to(dtype).clone()
        self.dtype = dtype
        self.mlp_form = mlp_form
        self.gradient_mode = gradient_mode
        self.view_embed_type = view_embed_type
        self.grid = t(grid).reshape(1, 1, *grid.shape[-3:]).requires_grad_(True)
        self.embeddings = t(embeddings).requires_grad_(True)
        self.W0, self.b0 = t(W0).requires_grad_(True), t(b0).requires_grad_(True)
        self.W1, self.b1 = t(W1).requires_grad_(True), t(b1).requires_grad_(True)
        self.variance = torch.tensor(float(variance), dtype=torch.float32).to(dtype).requires_grad_(True)
        self.s_max = s_max
        self.box = GridBox(tuple(self.grid.shape[-3:]), xyz_min, xyz_max, dtype=dtype)
        self.level_res = hash_level_resolutions(self.embeddings.shape[0], hash_base_resolution, hash_max_resolution)
        if occupancy_mask is None:
            occupancy_mask = torch.ones((occupancy_res,) * 3, dtype=torch.bool)
        self.occ_mask = torch.as_tensor(occupancy_mask, dtype=torch.bool)
        self.occ_min = torch.tensor(xyz_min, dtype=dtype)
        self.occ_max = torch.tensor(xyz_max, dtype=dtype)

    def params(self):
        return [self.grid, self.embeddings, self.W0, self.b0, self.W1, self.b1, self.variance]

    @property
    def s(self):
        return variance_s(self.variance, self.s_max)

    def query_sdf(self, xyz):
        return query_sdf(self.grid, self.box, xyz)

    def forward_points(self, xyz, view_dirs, active_levels=None):
        """voxneus.py:115-167."""
        sdf = self.query_sdf(xyz)
        grad, normal, corners, weights = query_normals(self.grid, self.box, xyz, self.gradient_mode)
        norm_xyz = (xyz - self.box.xyz_min) / (self.box.xyz_max - self.box.xyz_min + 1e-8)
        feats = hash_encode(torch.clamp(norm_xyz, min=0.0, max=1.0), self.embeddings, self.level_res, active_levels)
        mlp = radiance_mlp_bf16 if self.mlp_form == "bf16" else radiance_mlp
        rgb = mlp(feats, view_embed(view_dirs, self.view_embed_type), normal, self.W0, self.b0, self.W1, self.b1)
        return {"sdf": sdf, "grad": grad, "normal": normal, "corner_indices": corners, "weights": weights,
                "rgb": rgb, "hash_feats": feats}

    def sample(self, rays_o, rays_d, near, far, n_samples=64, n_importance=64, up_sample_steps=4,
               perturb=False, t_rand=None, u_steps=None, record=None, use_occupancy=True):
        occ = (self.occ_mask, self.occ_min, self.occ_max) if use_occupancy else None
        fn = lambda p: self.query_sdf(p).squeeze(-1)
        with torch.no_grad():
            return sample_multistep_sdf(rays_o, rays_d, near, far, fn, n_samples, n_importance, up_sample_steps,
                                        perturb, occ, t_rand, u_steps, record)

    def render(self, rays_o, rays_d, near, far, n_samples=64, n_importance=64, up_sample_steps=4,
               perturb=False, t_rand=None, u_steps=None, record=None, sphere_radius=1.0,
               background_rgb=None, cos_anneal_ratio=0.0, use_iter_cos=False, active_levels=None, zs=None):
        """voxneus.py:169-260.  ``zs`` = (z_vals, dists, mid_z) skips the sampler."""
        if zs is None:
            zs = self.sample(rays_o, rays_d, near, far, n_samples, n_importance, up_sample_steps, perturb,
                             t_rand, u_steps, record)
        z_vals, dists, mid = zs
        pts = rays_o.unsqueeze(1) + rays_d.unsqueeze(1) * mid.unsqueeze(-1)
        dirs = rays_d.unsqueeze(1).expand_as(pts)
        po = self.forward_points(pts, dirs, active_levels)
        out = volume_render_from_samples(po["rgb"], po["sdf"], po["normal"], dirs, mid, dists, self.s, pts=pts,
                                         sphere_radius=sphere_radius, background_rgb=background_rgb,
                                         cos_anneal_ratio=cos_anneal_ratio, use_iter_cos=use_iter_cos)
        out.update({"pts": pts, "sdf_samples": po["sdf"], "corner_indices": po["corner_indices"],
                    "grad_samples": po["grad"], "normal_samples": po["normal"], "z_vals": z_vals,
                    "dists": dists, "mid_z_vals": mid, "sample_rgb": po["rgb"]})
        return out

    def train_step(self, rays_o, rays_d, target_rgb, w_eik=0.1, w_curv=5e-4, s_amp_k=5.0, amplify=True,
                   huber_delta=0.1, zs=None, **render_kw):
        """trainer.py:188-326 in order: bounds, render, Huber loss, backward, V_R with the
        relaxed-sphere filter, eikonal + curvature closed-form gradients, s-amplification.
        Uses the trainer's separate-tensor branch (:292-319) with the regulariser weights
        applied on accumulation, which is the branch taken on the CPU."""
        for p in self.params():
            p.grad = None
        near, far = near_far_from_sphere(rays_o, rays_d, radius=1.0)
        out = self.render(rays_o, rays_d, near, far, zs=zs, **render_kw)
        loss = huber(out["rgb"], target_rgb, huber_delta)
        loss.backward()
        res = self.box.res
        verts, lin = extract_active_vertices(out["corner_indices"], res)
        world = self.box.xyz_min + verts.to(self.dtype) * self.box.voxel_size
        inside = torch.linalg.norm(world, dim=-1) < 1.2
        if inside.sum() > 0:
            verts, lin = verts[inside], lin[inside]
        data = self.grid.detach()[0, 0]
        vs = self.box.voxel_size
        le, ge = eikonal_loss_and_grad(data, verts, vs)
        lc, gc = curvature_loss_and_grad(data, verts, vs)
        with torch.no_grad():
            if w_eik > 0:
                self.grid.grad.add_((ge * w_eik).reshape(self.grid.shape))
            if w_curv > 0:
                self.grid.grad.add_(gc.reshape(self.grid.shape), alpha=w_curv)
        amplified = False
        if amplify and self.variance.grad is not None:
            g, amplified = amplify_s_gradient(self.variance.grad, s_amp_k)
            self.variance.grad = g
        return {"loss": loss.detach(), "loss_eik": le, "loss_curv": lc, "active_linear": lin,
                "active_vertices": verts, "amplified": amplified, "render": out, "near": near, "far": far}


# ----------------------------------------------------------------------------------------
# Mesh extraction (Code/voxneus/geometry/mesh.py:34-77)
# The reference calls scikit-image's marching_cubes, which is not installed here; this is
# marching cubes with one vertex per straddling grid edge at the linear-interpolation
# crossing (the vertex set scikit-image's Lewiner variant also produces) and triangles
# from a face-consistent case table.  Pinned by the mesh's properties (closed, oriented,
# Euler characteristic, enclosed volume and area, vertices on the zero crossing), not by
# scikit-image's triangulation.
# ----------------------------------------------------------------------------------------

def _mc_edge(c0, c1):
    d = c0 ^ c1
    a = {4: 0, 2: 1, 1: 2}[d]
    lo = c0 & ~d
    u, v = [(lo >> (2 - b)) & 1 for b in range(3) if b != a]
    return 4 * a + 2 * u + v


def _mc_edge_corners(e):
    a, uv = e // 4, ((e >> 1) & 1, e & 1)
    base = [0, 0, 0]
    for b, x in zip([b for b in range(3) if b != a], uv):
        base[b] = x
    c0 = (base[0] << 2) | (base[1] << 1) | base[2]
    return c0, c0 | (1 << (2 - a))


def _mc_edge_faces(e):
    """The two cube faces (axis, side) containing edge e."""
    c0, c1 = _mc_edge_corners(e)
    a = e // 4
    return {(b, (c0 >> (2 - b)) & 1) for b in range(3) if b != a}


def mc_case_table():
    """For every case (bit k set = corner k = 4ox+2oy+oz inside), the triangles as edge
    triples: on each face, walked counter-clockwise about its outward normal, every run of
    inside corners is cut off by a segment from the crossing leaving the run to the
    crossing entering it; segments chain into loops (edge by edge across faces), each
    loop oriented with its normal from inside to outside corners and fanned from its
    first edge."""
    import numpy as np
    pos = lambda k: np.array([(k >> 2) & 1, (k >> 1) & 1, k & 1], float)
    faces = []
    for a in range(3):
        for s in (0, 1):
            cs = [k for k in range(8) if ((k >> (2 - a)) & 1) == s]
            n = np.zeros(3)
            n[a] = 1.0 if s else -1.0
            t1 = np.zeros(3)
            t1[(a + 1) % 3] = 1.0
            t2 = np.cross(n, t1)
            ctr = sum(pos(k) for k in cs) / 4.0
            cs.sort(key=lambda k: math.atan2(float(np.dot(pos(k) - ctr, t2)), float(np.dot(pos(k) - ctr, t1))))
            faces.append(cs)
    table = []
    for case in range(256):
        inside = [(case >> k) & 1 for k in range(8)]
        succ = {}
        for cs in faces:
            for i in range(4):
                if inside[cs[i]] and not inside[cs[(i + 1) % 4]]:
                    j = i
                    while inside[cs[(j - 1) % 4]]:
                        j = (j - 1) % 4
                    succ[_mc_edge(cs[i], cs[(i + 1) % 4])] = _mc_edge(cs[(j - 1) % 4], cs[j])
        tris, seen = [], set()
        for e0 in range(12):
            if e0 not in succ or e0 in seen:
                continue
            loop, e = [e0], succ[e0]
            seen.add(e0)
            while e != e0:
                loop.append(e)
                seen.add(e)
                e = succ[e]
            mids = [sum(pos(c) for c in _mc_edge_corners(x)) / 2.0 for x in loop]
            nrm = sum(np.cross(mids[i], mids[(i + 1) % len(mids)]) for i in range(len(mids)))
            cor = [c for x in loop for c in _mc_edge_corners(x)]
            ins = np.mean([pos(c) for c in cor if inside[c]], 0)
            outs = np.mean([pos(c) for c in cor if not inside[c]], 0)
            if float(np.dot(nrm, outs - ins)) < 0:
                loop = loop[::-1]
            # fan from the first loop vertex whose fan has no triangle lying in a cube face
            # (two cells sharing the face would both cover it)
            m = len(loop)
            for k in range(m):
                cand = loop[k:] + loop[:k]
                fan = [(cand[0], cand[i], cand[i + 1]) for i in range(1, m - 1)]
                if not any(_mc_edge_faces(a) & _mc_edge_faces(b) & _mc_edge_faces(c) for a, b, c in fan):
                    break
            else:
                raise AssertionError(f"case {case}: no fan avoids the cube faces")
            tris += fan
        table.append(tris)
    return table


def marching_cubes(volume, level=0.0, table=None):
    """Vertices in grid-index space (float64 [V,3]; numbered by owning grid vertex, the
    edge's low end, in grid order, then by axis) and triangles (int32 [F,3]; by cell in
    grid order, then table order).  value < level is inside."""
    import numpy as np
    vol = np.asarray(volume, dtype=np.float32)
    X, Y, Z = vol.shape
    table = mc_case_table() if table is None else table
    inside = vol < np.float32(level)
    vmask = np.zeros(vol.shape, np.int64)
    vmask[:-1, :, :] |= (inside[:-1] != inside[1:]).astype(np.int64)
    vmask[:, :-1, :] |= (inside[:, :-1] != inside[:, 1:]).astype(np.int64) << 1
    vmask[:, :, :-1] |= (inside[:, :, :-1] != inside[:, :, 1:]).astype(np.int64) << 2
    flat = vmask.reshape(-1)
    cnt = ((flat & 1) + ((flat >> 1) & 1) + ((flat >> 2) & 1)).astype(np.int64)
    voff = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    owners, axes = [], []
    for a in range(3):
        idx = np.nonzero((flat >> a) & 1)[0]
        owners.append(idx)
        axes.append(np.full(idx.shape, a))
    owners, axes = np.concatenate(owners), np.concatenate(axes)
    order = np.lexsort((axes, owners))
    owners, axes = owners[order], axes[order]
    p = np.stack(np.unravel_index(owners, vol.shape), -1).astype(np.float64)
    step = np.array([Y * Z, Z, 1])[axes]
    v0 = vol.reshape(-1)[owners].astype(np.float64)
    v1 = vol.reshape(-1)[owners + step].astype(np.float64)
    t = (float(np.float32(level)) - v0) / (v1 - v0)
    p[np.arange(p.shape[0]), axes] += t
    cx, cy, cz = np.meshgrid(np.arange(X - 1), np.arange(Y - 1), np.arange(Z - 1), indexing="ij")
    base = (cx * Y * Z + cy * Z + cz).reshape(-1)
    case = np.zeros(base.shape, np.int64)
    for k in range(8):
        ox, oy, oz = (k >> 2) & 1, (k >> 1) & 1, k & 1
        case |= inside[ox:X - 1 + ox, oy:Y - 1 + oy, oz:Z - 1 + oz].reshape(-1).astype(np.int64) << k
    keys, faces = [], []
    for c in range(256):
        tris = table[c]
        if not tris:
            continue
        cells = base[case == c]
        if cells.size == 0:
            continue
        for ti, tri in enumerate(tris):
            ids = []
            for e in tri:
                a, u, v = e // 4, (e >> 1) & 1, e & 1
                off = [0, 0, 0]
                for b, x in zip([b for b in range(3) if b != a], (u, v)):
                    off[b] = x
                own = cells + off[0] * Y * Z + off[1] * Z + off[2]
                below = flat[own] & ((1 << a) - 1)
                ids.append(voff[own] + (below & 1) + ((below >> 1) & 1))
            faces.append(np.stack(ids, -1))
            keys.append(np.stack([cells, np.full(cells.shape, ti)], -1))
    if faces:
        faces, keys = np.concatenate(faces), np.concatenate(keys)
        faces = faces[np.lexsort((keys[:, 1], keys[:, 0]))].astype(np.int32)
    else:
        faces = np.zeros((0, 3), np.int32)
    return p, faces


def apply_gaussian_filter_to_grid(sdf_grid, sigma=1.0):
    """mesh.py:17-31: scipy.ndimage.gaussian_filter(sdf_grid, sigma, mode="nearest"), the
    library routine the reference calls (scipy is installed here; a float32 grid is
    filtered axis by axis, each 1-D pass accumulated in double and stored as float32)."""
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(sdf_grid, sigma=sigma, mode="nearest")


def gaussian_weights(sigma, truncate=4.0):
    """scipy's 1-D Gaussian correlation weights (scipy/ndimage/_filters.py
    gaussian_filter1d and _gaussian_kernel1d, order 0): radius int(truncate*sigma + 0.5),
    exp(-0.5/sigma^2 * x^2) over x = -r..r, divided by its sum, in float64."""
    import numpy as np
    r = int(truncate * float(sigma) + 0.5)
    x = np.arange(-r, r + 1)
    phi = np.exp(-0.5 / (float(sigma) * float(sigma)) * x ** 2)
    return phi / phi.sum(), r


def extract_mesh_marching_cubes(sdf_grid, xyz_min, xyz_max, level=0.0, post_filter_sigma=None):
    """mesh.py:34-77: world-space float32 vertices, index * (xyz_max - xyz_min) / (res - 1)
    + xyz_min, and int32 faces, after the optional Gaussian pre-filter (mesh.py:56-57)."""
    import numpy as np
    if post_filter_sigma is not None and post_filter_sigma > 0:
        sdf_grid = apply_gaussian_filter_to_grid(np.asarray(sdf_grid).copy(), post_filter_sigma)
    verts, faces = marching_cubes(sdf_grid, level)
    res = np.array(np.asarray(sdf_grid).shape, dtype=np.float64)
    mn = np.asarray(xyz_min, dtype=np.float32)
    mx = np.asarray(xyz_max, dtype=np.float32)
    scale = (mx - mn) / (res - 1.0)
    return (verts * scale + mn).astype(np.float32), faces
