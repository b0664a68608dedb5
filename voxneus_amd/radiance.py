"""Appearance modules with the reference's parameter layout: the SH and positional-
encoding view embeddings (models/embedder.py:12-87), the 2-layer radiance MLP
(models/radiance_mlp.py:11-99) and the single variance parameter (models/variance.py:11-46).

On the fused path these are evaluated inside the gfx950 render kernel; the modules hold
the parameters (``net.0.weight`` [64,51] with the SH embedding or [64,62] with the
positional encoding, ``net.0.bias``, ``net.2.weight`` [3,64], ``net.2.bias``,
``variance``) under the reference's state-dict names.  Called on its own (as
``VoxNeuS.forward_points`` does) the MLP runs the per-point gfx950 kernel.
"""

from __future__ import annotations

from typing import Optional, Union

import torch
import torch.nn as nn

from . import _native as N
from . import functional as Fn

SH_C = (0.28209479177387814, 0.4886025119029199, 1.0925484305920792, 0.31539156525252005,
        0.5462742152960396, 0.5900435899266435, 2.890611442640554, 0.4570457994644658,
        0.3731763325901154, 1.445305721320277)


class SphericalHarmonicsEmbedder(nn.Module):
    """Real SH bands 0-3 (16 outputs) of unit directions (embedder.py:55-87)."""

    def __init__(self, degree: int = 4):
        super().__init__()
        if degree != 4:
            raise NotImplementedError("the fused path implements degree 4 (16 outputs)")
        self.degree = degree
        self.out_dim = 16

    def forward(self, dirs: torch.Tensor) -> torch.Tensor:
        x, y, z = dirs[..., 0:1], dirs[..., 1:2], dirs[..., 2:3]
        C = SH_C
        return torch.cat([torch.full_like(x, C[0]), -C[1] * y, C[1] * z, -C[1] * x, C[2] * x * y, -C[2] * y * z,
                          C[3] * (2.0 * z * z - x * x - y * y), -C[2] * x * z, C[4] * (x * x - y * y),
                          -C[5] * y * (3.0 * x * x - y * y), C[6] * x * y * z,
                          -C[7] * y * (4.0 * z * z - x * x - y * y),
                          C[8] * z * (2.0 * z * z - 3.0 * x * x - 3.0 * y * y),
                          -C[7] * x * (4.0 * z * z - x * x - y * y), C[9] * z * (x * x - y * y),
                          -C[5] * x * (x * x - 3.0 * y * y)], -1)


class PositionalEmbedder(nn.Module):
    """gamma(x) = [x, sin(x), cos(x), sin(2x), cos(2x), ...] (embedder.py:12-44); the fused
    render evaluates it per ray for num_freqs 4 with the input (27 outputs)."""

    def __init__(self, input_dims: int = 3, num_freqs: int = 4, include_input: bool = True,
                 log_sampling: bool = True):
        super().__init__()
        if (input_dims, num_freqs, include_input, log_sampling) != (3, 4, True, True):
            raise NotImplementedError("the fused path implements 4 log-spaced frequencies with the input "
                                      "(27 outputs)")
        self.input_dims, self.num_freqs, self.include_input = input_dims, num_freqs, include_input
        self.register_buffer("freq_bands", 2.0 ** torch.linspace(0.0, num_freqs - 1, num_freqs))
        self.out_dim = input_dims + input_dims * num_freqs * 2

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        outputs = [x]
        for freq in self.freq_bands:
            outputs.append(torch.sin(x * freq))
            outputs.append(torch.cos(x * freq))
        return torch.cat(outputs, dim=-1)


class RadianceMLP(nn.Module):
    """sigmoid(W1 relu(W0 [hash32, view, normal3] + b0) + b1), kaiming-uniform (ReLU) init;
    view = the 16 SH values or the 27 positional-encoding values."""

    def __init__(self, hash_feature_dim: int = 32, view_embed_dim: int = 16, normal_dim: int = 3,
                 geom_feature_dim: int = 0, hidden_dim: int = 64, num_layers: int = 2, weight_norm: bool = False,
                 entangled: bool = False):
        super().__init__()
        if (hash_feature_dim, normal_dim, geom_feature_dim, hidden_dim, num_layers) != (
                32, 3, 0, 64, 2) or view_embed_dim not in (16, 27) or weight_norm or entangled:
            raise NotImplementedError("the fused gfx950 MLP is the reference default 51 -> 64 -> 3 (SH view "
                                      "embedding) or 62 -> 64 -> 3 (positional encoding), with no geometry "
                                      "features, no weight norm, not entangled")
        self.entangled = False
        self.geom_feature_dim = 0
        self.num_layers = 2
        self.weight_norm = False
        self.view_embed_dim = view_embed_dim
        self.in_dim = 35 + view_embed_dim
        self.net = nn.Sequential(nn.Linear(self.in_dim, 64), nn.ReLU(inplace=True), nn.Linear(64, 3), nn.Sigmoid())
        self.mlp_mode = N.VXN_MLP_FP32  # the owning VoxNeuS sets its form
        self.reset_parameters()

    def reset_parameters(self) -> None:
        for m in self.net.modules():
            if isinstance(m, nn.Linear):
                nn.init.kaiming_uniform_(m.weight, nonlinearity="relu")
                nn.init.zeros_(m.bias)

    def flat(self) -> torch.Tensor:
        """W0, b0, W1, b1 flattened (differentiable), the layout the kernels read."""
        return torch.cat([self.net[0].weight.reshape(-1), self.net[0].bias, self.net[2].weight.reshape(-1),
                          self.net[2].bias])

    def unflatten_grad(self, g: torch.Tensor) -> None:
        """Write a flat gradient (the ``flat()`` layout) into the four parameters' .grad
        (accumulating)."""
        sizes = (64 * self.in_dim, 64, 3 * 64, 3)
        params = (self.net[0].weight, self.net[0].bias, self.net[2].weight, self.net[2].bias)
        off = 0
        for p, n in zip(params, sizes):
            v = g[off:off + n].reshape(p.shape)
            if p.grad is None:
                p.grad = v.clone()
            else:
                p.grad.add_(v)
            off += n

    def forward(self, hash_features, view_embeddings, normals, geom_features=None, sdf=None):
        """radiance_mlp.py:81-99 on the device: the per-point MLP kernel (FP32, or the BF16
        form's rounding points when ``mlp_mode`` is VXN_MLP_BF16_MFMA)."""
        return Fn.PointMlpFunction.apply(hash_features, view_embeddings, normals, self.flat(), self.mlp_mode)


class SingleVarianceNetwork(nn.Module):
    """s = clamp(exp(10 v), 1e-4, s_max) (variance.py:11-46)."""

    def __init__(self, init_val: float = 0.3, s_max: float = 300.0):
        super().__init__()
        self.variance = nn.Parameter(torch.tensor(init_val, dtype=torch.float32))
        self._s_max = float(s_max)

    def forward(self, x: Union[int, torch.Tensor] = 1) -> torch.Tensor:
        s = self.s
        if isinstance(x, torch.Tensor):
            return torch.ones_like(x[..., :1]) * s
        if isinstance(x, int) and x > 1:
            return torch.full((x, 1), s.item(), device=self.variance.device)
        return s

    @property
    def s(self) -> torch.Tensor:
        return torch.clamp(torch.exp(self.variance * 10.0), min=1e-4, max=self._s_max)

    @property
    def std(self) -> torch.Tensor:
        return 1.0 / self.s
