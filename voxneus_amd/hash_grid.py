"""Multiresolution hash-grid encoder, drop-in for RadianceHashGrid / PyTorchHashGrid
(models/radiance_hash.py:25-231), backed by the gfx950 hash kernels."""

from __future__ import annotations

import math
from typing import List, Optional

import torch
import torch.nn as nn

from . import functional as Fn


def level_resolutions(n_levels: int = 16, base_resolution: int = 16, max_resolution: float = 2048.0) -> List[int]:
    """radiance_hash.py:49,108: b = exp((ln max - ln base)/(L-1)); N_l = floor(base * b**l) in double."""
    b = math.exp((math.log(max_resolution) - math.log(base_resolution)) / max(n_levels - 1, 1))
    return [math.floor(base_resolution * (b ** lvl)) for lvl in range(n_levels)]


def level_weights(n_levels: int, active_levels: Optional[float]) -> Optional[List[float]]:
    """NeuS2 lambda mask (radiance_hash.py:90-102); None when every level is fully active."""
    if active_levels is None:
        return None
    out = []
    for lvl in range(n_levels):
        li = lvl + 1.0
        if li <= active_levels:
            out.append(1.0)
        elif li - 1.0 < active_levels < li:
            out.append(active_levels - (li - 1.0))
        else:
            out.append(0.0)
    return out


class HashGridEncoder(nn.Module):
    """Parameters and level table of one hash grid; ``embeddings`` is [L, T, F]."""

    def __init__(self, n_levels=16, n_features_per_level=2, log2_hashmap_size=19, base_resolution=16,
                 max_resolution=2048.0, dtype=torch.float32):
        super().__init__()
        self.n_levels = n_levels
        self.n_features_per_level = n_features_per_level
        self.log2_hashmap_size = log2_hashmap_size
        self.hashmap_size = 2 ** log2_hashmap_size
        self.base_resolution = base_resolution
        self.max_resolution = max_resolution
        self.out_dim = n_levels * n_features_per_level
        self.b = math.exp((math.log(max_resolution) - math.log(base_resolution)) / max(n_levels - 1, 1))
        self.register_buffer("primes", torch.tensor([1, 2654435761, 805459861], dtype=torch.int64))
        self.embeddings = nn.Parameter(torch.empty(n_levels, self.hashmap_size, n_features_per_level, dtype=dtype))
        self.level_res = level_resolutions(n_levels, base_resolution, max_resolution)
        self.reset_parameters()

    def reset_parameters(self):
        nn.init.uniform_(self.embeddings, -1e-4, 1e-4)

    def forward(self, x01: torch.Tensor, active_levels: Optional[float] = None) -> torch.Tensor:
        return Fn.HashEncodeFunction.apply(self.embeddings, x01, self.level_res,
                                           level_weights(self.n_levels, active_levels))


# Name of the reference's pure-PyTorch encoder, for drop-in imports.
PyTorchHashGrid = HashGridEncoder


class RadianceHashGrid(nn.Module):
    """radiance_hash.py:169-231: clamps to [0,1] and encodes (the clamp runs in-kernel)."""

    def __init__(self, n_levels=16, n_features_per_level=2, log2_hashmap_size=19, base_resolution=16,
                 max_resolution=2048.0, force_pytorch=False, dtype=torch.float32):
        super().__init__()
        self.n_levels = n_levels
        self.n_features_per_level = n_features_per_level
        self.out_dim = n_levels * n_features_per_level
        self.use_tcnn = False
        self.encoder = HashGridEncoder(n_levels, n_features_per_level, log2_hashmap_size, base_resolution,
                                       max_resolution, dtype)

    def forward(self, norm_xyz: torch.Tensor, active_levels: Optional[float] = None) -> torch.Tensor:
        return self.encoder(norm_xyz, active_levels=active_levels)
