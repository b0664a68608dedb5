"""The positional-encoding view embedding on the CPU: the oracle's encoding pinned by its
closed form (layout [d, sin d, cos d, sin 2d, cos 2d, ..., cos 8d], the input first, each
band a 3-vector, sin^2 + cos^2 = 1), the package module equal to it, and the model and
MLP shapes of view_embed_type="pe" (voxneus.py:78-87, radiance_mlp.py:36)."""

import math

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R


def test_pe_layout_and_values():
    d = torch.tensor([[0.3, -0.8, 0.52], [1.0, 0.0, -1.0]], dtype=torch.float32)
    pe = R.pe_embed(d)
    assert pe.shape == (2, 27)
    assert torch.equal(pe[:, :3], d)
    for f in range(4):
        b = 2.0 ** f
        for c in range(3):
            for r in range(2):
                x = float(d[r, c]) * b
                assert abs(float(pe[r, 3 + 6 * f + c]) - math.sin(x)) < 2e-7
                assert abs(float(pe[r, 6 + 6 * f + c]) - math.cos(x)) < 2e-7
    s2 = pe[:, 3::6] ** 2 + pe[:, 6::6] ** 2
    assert torch.allclose(s2, torch.ones_like(s2), atol=1e-6)
    with pytest.raises(ValueError):
        R.view_embed(d, "fourier")


def test_package_embedder_equals_oracle():
    from voxneus_amd.radiance import PositionalEmbedder
    d = torch.nn.functional.normalize(torch.randn(100, 3, generator=torch.Generator().manual_seed(1)), dim=-1)
    pe = PositionalEmbedder(num_freqs=4, include_input=True)
    assert pe.out_dim == 27
    assert torch.equal(pe(d), R.pe_embed(d))
    with pytest.raises(NotImplementedError):
        PositionalEmbedder(num_freqs=6)


def test_pe_model_shapes():
    from voxneus_amd import _native as N
    from voxneus_amd.radiance import RadianceMLP
    mlp = RadianceMLP(32, 27, 3)
    assert tuple(mlp.net[0].weight.shape) == (64, 62)
    assert mlp.flat().numel() == N.MLP_PARAMS_PE == 64 * 62 + 64 + 3 * 64 + 3
    g = torch.arange(N.MLP_PARAMS_PE, dtype=torch.float32)
    mlp.unflatten_grad(g)
    assert torch.equal(mlp.net[0].weight.grad.reshape(-1), g[:64 * 62])
    assert torch.equal(mlp.net[2].bias.grad, g[-3:])
    assert N.view_embed_code("PE") == N.VXN_VIEW_PE and N.view_embed_code("sh") == N.VXN_VIEW_SH
    with pytest.raises(ValueError):
        N.view_embed_code("fourier")
    with pytest.raises(NotImplementedError):
        RadianceMLP(32, 20, 3)


def test_pe_oracle_mlp_matches_manual_concat():
    """The oracle's MLP on the PE embedding is F.linear over [hash, pe, normal] (the
    reference's concatenation order, radiance_mlp.py:90)."""
    g = torch.Generator().manual_seed(2)
    feats, nrm = torch.randn(10, 32, generator=g), torch.randn(10, 3, generator=g)
    d = torch.nn.functional.normalize(torch.randn(10, 3, generator=g), dim=-1)
    W0, b0 = torch.randn(64, 62, generator=g), torch.randn(64, generator=g)
    W1, b1 = torch.randn(3, 64, generator=g), torch.randn(3, generator=g)
    rgb = R.radiance_mlp(feats, R.pe_embed(d), nrm, W0, b0, W1, b1)
    x = torch.cat([feats, R.pe_embed(d), nrm], -1).double()
    ref = torch.sigmoid(torch.relu(x @ W0.double().T + b0.double()) @ W1.double().T + b1.double())
    assert torch.allclose(rgb.double(), ref, atol=1e-5)
