"""Checkpoint layout and grid super-resolution on the CPU: the oracle's upsampling is
exact on trilinear fields and keeps the corner values (align_corners), and checkpoints
keep the reference's layout and load without executing anything from the file."""

import os

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R


def test_upsample_reproduces_trilinear_fields_and_corners():
    c = torch.tensor([0.2, -0.7, 0.4, 1.1, 0.3, -0.5, 0.8, 0.6], dtype=torch.float64)
    f = lambda x, y, z: c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z + c[6] * x * z + \
        c[7] * x * y * z
    def lattice(n):
        ax = [torch.linspace(-1, 1, k, dtype=torch.float64) for k in n]
        return torch.meshgrid(*ax, indexing="ij")
    src = f(*lattice((9, 12, 7)))[None, None]
    up = R.upsample_sdf_grid(src, (17, 23, 31))
    assert torch.allclose(up[0, 0], f(*lattice((17, 23, 31))), atol=1e-12)
    g = torch.randn(1, 1, 10, 11, 12)
    u = R.upsample_sdf_grid(g, (19, 21, 23))[0, 0]
    for i in (0, -1):
        for j in (0, -1):
            for k in (0, -1):
                assert u[i, j, k] == g[0, 0, i, j, k]
    assert torch.equal(R.upsample_sdf_grid(g, (10, 11, 12)), g)


def test_checkpoint_layout_and_safe_round_trip(tmp_path):
    from voxneus_amd.checkpoint import load_checkpoint, save_checkpoint
    from voxneus_amd.voxneus import VoxNeuS
    torch.manual_seed(0)
    m = VoxNeuS(grid_res=16, hash_log2_size=10)
    with torch.no_grad():
        for p in m.parameters():
            p.add_(torch.randn_like(p))
    cfg = {"model": {"grid_res": 16, "s_max": 250.0, "gradient_mode": "analytical"}, "training": {"lr": 0.1}}
    path = os.path.join(tmp_path, "ck", "step_5.pt")
    save_checkpoint(path, 5, m, {}, cfg, {"psnr": 30.5, "s_amplified": True})
    ck = torch.load(path, map_location="cpu", weights_only=True)
    assert set(ck) == {"step", "model_state", "config", "metrics", "grid_res"}
    assert ck["step"] == 5 and ck["grid_res"] == [16, 16, 16] and ck["config"] == cfg
    keys = set(ck["model_state"])
    for k in ("sdf_grid.grid", "sdf_grid.xyz_min", "radiance_hash.encoder.embeddings", "radiance_hash.encoder.primes",
              "radiance_mlp.net.0.weight", "radiance_mlp.net.2.bias", "variance_net.variance",
              "occupancy_grid.density_grid", "occupancy_grid.occupancy_mask"):
        assert k in keys
    m2 = VoxNeuS(grid_res=16, hash_log2_size=10)
    out = load_checkpoint(path, m2)
    assert out["step"] == 5
    for (k, a), (_, b) in zip(m.state_dict().items(), m2.state_dict().items()):
        assert torch.equal(a, b), k
    with pytest.raises(FileNotFoundError):
        load_checkpoint(os.path.join(tmp_path, "missing.pt"), m2)
