"""Grid super-resolution and the checkpoint -> resident scene path on the GPU: the
upsampling kernel equals the oracle (PyTorch's CPU F.interpolate) bit for bit, loading a
checkpoint of another resolution rescales the grid first, and a scene loaded from a
checkpoint serves renders and queries identical to the single-scene calls on the
model's own tensors."""

import os

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R
from voxneus_amd import scenes as S

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


@pytest.fixture(scope="module")
def Fn():
    from voxneus_amd import functional
    return functional


@pytest.mark.parametrize("src,dst", [((13, 17, 11), (29, 40, 23)), ((96, 96, 96), (160, 160, 160)),
                                     ((40, 40, 40), (17, 17, 17)), ((8, 9, 10), (8, 9, 10)), ((5, 1, 6), (9, 4, 1))])
def test_upsample_matches_oracle(Fn, src, dst):
    g = torch.Generator().manual_seed(sum(src) + sum(dst))
    grid = torch.randn(1, 1, *src, generator=g)
    ref = R.upsample_sdf_grid(grid, dst)[0, 0]
    out = Fn.upsample_grid(grid.to(DEV), dst)
    assert torch.equal(out.cpu(), ref)


def _random_model(grid_res, seed):
    from voxneus_amd.voxneus import VoxNeuS
    torch.manual_seed(seed)
    m = VoxNeuS(grid_res=grid_res, hash_log2_size=14).to(DEV)
    with torch.no_grad():
        m.radiance_hash.encoder.embeddings.uniform_(-0.05, 0.05)
        for p in m.radiance_mlp.parameters():
            p.mul_(3.0)
        m.variance_net.variance.fill_(0.42)
        occ = m.occupancy_grid.occupancy_mask
        occ.copy_(torch.rand(occ.shape, device=DEV) < 0.7)
    return m


def test_load_checkpoint_rescales_the_grid(Fn, tmp_path):
    from voxneus_amd.checkpoint import load_checkpoint, save_checkpoint
    from voxneus_amd.voxneus import VoxNeuS
    m = _random_model(24, 1)
    path = os.path.join(tmp_path, "ck.pt")
    save_checkpoint(path, 3, m, {}, {"model": {"grid_res": 24}})
    m2 = VoxNeuS(grid_res=16, hash_log2_size=14).to(DEV)
    want = R.upsample_sdf_grid(m2.sdf_grid.grid.detach().cpu(), (40, 40, 40))
    m2.sdf_grid.scale_volume_grid(40)
    assert torch.equal(m2.sdf_grid.grid.detach().cpu(), want)
    assert [int(v) for v in m2.sdf_grid.res] == [40, 40, 40]
    load_checkpoint(path, m2, device=DEV)
    assert [int(v) for v in m2.sdf_grid.res] == [24, 24, 24]
    assert torch.equal(m2.sdf_grid.grid, m.sdf_grid.grid)


def test_resident_scene_from_checkpoint_serves_like_the_model(Fn, tmp_path):
    from voxneus_amd import _native as N
    from voxneus_amd.checkpoint import load_resident_scene, save_checkpoint
    from voxneus_amd.serving import QueryRequest, RenderRequest, Server, pinhole_rays
    m = _random_model(48, 2)
    path = os.path.join(tmp_path, "scene.pt")
    save_checkpoint(path, 7, m, {}, {"model": {"grid_res": 48, "s_max": 300.0, "gradient_mode": "analytical"}})
    sc = load_resident_scene(path, DEV)
    assert sc.gradient_mode == N.VXN_GRADIENT_ANALYTICAL
    assert sc.s == float(m.s.detach())
    cam = S.hemisphere_cameras(49)[5]
    req = RenderRequest(0, cam, 40, 30, 60.0, 60.0, 20.0, 15.0)
    pts = torch.rand(3000, 3, device=DEV) * 2.2 - 1.1
    res = Server([sc], ray_capacity=1024, point_capacity=2048, ray_arena=4096, point_arena=4096).serve(
        [req, QueryRequest(0, pts)], realtime=False)
    grid3 = m.sdf_grid.grid.detach().reshape(48, 48, 48)
    box6 = m.sdf_grid.box6
    o, d = pinhole_rays(req, DEV)
    z, dists, mid = Fn.upsample(o, d, None, None, grid3, box6, 64, 64, 4, False, m.occupancy_grid.occupancy_mask,
                                m.occupancy_grid.box6, bound_radius=1.0)
    cfg = Fn.RenderConfig(box6, m.radiance_hash.encoder.level_res, gradient_mode=N.VXN_GRADIENT_ANALYTICAL)
    mlp = torch.cat([p.detach().reshape(-1) for p in (m.radiance_mlp.net[0].weight, m.radiance_mlp.net[0].bias,
                                                      m.radiance_mlp.net[2].weight, m.radiance_mlp.net[2].bias)])
    out = Fn.RenderFunction.apply(grid3, m.radiance_hash.encoder.embeddings.detach(), mlp, m.s.detach(), o, d, mid,
                                  dists, cfg, False, None)
    assert torch.equal(res[0].outputs["rgb"], out[0].reshape(res[0].outputs["rgb"].shape))
    assert torch.equal(res[0].outputs["depth"], out[1].reshape(res[0].outputs["depth"].shape))
    q = Fn.grid_query_raw(grid3, m.sdf_grid.box, pts, want_sdf=True, want_grad=False, want_normal=True,
                          gradient_mode=N.VXN_GRADIENT_ANALYTICAL)
    assert torch.equal(res[1].outputs["sdf"].reshape(-1), q["sdf"]) and torch.equal(res[1].outputs["normal"],
                                                                                     q["normal"])
