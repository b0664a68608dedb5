"""Ray setup upstream of the path (rendering/rays.py), on the CPU: the reference's own
checks (test_rays.py:14-58) and the serving loop's pinhole rays equal get_rays."""

import math

import numpy as np
import torch

from voxneus_amd.serving import RenderRequest, pinhole_rays
from voxneus_amd.training import get_rays, near_far_from_sphere


def test_reference_ray_checks():
    H = W = 100
    K = np.array([[120.0, 0.0, 50.0], [0.0, 120.0, 50.0], [0.0, 0.0, 1.0]], np.float32)
    c2w = torch.eye(4)
    c2w[2, 3] = 3.0
    o, d = get_rays(H, W, K, c2w)
    assert o.shape == (H, W, 3) and d.shape == (H, W, 3)
    assert torch.allclose(o[..., 2], torch.tensor(3.0))
    assert torch.allclose(torch.linalg.norm(d, dim=-1), torch.ones(H, W), atol=1e-5)
    # the pixel-centre direction before normalisation, in closed form
    i, j = 17, 63
    want = np.array([(i + 0.5 - 50.0) / 120.0, (j + 0.5 - 50.0) / 120.0, 1.0])
    assert np.allclose(d[j, i].numpy(), want / np.linalg.norm(want), atol=1e-6)
    near, far = near_far_from_sphere(torch.tensor([[0.0, 0.0, 3.0]]), torch.tensor([[0.0, 0.0, -1.0]]), radius=1.0)
    assert abs(near.item() - 2.0) < 1e-3 and abs(far.item() - 4.0) < 1e-3


def test_serving_rays_are_get_rays():
    th = 0.7
    c2w = np.eye(4, dtype=np.float32)
    c2w[:3, :3] = [[math.cos(th), 0, math.sin(th)], [0, 1, 0], [-math.sin(th), 0, math.cos(th)]]
    c2w[:3, 3] = [1.5, -0.2, 2.0]
    req = RenderRequest(0, c2w, 40, 30, 55.0, 57.0, 19.5, 15.5)
    o, d = pinhole_rays(req, "cpu")
    K = np.array([[55.0, 0, 19.5], [0, 57.0, 15.5], [0, 0, 1]], np.float32)
    ro, rd = get_rays(30, 40, K, torch.from_numpy(c2w))
    assert torch.equal(o, ro.reshape(-1, 3)) and torch.allclose(d, rd.reshape(-1, 3), atol=1e-7)
