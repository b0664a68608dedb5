"""Marching cubes, the Gaussian pre-filter and grid super-resolution on the GPU: meshes
equal the CPU oracle's (vertices bit-exact, same triangles in the same order) on closed
shapes, random fields full of ambiguous cells and a 128^3 scene grid, with and without
the pre-filter; the filter equals scipy.ndimage.gaussian_filter(mode="nearest") bit for
bit; the reference's own upsampling checks (test_grid_upsampling.py) hold on the
device."""

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R
from voxneus_amd import scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


def _lattice(res):
    ax = np.linspace(-1.0, 1.0, res)
    return np.meshgrid(ax, ax, ax, indexing="ij")


def _check(vol, level, lo=(-1.0, -1.0, -1.0), hi=(1.0, 1.0, 1.0), sigma=None):
    from voxneus_amd.mesh import extract_mesh_marching_cubes
    v, f = extract_mesh_marching_cubes(vol, np.array(lo, np.float32), np.array(hi, np.float32), level, sigma)
    rv, rf = R.extract_mesh_marching_cubes(vol, lo, hi, level, sigma)
    assert v.dtype == np.float32 and f.dtype == np.int32
    assert np.array_equal(v, rv) and np.array_equal(f, rf)
    return v, f


def test_closed_shapes_match_oracle():
    X, Y, Z = _lattice(48)
    v, f = _check((np.sqrt(X * X + Y * Y + Z * Z) - 0.6).astype(np.float32), 0.0)
    assert f.shape[0] > 1000
    _check((np.sqrt((np.sqrt(X * X + Y * Y) - 0.55) ** 2 + Z * Z) - 0.2).astype(np.float32), 0.0,
           (-1.5, -1.0, -0.5), (1.5, 1.0, 0.5))


def test_random_fields_match_oracle():
    rng = np.random.default_rng(5)
    for shape, level in (((30, 31, 29), 0.1), ((2, 2, 2), 0.0), ((17, 5, 40), -0.3)):
        _check(rng.normal(size=shape).astype(np.float32), level)
    v, f = _check(np.ones((9, 9, 9), np.float32), 0.0)
    assert v.shape == (0, 3) and f.shape == (0, 3)


def test_scene_grid_matches_oracle():
    sc = scenes.make_scene("multi", 128, 17, seed=2)
    v, f = _check(np.ascontiguousarray(sc.grid, np.float32), 0.0)
    assert f.shape[0] > 20000


def test_reference_upsampling_checks():
    """test_grid_upsampling.py:13-39: the resolution schedule, and a 32^3 sphere of radius
    0.5 upsampled to 64^3 still has SDF ~ 0 on the sphere (to 2 places)."""
    from voxneus_amd.sdf_grid import DenseSDFGrid, get_resolution_for_step
    assert [get_resolution_for_step(s) for s in (0, 9999, 10000, 29999, 30000, 40000)] == [96, 96, 160, 160,
                                                                                          320, 320]
    grid = DenseSDFGrid(resolution=32, init_mode="sphere", sphere_radius=0.5).to(DEV)
    grid.scale_volume_grid(64)
    assert list(grid.grid.shape) == [1, 1, 64, 64, 64]
    pts = torch.tensor([[0.5, 0.0, 0.0], [0.0, 0.5, 0.0], [0.0, 0.0, 0.5], [-0.5, 0.0, 0.0]], device=DEV)
    sdf = grid.query_sdf(pts)
    assert (sdf.abs() < 0.005).all()


@pytest.mark.parametrize("shape", [(17, 23, 29), (64, 64, 64), (1, 9, 40), (2, 30, 3)])
@pytest.mark.parametrize("sigma", [0.1, 0.7, 1.0, 1.5, 2.3])
def test_gaussian_filter_matches_scipy(shape, sigma):
    from voxneus_amd.mesh import apply_gaussian_filter_to_grid
    rng = np.random.default_rng(sum(shape) + int(sigma * 10))
    vol = rng.normal(size=shape).astype(np.float32)
    got = apply_gaussian_filter_to_grid(vol, sigma)
    assert got.dtype == np.float32 and got.shape == vol.shape
    assert np.array_equal(got, R.apply_gaussian_filter_to_grid(vol, sigma))
    # a device tensor stays on the device
    t = apply_gaussian_filter_to_grid(torch.from_numpy(vol).to(DEV), sigma)
    assert t.is_cuda and np.array_equal(t.cpu().numpy(), got)


def test_filtered_meshes_match_oracle():
    X, Y, Z = _lattice(48)
    rng = np.random.default_rng(9)
    noisy = (np.sqrt(X * X + Y * Y + Z * Z) - 0.6 + 0.02 * rng.normal(size=X.shape)).astype(np.float32)
    v, f = _check(noisy, 0.0, sigma=1.0)
    r = np.linalg.norm(v, axis=1)
    assert f.shape[0] > 1000 and abs(float(r.mean()) - 0.6) < 0.02
    sc = scenes.make_scene("multi", 128, 17, seed=2)
    _check(np.ascontiguousarray(sc.grid, np.float32), 0.0, sigma=1.5)
