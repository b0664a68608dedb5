"""Pin the oracle's occupancy update (occupancy_grid.py:15-22, 62-119) on the CPU:
the logistic density is a probability density of the SDF value (unit mass, peak s/4 at
the surface), and the update's rules, checked against float64 restatements of each rule
on a field whose trilinear SDF is exact (a plane)."""

import math

import numpy as np
import torch

from oracle import ref_cpu as R


def test_logistic_density_is_a_unit_density_peaking_at_the_surface():
    s = 37.0
    d = torch.linspace(-2.0, 2.0, 400001, dtype=torch.float64)
    phi = R.logistic_density(d, torch.tensor(s, dtype=torch.float64))
    mass = float(torch.trapezoid(phi, d))
    assert abs(mass - 1.0) < 1e-9
    assert abs(float(phi[200000]) - s / 4.0) < 1e-12 and int(phi.argmax()) == 200000
    assert torch.allclose(phi, torch.flip(phi, [0]))


def _plane_grid(res=17):
    """f(x) = 0.6 x - 0.3 y + 0.2 z - 0.1 on the vertices: trilinear interpolation is exact."""
    ax = torch.linspace(-1, 1, res, dtype=torch.float64)
    X, Y, Z = torch.meshgrid(ax, ax, ax, indexing="ij")
    return (0.6 * X - 0.3 * Y + 0.2 * Z - 0.1)[None, None], R.GridBox(res, dtype=torch.float64)


def test_update_rules_on_a_plane():
    grid, box = _plane_grid()
    res = 8
    mn, mx = torch.full((3,), -1.0, dtype=torch.float64), torch.full((3,), 1.0, dtype=torch.float64)
    s = torch.tensor(20.0, dtype=torch.float64)
    g = torch.Generator().manual_seed(0)
    ix, iy, iz = (torch.randint(0, res, (300,), generator=g) for _ in range(3))
    ix[:3], iy[:3], iz[:3] = 2, 5, 1  # a cell drawn three times
    dens = torch.zeros((res,) * 3, dtype=torch.float64)
    mask = torch.ones((res,) * 3, dtype=torch.bool)
    R.occupancy_update(dens, mask, mn, mx, grid, box, s, ix, iy, iz, first_update=True)
    c = lambda i: -1.0 + (i.double() + 0.5) / res * 2.0
    d = 0.6 * c(ix) - 0.3 * c(iy) + 0.2 * c(iz) - 0.1
    sig = 1.0 / (1.0 + torch.exp(-20.0 * d))
    phi = 20.0 * sig * (1.0 - sig)
    near = d.abs() < 2.0 * math.sqrt(3.0) * (2.0 / res)
    want = torch.where(near, torch.maximum(phi, torch.full_like(phi, 0.02)), phi)
    assert torch.allclose(dens[ix, iy, iz], want, atol=1e-12)
    assert torch.equal(mask[ix, iy, iz], want > 0.01)
    untouched = torch.ones((res,) * 3, dtype=torch.bool)
    untouched[ix, iy, iz] = False
    assert untouched.any() and bool(mask[untouched].all()) and not dens[untouched].any()
    # second update: max(decay * previous, new), here with s lowered so the decay term wins
    prev = dens.clone()
    R.occupancy_update(dens, mask, mn, mx, grid, box, torch.tensor(2.0, dtype=torch.float64), ix, iy, iz,
                       first_update=False)
    sig2 = 1.0 / (1.0 + torch.exp(-2.0 * d))
    phi2 = torch.where(near, torch.maximum(2.0 * sig2 * (1.0 - sig2), torch.full_like(d, 0.02)),
                       2.0 * sig2 * (1.0 - sig2))
    want2 = torch.maximum(prev[ix, iy, iz] * 0.95, phi2)
    assert torch.allclose(dens[ix, iy, iz], want2, atol=1e-12)
    assert bool(((prev[ix, iy, iz] * 0.95) > phi2).any()) and bool(((prev[ix, iy, iz] * 0.95) < phi2).any())


def test_oracle_update_matches_reference_fixture():
    """Three OccupancyGrid.update_from_sdf updates written by the reference itself (32^3
    occupancy grid over the config-1 SDF grid, its seeded torch.randint draws, s = e^3, e^3,
    60): the oracle given the same draws reproduces the density grid and the mask after
    every update bit for bit (duplicate draws included: PyTorch's index_put keeps the last)."""
    from tests.conftest import load_golden
    from voxneus_amd import scenes

    g = load_golden("unit_occupancy.npz")
    grid = torch.from_numpy(scenes.config1_scene(seed=0).grid)
    dens, mask = torch.zeros((32,) * 3), torch.ones((32,) * 3, dtype=torch.bool)
    mn, mx = torch.full((3,), -1.0), torch.full((3,), 1.0)
    for k, s in enumerate(g["s_vals"].tolist()):
        idx = [torch.from_numpy(g[f"idx{k}"][a].astype(np.int64)) for a in range(3)]
        R.occupancy_update(dens, mask, mn, mx, grid[None, None], R.GridBox(64), torch.tensor(s), *idx,
                           first_update=k == 0)
        ref_mask = np.unpackbits(g[f"mask{k}"]).astype(bool).reshape(32, 32, 32)
        assert np.array_equal(dens.numpy(), g[f"density{k}"]), k
        assert np.array_equal(mask.numpy(), ref_mask), k
    assert 0 < int(mask.sum()) < mask.numel()
