"""GPU parity of the occupancy EMA update (occupancy_grid.py:62-119) with the CPU oracle
on the same cells: densities and masks bit-exact (cell counts are multiples of 16, so
PyTorch's CPU sigmoid takes its vectorised path for every cell, the one the kernel
reproduces)."""

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R
from voxneus_amd import scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
BOX6 = (-1.0, -1.0, -1.0, 1.0, 1.0, 1.0)


@pytest.fixture(scope="module")
def Fn():
    from voxneus_amd import functional
    return functional


def _grid():
    sc = scenes.config1_scene(seed=0)
    return torch.from_numpy(sc.grid)


def test_indexed_update_matches_oracle(Fn):
    grid = _grid()
    res = 48
    g = torch.Generator().manual_seed(3)
    n = 16 * 2000
    idx = [torch.randint(0, res, (n,), generator=g) for _ in range(3)]
    for t in idx:
        t[:16] = t[16:32]  # duplicates
    mn, mx = torch.full((3,), -1.0), torch.full((3,), 1.0)
    dens, mask = torch.zeros((res,) * 3), torch.ones((res,) * 3, dtype=torch.bool)
    dd, md = dens.to(DEV), mask.to(DEV)
    grid_d = grid.to(DEV)
    for k, s in enumerate((20.0, 3.0, 55.0)):
        R.occupancy_update(dens, mask, mn, mx, grid[None, None], R.GridBox(64), torch.tensor(s), *idx,
                           first_update=k == 0)
        Fn.occupancy_update(dd, md, BOX6, grid_d, BOX6, s, 0.95, 0.01, k == 0, tuple(t.to(DEV) for t in idx))
        assert torch.equal(dd.cpu(), dens) and torch.equal(md.cpu(), mask), k
        if k == 0:  # a sharp s leaves cells far from the surface empty
            assert 0 < int(mask.sum()) < mask.numel()


def test_full_sweep_matches_oracle(Fn):
    grid = _grid()
    res = 32
    ax = torch.arange(res)
    X, Y, Z = torch.meshgrid(ax, ax, ax, indexing="ij")
    idx = (X.reshape(-1), Y.reshape(-1), Z.reshape(-1))
    mn, mx = torch.full((3,), -1.0), torch.full((3,), 1.0)
    dens, mask = torch.zeros((res,) * 3), torch.ones((res,) * 3, dtype=torch.bool)
    dd, md = dens.to(DEV), mask.to(DEV)
    for k, s in enumerate((10.0, 40.0)):
        R.occupancy_update(dens, mask, mn, mx, grid[None, None], R.GridBox(64), torch.tensor(s), *idx,
                           first_update=k == 0)
        Fn.occupancy_update(dd, md, BOX6, grid.to(DEV), BOX6, s, 0.95, 0.01, k == 0, None)
        assert torch.equal(dd.cpu(), dens) and torch.equal(md.cpu(), mask), k


def test_model_update_draws_cells_like_the_reference(Fn):
    """VoxNeuS.update_occupancy_grid: three torch.randint draws of max(res^3/4, 1024) cells
    from the generator; the same draws on the oracle give the same grid."""
    from voxneus_amd.rendering import OccupancyGrid
    from voxneus_amd.sdf_grid import DenseSDFGrid
    grid = _grid()
    sg = DenseSDFGrid(64).to(DEV)
    with torch.no_grad():
        sg.grid.copy_(grid.reshape(sg.grid.shape))
    occ = OccupancyGrid(32).to(DEV)
    gen = torch.Generator(device=DEV)
    gen.manual_seed(11)
    occ.update_from_sdf(sg.query_sdf, torch.tensor(25.0, device=DEV), 0.25, generator=gen)
    gen.manual_seed(11)
    n = max(int(32 ** 3 * 0.25), 1024)
    idx = [torch.randint(0, 32, (n,), device=DEV, generator=gen).cpu() for _ in range(3)]
    dens, mask = torch.zeros((32,) * 3), torch.ones((32,) * 3, dtype=torch.bool)
    R.occupancy_update(dens, mask, torch.full((3,), -1.0), torch.full((3,), 1.0), grid[None, None], R.GridBox(64),
                       torch.tensor(25.0), *idx, first_update=True)
    assert torch.equal(occ.density_grid.cpu(), dens) and torch.equal(occ.occupancy_mask.cpu(), mask)
    assert occ.total_updates == 1 and 0.0 < occ.occupied_percentage < 100.0


def test_update_matches_reference_fixture(Fn):
    """The device update, given the reference's own draws, reproduces the density grid and
    mask the reference wrote after each of three updates (unit_occupancy.npz) bit for bit."""
    from tests.conftest import load_golden

    g = load_golden("unit_occupancy.npz")
    grid_d = _grid().to(DEV)
    dd, md = torch.zeros((32,) * 3, device=DEV), torch.ones((32,) * 3, dtype=torch.bool, device=DEV)
    for k, s in enumerate(g["s_vals"].tolist()):
        idx = tuple(torch.from_numpy(g[f"idx{k}"][a].astype(np.int64)).to(DEV) for a in range(3))
        Fn.occupancy_update(dd, md, BOX6, grid_d, BOX6, s, 0.95, 0.01, k == 0, idx)
        ref_mask = np.unpackbits(g[f"mask{k}"]).astype(bool).reshape(32, 32, 32)
        assert np.array_equal(dd.cpu().numpy(), g[f"density{k}"]), k
        assert np.array_equal(md.cpu().numpy(), ref_mask), k
