"""vxn_reg_eik_curv_vr_noalloc (the regulariser pass with its partial sums in a static
device array instead of a workspace) against vxn_reg_eik_curv_vr, through the C ABI, on
two calls in a row, the second on a smaller grid: V_R counts and the FP64 losses
bit-identical (fixed-order reductions), the gradient within 1e-6 of the largest magnitude
(the tiles' float atomics add in another order)."""

import ctypes

import numpy as np
import pytest
import torch

from voxneus_amd import _native as N
from voxneus_amd import functional as Fn

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
BOX6 = (-1.0, -1.0, -1.0, 1.0, 1.0, 1.0)


def _case(res, n_corners, seed):
    rng = np.random.default_rng(seed)
    data = torch.from_numpy(rng.normal(size=(res, res, res)).astype(np.float32)).to(DEV)
    corners = torch.from_numpy(rng.integers(0, res, size=(n_corners, 8, 3))).to(DEV)
    bm = torch.zeros(Fn.vr_bitmask_words((res, res, res)), device=DEV, dtype=torch.int32)
    Fn.vr_mark_corners(corners, (res, res, res), bm)
    return data, bm


def _call(noalloc, data, bm, relax, w_eik, w_curv):
    X, Y, Z = data.shape
    dev = data.device
    grad = torch.full((X, Y, Z), 0.25, device=dev)
    losses = torch.full((2,), -1.0, device=dev, dtype=torch.float64)
    counts = torch.full((2,), -1, device=dev, dtype=torch.int32)
    eps = Fn.voxel_size_of(BOX6, (X, Y, Z))
    L = N.lib()
    args = (data.data_ptr(), X, Y, Z, N.f3(eps), N.box6(*BOX6), float(relax), bm.data_ptr(), float(w_eik),
            float(w_curv), grad.data_ptr(), losses.data_ptr(), counts.data_ptr())
    s = torch.cuda.current_stream().cuda_stream
    if noalloc:
        N.check(L.vxn_reg_eik_curv_vr_noalloc(*args, s), "reg_eik_curv_vr_noalloc")
    else:
        nb = ctypes.c_int64(0)
        N.check(L.vxn_reg_vr_workspace_bytes(X, Y, Z, ctypes.byref(nb)), "reg_vr_workspace_bytes")
        ws = torch.empty(int(nb.value), device=dev, dtype=torch.uint8)
        N.check(L.vxn_reg_eik_curv_vr(*args, ws.data_ptr(), ws.numel(), s), "reg_eik_curv_vr")
    torch.cuda.synchronize()
    return grad, losses, counts


@pytest.mark.parametrize("relax", [1.2, 0.3])
def test_noalloc_equals_workspace_entry_point_twice(relax):
    for res, n, seed in ((56, 40000, 1), (24, 3000, 2)):  # the second call on a smaller grid
        print(f"noalloc regulariser: grid {res}^3, {n} cells marked, relax {relax}", flush=True)
        data, bm = _case(res, n, seed)
        g_ws, l_ws, c_ws = _call(False, data, bm, relax, 0.1, 5e-4)
        g_na, l_na, c_na = _call(True, data, bm, relax, 0.1, 5e-4)
        assert torch.equal(c_na, c_ws) and int(c_ws[0]) > 0
        assert torch.equal(l_na, l_ws)
        d_ws, d_na = (g_ws - 0.25).cpu().numpy(), (g_na - 0.25).cpu().numpy()
        assert np.abs(d_na - d_ws).max() <= 1e-6 * np.abs(d_ws).max()
        assert np.array_equal(d_na != 0, d_ws != 0)
