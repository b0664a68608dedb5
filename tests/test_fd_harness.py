"""The finite-difference harness of the GPU gradient checks, pinned on the CPU: its
grouped, locality-restricted FP64 central differences must equal the FP64 oracle's own
autograd gradient (same frozen z) on grid voxels, hash entries and the variance."""

import numpy as np
import torch

from oracle import ref_cpu as R
from tests import test_gpu_fd as T
from voxneus_amd import scenes


def test_fd_harness_matches_fp64_autograd():
    sc = scenes.make_scene("multi", 64, 15, seed=2)
    o, d, tgt = scenes.config_rays("multi", 24, seed=21, shape=sc.shape)
    res = R.hash_level_resolutions(16, 16, sc.hash_max_resolution)
    m32 = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    ot, dt = torch.from_numpy(o), torch.from_numpy(d)
    near, far = R.near_far_from_sphere(ot, dt)
    st = {"sc": sc, "o": o, "d": d, "t": tgt, "res": res, "zs": tuple(z.float() for z in m32.sample(ot, dt, near, far))}
    n_rays, T.N_RAYS = T.N_RAYS, 24
    try:
        f64 = T._FP64(st)
        out = f64.m.render(f64.o, f64.d, None, None, zs=f64.zs)
        R.huber(out["rgb"], f64.t, T.DELTA).backward()
        gg, gt, gv = f64.m.grid.grad.numpy()[0, 0].copy(), f64.m.embeddings.grad.numpy().copy(), f64.m.variance.grad.item()
        P = T._sample_points(st)
        c0 = torch.floor(torch.clamp((P + 1.0) / (2.0 + 1e-8) * 63.0, 0.0, 63.0 - 1e-4)).to(torch.int64).numpy()
        rng = np.random.default_rng(3)
        touched = np.argwhere(gg != 0)
        vox = touched[rng.choice(len(touched), size=60, replace=False)]
        vsets = [np.nonzero(np.all((c0 - 1 <= v) & (v <= c0 + 2), axis=-1).any(-1))[0] for v in vox]
        fd = f64.fd([(0, 0, int(a), int(b), int(c)) for a, b, c in vox], vsets, f64.m.grid)
        assert np.percentile(np.abs(fd - gg[tuple(vox.T)]) / np.abs(fd).max(), 95) < 1e-6
        nz = np.argwhere(gt != 0)
        ent = nz[rng.choice(len(nz), size=20, replace=False)]
        x01 = torch.clamp((P + 1.0) / (2.0 + 1e-8), 0.0, 1.0)
        esets = []
        for l, e, _ in ent:
            b = torch.floor(x01 * float(res[l])).to(torch.int64)
            hit = torch.zeros(P.shape[:2], dtype=torch.bool)
            for k in range(8):
                hit |= R.hash_index(b + torch.tensor([(k >> 2) & 1, (k >> 1) & 1, k & 1]), sc.embeddings.shape[1]) == int(e)
            esets.append(np.nonzero(hit.any(-1).numpy())[0])
        fdt = f64.fd([tuple(int(x) for x in c) for c in ent], esets, f64.m.embeddings)
        assert np.abs(fdt - gt[tuple(ent.T)]).max() <= 1e-6 * np.abs(fdt).max()
        fdv = f64.fd([()], [np.arange(24)], f64.m.variance)[0]
        assert abs(fdv - gv) <= 1e-6 * abs(gv)
    finally:
        T.N_RAYS = n_rays
