"""The hand-written backward against FP64 central finite differences (h = 1e-4) with the
sample positions (z) and V_R frozen, on the differentiation-check scene: 128^3 grid of the
seeded multi-primitive SDF, 16 x 2^17 hash, 64 hemisphere rays x 128 samples, Huber colour
loss (delta 0.1), 1000 seeded grid voxels, 100 hash entries and the variance v (s = exp(10 v)).

The FP64 side is the oracle evaluated in float64 on the CPU.  A grid voxel only changes
the rays with a sample whose 4x4x4 stencil brick holds it, and a hash entry only the rays
with a sample hashing to it, so each central difference re-renders just those rays;
coordinates with disjoint ray sets share one pair of renders.  The eikonal and curvature
gradients are checked against central differences of their losses (mean over the
interior of V_R = {|sdf| < 3 voxels}); the closed forms scale by 2/|V_R| where the loss
averages over interior vertices only (eikonal.py:44,152), so the FD side is multiplied by
|V_int| / |V_R| (1 here: V_R has no boundary vertices).
"""

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R
from voxneus_amd import functional as Fn
from voxneus_amd import scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
BOX6 = (-1.0, -1.0, -1.0, 1.0, 1.0, 1.0)
H = 1e-4
N_RAYS = 64
DELTA = 0.1


@pytest.fixture(scope="module")
def setup():
    sc = scenes.make_scene("multi", 128, 17, seed=2)
    o, d, tgt = scenes.config_rays("multi", N_RAYS, seed=21, shape=sc.shape)
    res = R.hash_level_resolutions(16, 16, sc.hash_max_resolution)
    og, dg = torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV)
    occ = torch.ones((128,) * 3, dtype=torch.bool, device=DEV)
    z, dists, mid = Fn.upsample(og, dg, None, None, torch.from_numpy(sc.grid).to(DEV), BOX6, 64, 64, 4, False, occ,
                                BOX6, bound_radius=1.0)
    return {"sc": sc, "o": o, "d": d, "t": tgt, "res": res, "zs": (z.cpu(), dists.cpu(), mid.cpu())}


def _hip_grads(st, bf16=False):
    sc = st["sc"]
    from voxneus_amd import _native as N

    grid = torch.from_numpy(sc.grid).to(DEV).requires_grad_(True)
    emb = torch.from_numpy(sc.embeddings).to(DEV)
    gt = torch.zeros(emb.shape, device=DEV)
    if bf16:
        emb = emb.to(torch.bfloat16)
    mlp = torch.cat([torch.from_numpy(sc.mlp[k]).reshape(-1) for k in ("W0", "b0", "W1", "b1")]).to(DEV)
    mlp.requires_grad_(True)
    v = torch.tensor(0.3, device=DEV, requires_grad=True)
    s = torch.clamp(torch.exp(v * 10.0), 1e-4, 300.0)
    cfg = Fn.RenderConfig(BOX6, st["res"], mlp_mode=N.VXN_MLP_BF16_MFMA if bf16 else N.VXN_MLP_FP32,
                          table_grad_out=gt)
    z, dists, mid = (t.to(DEV) for t in st["zs"])
    out = Fn.RenderFunction.apply(grid, emb, mlp, s, torch.from_numpy(st["o"]).to(DEV),
                                  torch.from_numpy(st["d"]).to(DEV), mid, dists, cfg, False, None)
    loss = R.huber(out[0], torch.from_numpy(st["t"]).to(DEV), DELTA)
    loss.backward()
    return grid.grad.cpu().numpy(), gt.cpu().numpy(), v.grad.item(), mlp.grad.cpu().numpy()


class _FP64:
    """The oracle in float64 with frozen z: per-ray Huber sums of any subset of rays."""

    def __init__(self, st, table=None):
        sc = st["sc"]
        emb = sc.embeddings if table is None else table
        self.m = R.RefModel(sc.grid, emb, **sc.mlp, hash_max_resolution=sc.hash_max_resolution,
                            dtype=torch.float64)
        self.o, self.d = torch.from_numpy(st["o"]).double(), torch.from_numpy(st["d"]).double()
        self.t = torch.from_numpy(st["t"]).double()
        self.zs = tuple(z.double() for z in st["zs"])

    def ray_losses(self, rays):
        rays = torch.as_tensor(rays, dtype=torch.int64)
        with torch.no_grad():
            out = self.m.render(self.o[rays], self.d[rays], None, None, zs=tuple(z[rays] for z in self.zs))
            diff = (out["rgb"] - self.t[rays]).abs()
            hub = torch.where(diff < DELTA, 0.5 * diff * diff / DELTA, diff - 0.5 * DELTA)
        return hub.sum(-1).numpy()

    def fd(self, coords, ray_sets, param, h=H):
        """Central differences of the mean Huber loss w.r.t. param[c] for each coordinate,
        grouping coordinates whose ray sets are disjoint."""
        scale = 1.0 / (3.0 * N_RAYS)
        out = np.zeros(len(coords))
        left = list(range(len(coords)))
        while left:
            taken, group, rest = set(), [], []
            for i in left:
                rs = set(ray_sets[i].tolist())
                if rs & taken:
                    rest.append(i)
                else:
                    group.append(i)
                    taken |= rs
            rays = np.array(sorted(taken), np.int64)
            owner = {}
            for i in group:
                for r in ray_sets[i].tolist():
                    owner[r] = i
            vals = {}
            for sgn in (1.0, -1.0):
                with torch.no_grad():
                    for i in group:
                        param[coords[i]] += sgn * h
                vals[sgn] = self.ray_losses(rays)
                with torch.no_grad():
                    for i in group:
                        param[coords[i]] -= sgn * h
            delta = (vals[1.0] - vals[-1.0]) / (2.0 * h) * scale
            for k, r in enumerate(rays.tolist()):
                out[owner[r]] += delta[k]
            left = rest
        return out


def _sample_points(st):
    z, dists, mid = st["zs"]
    o, d = torch.from_numpy(st["o"]).double(), torch.from_numpy(st["d"]).double()
    return o[:, None, :] + d[:, None, :] * mid.double()[..., None]  # [R, S, 3]


def _report(name, g_hip, g_fd, smooth=None):
    err = np.abs(g_hip - g_fd)
    scale = np.abs(g_fd).max()
    rel = err / scale
    if smooth is not None:
        rel = rel[smooth]
    print(f"{name}: n={len(rel)} of {len(g_fd)} max|fd|={scale:.3e} max err/scale={rel.max():.2e} "
          f"p99={np.percentile(rel, 99):.2e} median={np.median(rel):.2e}")
    return rel


def _smooth(f64, coords, sets, param, fd_h):
    """Coordinates where the loss is smooth at the FD scale: central differences at h and
    h/4 agree to 1e-6 of the largest (a ReLU, Huber or transmittance-cutoff kink within h
    breaks this; about 2% of the seeded voxels sit next to one)."""
    fd_q = f64.fd(coords, sets, param, h=H / 4)
    ok = np.abs(fd_h - fd_q) <= 1e-6 * np.abs(fd_h).max()
    print(f"  {(~ok).sum()} of {len(ok)} coordinates are non-smooth at h and left out")
    assert ok.mean() >= 0.97
    return ok


def test_render_backward_vs_fp64_finite_differences(setup):
    st = setup
    gg, gt, gv, _ = _hip_grads(st)
    f64 = _FP64(st)
    P = _sample_points(st)
    rng = np.random.default_rng(12)
    # grid voxels: 1000 of those the rays touch
    coords = (P + 1.0) / (2.0 + 1e-8) * 127.0
    c0 = torch.floor(torch.clamp(coords, 0.0, 127.0 - 1e-4)).to(torch.int64).numpy()  # [R, S, 3]
    touched = np.argwhere(gg != 0)
    vox = touched[rng.choice(len(touched), size=1000, replace=False)]
    vsets = []
    for v in vox:
        inside = np.all((c0 - 1 <= v) & (v <= c0 + 2), axis=-1)  # [R, S]
        vsets.append(np.nonzero(inside.any(-1))[0])
    gcoords = [(0, 0, int(a), int(b), int(c)) for a, b, c in vox]
    fd_grid = f64.fd(gcoords, vsets, f64.m.grid)
    ok = _smooth(f64, gcoords, vsets, f64.m.grid, fd_grid)
    rel = _report("grid", gg[tuple(vox.T)], fd_grid, ok)
    assert rel.max() < 1e-4 and np.percentile(rel, 99) < 2e-5
    # hash entries: 100 of those with a gradient
    nz = np.argwhere(gt != 0)
    ent = nz[rng.choice(len(nz), size=100, replace=False)]
    esets = _hash_entry_sets(st, ent)
    hcoords = [tuple(int(x) for x in c) for c in ent]
    fd_tab = f64.fd(hcoords, esets, f64.m.embeddings)
    ok = _smooth(f64, hcoords, esets, f64.m.embeddings, fd_tab)
    rel = _report("hash", gt[tuple(ent.T)], fd_tab, ok)
    # hash-entry gradients are small sums of many FP32 contributions: FP32 noise ~5e-5
    assert rel.max() < 2e-4 and np.percentile(rel, 99) < 1e-4
    # the variance v (every ray)
    fd_v = f64.fd([()], [np.arange(N_RAYS)], f64.m.variance)[0]
    print(f"v: hip {gv:.6e} fd {fd_v:.6e}")
    assert abs(gv - fd_v) <= 1e-4 * abs(fd_v)
    # and the whole gradient against the FP64 autograd of the same loss
    out = f64.m.render(f64.o, f64.d, None, None, zs=f64.zs)
    R.huber(out["rgb"], f64.t, DELTA).backward()
    ga, ta = f64.m.grid.grad.numpy()[0, 0], f64.m.embeddings.grad.numpy()
    e_g = np.abs(gg - ga).max() / np.abs(ga).max()
    e_t = np.abs(gt - ta).max() / np.abs(ta).max()
    print(f"full gradient vs fp64 autograd: grid {e_g:.2e} hash {e_t:.2e}")
    assert e_g < 1e-4 and e_t < 3e-4


def _hash_entry_sets(st, ent):
    """For each (level, entry, feature): the rays with a sample that hashes to it."""
    P = _sample_points(st)
    T = st["sc"].embeddings.shape[1]
    x01 = torch.clamp((P + 1.0) / (2.0 + 1e-8), 0.0, 1.0)
    esets = []
    for l, e, _ in ent:
        s = x01 * float(st["res"][l])
        b = torch.floor(s).to(torch.int64)
        hit = torch.zeros(P.shape[:2], dtype=torch.bool)
        for k in range(8):
            c = b + torch.tensor([(k >> 2) & 1, (k >> 1) & 1, k & 1])
            hit |= R.hash_index(c, T) == int(e)
        esets.append(np.nonzero(hit.any(-1).numpy())[0])
    return esets


def test_bf16_render_backward_vs_fp64_finite_differences(setup):
    """The BF16 form's gradient error against FP64 central differences of the model with
    the same (bf16-valued) hash table, stated and bounded: 300 grid voxels, 100 hash
    entries, 24 entries of W0 (hash, SH and normal columns) and 6 of W1, and v.  The
    MLP's weights enter the BF16 form rounded to bf16, so their gradients carry bf16-scale
    error beside the FP32 form's; a ReLU or Huber branch that the bf16 rounding flips
    moves a single coordinate by a whole term, so the bounds are on the 99th percentile
    with a looser maximum."""
    st = setup
    gg, gt, gv, gm = _hip_grads(st, bf16=True)
    table = torch.from_numpy(st["sc"].embeddings).to(torch.bfloat16).float().numpy()
    f64 = _FP64(st, table)
    P = _sample_points(st)
    rng = np.random.default_rng(13)
    coords = (P + 1.0) / (2.0 + 1e-8) * 127.0
    c0 = torch.floor(torch.clamp(coords, 0.0, 127.0 - 1e-4)).to(torch.int64).numpy()
    touched = np.argwhere(gg != 0)
    vox = touched[rng.choice(len(touched), size=300, replace=False)]
    vsets = [np.nonzero(np.all((c0 - 1 <= v) & (v <= c0 + 2), axis=-1).any(-1))[0] for v in vox]
    gcoords = [(0, 0, int(a), int(b), int(c)) for a, b, c in vox]
    fd_grid = f64.fd(gcoords, vsets, f64.m.grid)
    ok = _smooth(f64, gcoords, vsets, f64.m.grid, fd_grid)
    rel = _report("bf16 grid", gg[tuple(vox.T)], fd_grid, ok)
    assert np.percentile(rel, 99) < 5e-3 and rel.max() < 2e-2
    # hash entries: 100 of those with a gradient
    nz = np.argwhere(gt != 0)
    ent = nz[rng.choice(len(nz), size=100, replace=False)]
    esets = _hash_entry_sets(st, ent)
    hcoords = [tuple(int(x) for x in c) for c in ent]
    fd_tab = f64.fd(hcoords, esets, f64.m.embeddings)
    ok = _smooth(f64, hcoords, esets, f64.m.embeddings, fd_tab)
    rel = _report("bf16 hash", gt[tuple(ent.T)], fd_tab, ok)
    assert np.percentile(rel, 99) < 5e-2 and rel.max() < 1e-1
    # MLP weights: every ray depends on them, so each coordinate is its own pair of renders
    w0 = [(int(i), int(j)) for i, j in zip(rng.integers(0, 64, 24), np.r_[rng.integers(0, 32, 12),
                                                                            rng.integers(32, 48, 8),
                                                                            rng.integers(48, 51, 4)])]
    w1 = [(int(i), int(j)) for i, j in zip(rng.integers(0, 3, 6), rng.integers(0, 64, 6))]
    all_rays = np.arange(N_RAYS)
    fd_w0 = f64.fd(w0, [all_rays] * len(w0), f64.m.W0)
    fd_w1 = f64.fd(w1, [all_rays] * len(w1), f64.m.W1)
    g_w0 = np.array([gm[i * 51 + j] for i, j in w0])
    g_w1 = np.array([gm[3328 + i * 64 + j] for i, j in w1])
    rel0 = np.abs(g_w0 - fd_w0) / np.abs(fd_w0).max()
    rel1 = np.abs(g_w1 - fd_w1) / np.abs(fd_w1).max()
    print(f"bf16 W0: n={len(w0)} max|fd|={np.abs(fd_w0).max():.3e} max err/scale={rel0.max():.2e} "
          f"median={np.median(rel0):.2e}; W1: n={len(w1)} max|fd|={np.abs(fd_w1).max():.3e} "
          f"max err/scale={rel1.max():.2e}")
    assert rel0.max() < 5e-2 and rel1.max() < 2e-2
    fd_v = f64.fd([()], [np.arange(N_RAYS)], f64.m.variance)[0]
    print(f"bf16 v: hip {gv:.6e} fd {fd_v:.6e}")
    assert abs(gv - fd_v) <= 5e-2 * abs(fd_v)


def test_regulariser_gradients_vs_fp64_finite_differences(setup):
    sc = setup["sc"]
    res = 128
    eps = 2.0 / (res - 1)
    data = sc.grid.astype(np.float64)
    vr = np.argwhere(np.abs(sc.grid) < 3.0 * eps)  # V_R: vertices within 3 voxels of the zero level
    lin = (vr[:, 0] * res + vr[:, 1]) * res + vr[:, 2]
    bm = np.zeros((res ** 3 + 31) // 32, np.uint32)
    np.bitwise_or.at(bm, lin // 32, (1 << (lin % 32)).astype(np.uint32))
    lst, counts = Fn.vr_compact(torch.from_numpy(bm.view(np.int32)).to(DEV), (res,) * 3, BOX6, relax_radius=1.2)
    n_all, n_in = counts.tolist()
    assert n_all == len(vr) and n_in == len(vr)
    g3 = torch.from_numpy(sc.grid).to(DEV)
    vs = Fn.voxel_size_of(BOX6, (res,) * 3)
    ge, gc = torch.zeros_like(g3), torch.zeros_like(g3)
    Fn.reg_eik_curv(g3, vs, lst, counts, 1.0, 0.0, ge)
    Fn.reg_eik_curv(g3, vs, lst, counts, 0.0, 1.0, gc)
    ge, gc = ge.cpu().numpy(), gc.cpu().numpy()
    interior = np.all((vr > 0) & (vr < res - 1), axis=-1)
    vint = vr[interior]
    quirk = len(vint) / len(vr)
    is_int = np.zeros((res,) * 3, bool)
    is_int[tuple(vint.T)] = True
    e = np.array(vs, np.float64)

    def eik_term(D, u):
        n = [(D[u[0] + 1, u[1], u[2]] - D[u[0] - 1, u[1], u[2]]) / (2 * e[0]),
             (D[u[0], u[1] + 1, u[2]] - D[u[0], u[1] - 1, u[2]]) / (2 * e[1]),
             (D[u[0], u[1], u[2] + 1] - D[u[0], u[1], u[2] - 1]) / (2 * e[2])]
        return (np.sqrt(sum(x * x for x in n)) - 1.0) ** 2

    def curv_term(D, u):
        f = D[tuple(u)]
        lx = (D[u[0] + 1, u[1], u[2]] + D[u[0] - 1, u[1], u[2]] - 2 * f) / e[0] ** 2
        ly = (D[u[0], u[1] + 1, u[2]] + D[u[0], u[1] - 1, u[2]] - 2 * f) / e[1] ** 2
        lz = (D[u[0], u[1], u[2] + 1] + D[u[0], u[1], u[2] - 1] - 2 * f) / e[2] ** 2
        return lx * lx + ly * ly + lz * lz

    rng = np.random.default_rng(14)
    cand = np.argwhere((ge != 0) | (gc != 0))
    vox = cand[rng.choice(len(cand), size=1000, replace=False)]
    offs = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
    fd_e, fd_c = np.zeros(len(vox)), np.zeros(len(vox))
    for i, w in enumerate(vox):
        aff = [w + o for o in offs if np.all(w + o >= 0) and np.all(w + o < res) and is_int[tuple(w + o)]]
        for sgn in (1.0, -1.0):
            data[tuple(w)] += sgn * H
            fd_e[i] += sgn * sum(eik_term(data, u) for u in aff)
            fd_c[i] += sgn * sum(curv_term(data, u) for u in aff)
            data[tuple(w)] -= sgn * H
    fd_e *= quirk / (2 * H * len(vint))
    fd_c *= quirk / (2 * H * len(vint))
    rel = _report("eikonal", ge[tuple(vox.T)], fd_e)
    assert np.percentile(rel, 99) < 1e-4 and rel.max() < 1e-3
    rel = _report("curvature", gc[tuple(vox.T)], fd_c)
    assert np.percentile(rel, 99) < 1e-4 and rel.max() < 1e-3
