"""GPU parity of the gfx950 kernels against the reference's golden vectors and the CPU
oracle (same seeded inputs).  Integer outputs (corners, hash features built from exact
products, searchsorted indices, z values, V_R) must be bit-identical; float gradients
are held to 1e-5 of the largest reference magnitude (FP32 form, atomics reorder sums).
"""

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R
from tests.conftest import load_golden
from voxneus_amd import scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
BOX = ((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
BOX6 = (-1.0, -1.0, -1.0, 1.0, 1.0, 1.0)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.fixture(scope="module")
def Fn():
    from voxneus_amd import functional
    return functional


def test_grid_query_matches_reference(Fn):
    g = load_golden("unit_grid.npz")
    sc = scenes.config1_scene(seed=0)
    grid = torch.from_numpy(sc.grid).to(DEV).requires_grad_(True)
    xyz = torch.from_numpy(g["pts"]).to(DEV)
    o = Fn.grid_query_raw(grid.detach(), BOX, xyz, True, True, True, True, True)
    assert np.array_equal(o["sdf"].cpu().numpy(), g["sdf"][:, 0])
    assert np.array_equal(o["corners"].cpu().numpy().astype(np.int16), g["corners"])
    assert np.array_equal(o["weights"].cpu().numpy(), g["weights"])
    assert rel_err(o["grad"].cpu(), g["grad"]) < 1e-6
    assert np.abs(o["normal"].cpu().numpy() - g["normal"]).max() < 1e-6
    sdf, gv = Fn.GridQueryFunction.apply(grid, xyz, BOX)
    loss = (gv * torch.from_numpy(g["r_g"]).to(DEV)).sum() + (sdf * torch.from_numpy(g["r_s"][:, 0]).to(DEV)).sum()
    loss.backward()
    ref = np.zeros(grid.numel(), np.float32)
    ref[g["grid_grad_idx"]] = g["grid_grad_val"]
    assert rel_err(grid.grad.cpu().numpy().reshape(-1), ref) < 1e-5


def test_gradient_interp_reference_shape(Fn):
    """Pattern of test_cuda_parity.py:31-70: random 32^3 grid, coords in [0, res-1.001]."""
    torch.manual_seed(42)
    res = 32
    grid = torch.randn(1, 1, res, res, res)
    coords = torch.rand(256, 3) * (res - 1.001)
    vs = torch.tensor([2.0 / (res - 1)] * 3)
    gres = torch.tensor([res] * 3, dtype=torch.float32)
    ref, _, _, _ = R.interpolate_gradients(grid, coords, vs, gres)
    gg = grid.to(DEV).requires_grad_(True)
    out = Fn.GradientInterpFunction.apply(gg, coords.to(DEV), vs)
    assert (out.detach().cpu() - ref).abs().max().item() < 1e-5
    gr = grid.clone().requires_grad_(True)
    rg, _, _, _ = R.interpolate_gradients(gr, coords, vs, gres)
    (rg ** 2).sum().backward()
    (out ** 2).sum().backward()
    assert (gg.grad.cpu() - gr.grad).abs().max().item() < 1e-4


def test_hash_matches_reference(Fn):
    g = load_golden("unit_hash.npz")
    sc = scenes.config1_scene(seed=0)
    emb = torch.from_numpy(sc.embeddings).to(DEV).requires_grad_(True)
    x = torch.from_numpy(g["x"]).to(DEV)
    res = g["level_res"].tolist()
    f = Fn.HashEncodeFunction.apply(emb, x, res, None)
    assert np.array_equal(f.detach().cpu().numpy(), g["feats"])
    from voxneus_amd.hash_grid import level_weights
    fl = Fn.HashEncodeFunction.apply(emb.detach(), x, res, level_weights(16, 5.5))
    assert np.array_equal(fl.cpu().numpy(), g["feats_lambda"])
    (f * torch.from_numpy(g["r"]).to(DEV)).sum().backward()
    eg = emb.grad.cpu().numpy().reshape(-1)
    assert int(np.count_nonzero(eg)) == int(g["emb_grad_nnz"])
    assert rel_err(eg[g["emb_grad_idx"]], g["emb_grad_val"]) < 1e-5
    assert abs(eg.astype(np.float64).sum() - float(g["emb_grad_sum"])) < 1e-5 * float(g["emb_grad_abs"])


def test_hash_bf16_table_reads_exact_values(Fn):
    sc = scenes.config1_scene(seed=0)
    emb = torch.from_numpy(sc.embeddings).to(torch.bfloat16)
    x = torch.rand(2048, 3)
    res = R.hash_level_resolutions(16, 16, sc.hash_max_resolution)
    ref = R.hash_encode(x, emb.float(), res)
    out = Fn.HashEncodeFunction.apply(emb.to(DEV), x.to(DEV), res, None)
    assert np.array_equal(out.cpu().numpy(), ref.numpy())


def _sampler_inputs(g):
    sc = scenes.config1_scene(seed=0)
    occ = np.unpackbits(g["occ_mask"])[: 32 ** 3].astype(bool).reshape(32, 32, 32)
    return sc, occ


def test_sampler_bit_exact_deterministic(Fn):
    g = load_golden("unit_sampler.npz")
    sc, occ = _sampler_inputs(g)
    o, d = torch.from_numpy(g["rays_o"]).to(DEV), torch.from_numpy(g["rays_d"]).to(DEV)
    near, far = torch.from_numpy(g["near"]).to(DEV), torch.from_numpy(g["far"]).to(DEV)
    grid = torch.from_numpy(sc.grid).to(DEV)
    z, dists, mid, inds = Fn.upsample(o, d, near, far, grid, BOX, occ_mask=torch.from_numpy(occ).to(DEV),
                                      return_inds=True)
    # searchsorted indices against the oracle's, step by step
    m = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, occupancy_mask=occ)
    rec = []
    m.sample(o.cpu(), d.cpu(), near.cpu(), far.cpu(), record=rec)
    ref_inds = np.stack([r["inds"].numpy() for r in rec if "inds" in r])
    got = inds.cpu().numpy()
    mism = int((got != ref_inds).sum())
    assert mism == 0, f"{mism} searchsorted indices differ"
    assert np.array_equal(z.cpu().numpy(), g["z_det"])
    assert np.array_equal(dists.cpu().numpy(), g["dists_det"])
    assert np.array_equal(mid.cpu().numpy(), g["mid_det"])


@pytest.mark.parametrize("res", [64, 48])
def test_sampler_vr_cell_marks_equal_corner_marks(Fn, res):
    """The sampler's V_R with vr_cells (one mark per sample cell, then the bitmask written
    whole as those cells' corners; grids whose Z is a multiple of 32) is the corner marks'
    bitmask bit for bit, overwriting what the bitmask held; with Z not a multiple of 32 the
    sampler marks the corners as without vr_cells and the cells stay zero."""
    sc = scenes.make_scene("multi", res, 15, seed=3)
    o, d = (torch.from_numpy(a).to(DEV) for a in scenes.hemisphere_ray_batch(2048, seed=7))
    grid = torch.from_numpy(sc.grid).to(DEV)
    words = Fn.vr_bitmask_words((res, res, res))
    ref = torch.zeros(words, device=DEV, dtype=torch.int32)
    z0 = Fn.upsample(o, d, None, None, grid, BOX, perturb=False, bound_radius=1.0, vr_bitmask=ref)[0]
    cells = torch.zeros(words, device=DEV, dtype=torch.int32)
    fill = 0x5A5A5A5A if res % 32 == 0 else 0
    got = torch.full((words,), fill, device=DEV, dtype=torch.int32)
    z1 = Fn.upsample(o, d, None, None, grid, BOX, perturb=False, bound_radius=1.0, vr_bitmask=got, vr_cells=cells)[0]
    torch.cuda.synchronize()
    assert torch.equal(z0, z1)
    n_ref = int(np.unpackbits(ref.cpu().numpy().view(np.uint8)).sum())
    assert n_ref > 1000
    assert torch.equal(got, ref), "V_R from the cell marks differs from the corner marks"
    if res % 32 == 0:
        assert int(cells.ne(0).sum()) > 0
    else:
        assert int(cells.ne(0).sum()) == 0


def test_sampler_bit_exact_injected_uniforms(Fn):
    g = load_golden("unit_sampler.npz")
    sc, occ = _sampler_inputs(g)
    o, d = torch.from_numpy(g["rays_o"]).to(DEV), torch.from_numpy(g["rays_d"]).to(DEV)
    near, far = torch.from_numpy(g["near"]).to(DEV), torch.from_numpy(g["far"]).to(DEV)
    grid = torch.from_numpy(sc.grid).to(DEV)
    z, dists, mid = Fn.upsample(o, d, near, far, grid, BOX, perturb=True, occ_mask=torch.from_numpy(occ).to(DEV),
                                t_rand=torch.from_numpy(g["t_rand"]).to(DEV),
                                u_rand=torch.from_numpy(g["u_rnd"]).to(DEV))
    assert np.array_equal(z.cpu().numpy(), g["z_rnd"])
    assert np.array_equal(mid.cpu().numpy(), g["mid_rnd"])


def _mlp_flat(sc):
    m = sc.mlp
    return torch.cat([torch.from_numpy(m[k]).reshape(-1) for k in ("W0", "b0", "W1", "b1")])


def _golden_scene(name):
    if name == "step_c1.npz":
        return scenes.config1_scene(seed=0)
    return scenes.make_scene("multi", 128, 17, seed=2)


@pytest.mark.parametrize("name", ["step_c1.npz", "step_c2.npz"])
def test_render_forward_matches_reference(Fn, name):
    g = load_golden(name)
    sc = _golden_scene(name)
    cfg = Fn.RenderConfig(BOX6, R.hash_level_resolutions(16, 16, sc.hash_max_resolution))
    grid = torch.from_numpy(sc.grid).to(DEV)
    emb = torch.from_numpy(sc.embeddings).to(DEV)
    s = torch.tensor(float(np.float32(np.exp(np.float32(0.3) * np.float32(10.0)))), device=DEV)
    s = torch.clamp(torch.exp(torch.tensor(0.3) * 10.0), 1e-4, 300.0).to(DEV)
    outs = Fn.RenderFunction.apply(grid, emb, _mlp_flat(sc).to(DEV), s, torch.from_numpy(g["rays_o"]).to(DEV),
                                   torch.from_numpy(g["rays_d"]).to(DEV), torch.from_numpy(g["mid"]).to(DEV),
                                   torch.from_numpy(g["dists"]).to(DEV), cfg, True, None)
    rgb, depth, normal, opac, w, alpha, T, sdf_s, grad_s, nrm_s, corners = [o.detach().cpu().numpy() if o is not None
                                                                           else None for o in outs]
    assert np.array_equal(sdf_s, g["sdf_samples"])
    assert np.array_equal(corners[..., 0, :].astype(np.int16), g["corner_c0"])
    assert np.abs(grad_s - g["grad_samples"]).max() <= 1e-6 * max(1.0, np.abs(g["grad_samples"]).max())
    for k, v in (("alpha", alpha), ("transmittance", T), ("weights", w), ("rgb", rgb), ("depth", depth),
                 ("opacity", opac)):
        assert np.abs(v - g[k]).max() < 1e-5, k
    # normal = normalize(sum w n) (volume_render.py:123-133): its rounding error grows as
    # 1 / |sum w n|, so the unnormalised composite is held to 1e-5 relative, and the unit
    # normal itself to 1e-5 on the rays that hit a surface (|sum w n| >= 0.1)
    m = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    with torch.no_grad():
        ro = m.render(torch.from_numpy(g["rays_o"]), torch.from_numpy(g["rays_d"]), None, None,
                      zs=tuple(torch.from_numpy(g[k]) for k in ("z_vals", "dists", "mid")))
    nsum = (ro["weights"][..., None] * ro["normal_samples"]).sum(1).norm(dim=-1).numpy()
    err = np.abs(normal - g["normal"]).max(-1)
    assert (err * nsum).max() <= 1e-5 * nsum.max()
    assert err[nsum >= 0.1].max() <= 1e-5


def test_render_backward_matches_oracle_autograd(Fn):
    """Autograd through the fused render vs the oracle's PyTorch autograd on the same z."""
    g = load_golden("step_c1.npz")
    sc = scenes.config1_scene(seed=0)
    m = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    o, d = torch.from_numpy(g["rays_o"]), torch.from_numpy(g["rays_d"])
    zs = tuple(torch.from_numpy(g[k]) for k in ("z_vals", "dists", "mid"))
    rng = np.random.default_rng(3)
    rr = [torch.from_numpy(rng.normal(size=sh).astype(np.float32)) for sh in ((64, 3), (64, 1), (64, 3), (64, 1))]
    out = m.render(o, d, None, None, zs=zs)
    L = (out["rgb"] * rr[0]).sum() + (out["depth"] * rr[1]).sum() + (out["normal"] * rr[2]).sum() + (
        out["opacity"] * rr[3]).sum()
    L.backward()
    cfg = Fn.RenderConfig(BOX6, m.level_res)
    grid = torch.from_numpy(sc.grid).to(DEV).requires_grad_(True)
    emb = torch.from_numpy(sc.embeddings).to(DEV).requires_grad_(True)
    mlp = _mlp_flat(sc).to(DEV).requires_grad_(True)
    v = torch.tensor(0.3, device=DEV, requires_grad=True)
    s = torch.clamp(torch.exp(v * 10.0), min=1e-4, max=300.0)
    outs = Fn.RenderFunction.apply(grid, emb, mlp, s, o.to(DEV), d.to(DEV), zs[2].to(DEV), zs[1].to(DEV), cfg, False,
                                   None)
    Lg = sum((outs[i] * rr[k].to(DEV)).sum() for k, i in enumerate((0, 1, 2, 3)))
    Lg.backward()
    assert rel_err(grid.grad.cpu().numpy(), m.grid.grad.numpy()[0, 0]) < 1e-5
    assert rel_err(emb.grad.cpu().numpy(), m.embeddings.grad.numpy()) < 1e-5
    ref_mlp = torch.cat([m.W0.grad.reshape(-1), m.b0.grad, m.W1.grad.reshape(-1), m.b1.grad]).numpy()
    assert rel_err(mlp.grad.cpu().numpy(), ref_mlp) < 1e-5
    assert abs(v.grad.item() - m.variance.grad.item()) <= 1e-5 * max(1e-3, abs(m.variance.grad.item()))


@pytest.mark.parametrize("log2_T", [19, 21, 22])
def test_render_backward_hash_gradient_large_tables(Fn, log2_T):
    """The hash-table gradient with 2^19 entries per level (binned: 256 bins of 2048
    entries per level), 2^21 (1024 bins per level, the binned path's limit: 68 KB of LDS
    in the write pass) and 2^22 (past it: the atomic scatter)."""
    g = load_golden("step_c1.npz")
    sc = scenes.config1_scene(seed=0)
    emb = (torch.rand((16, 1 << log2_T, 2), generator=torch.Generator().manual_seed(log2_T)) - 0.5) * 0.2
    m = R.RefModel(sc.grid, emb.numpy(), **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    o, d = torch.from_numpy(g["rays_o"]), torch.from_numpy(g["rays_d"])
    zs = tuple(torch.from_numpy(g[k]) for k in ("z_vals", "dists", "mid"))
    rr = torch.from_numpy(np.random.default_rng(5).normal(size=(64, 3)).astype(np.float32))
    out = m.render(o, d, None, None, zs=zs)
    (out["rgb"] * rr).sum().backward()
    cfg = Fn.RenderConfig(BOX6, m.level_res)
    ed = emb.to(DEV).requires_grad_(True)
    s = torch.clamp(torch.exp(torch.tensor(0.3, device=DEV) * 10.0), min=1e-4, max=300.0)
    outs = Fn.RenderFunction.apply(torch.from_numpy(sc.grid).to(DEV), ed, _mlp_flat(sc).to(DEV), s, o.to(DEV),
                                   d.to(DEV), zs[2].to(DEV), zs[1].to(DEV), cfg, False, None)
    (outs[0] * rr.to(DEV)).sum().backward()
    ref = m.embeddings.grad.numpy()
    got = ed.grad.cpu().numpy()
    assert np.array_equal(got != 0, ref != 0)
    assert rel_err(got, ref) < 1e-5


def test_render_backward_hash_gradient_split_pairs(Fn):
    """Every sample at x = 0.75384 (x01 = 0.87692): at levels 14 and 15 (res 4670, 7006)
    its cells sit at x = 4095 and 6143, whose corner pairs straddle two 2048-entry bins, so
    each sample gives 8 records per level and a chunk of 256 samples 2048, past the write
    pass's 1024 staged slots (the rest are stored directly).  The hash-table gradient still
    equals the oracle's autograd."""
    sc = scenes.config1_scene(seed=0)
    m = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    assert [int(m.level_res[l] * 0.87692) for l in (14, 15)] == [4095, 6143]
    n = 64
    zr = torch.linspace(-0.3, 0.3, n)
    o = torch.stack([torch.full((n,), 0.75384), torch.full((n,), -0.6), zr], 1).float()
    d = torch.tensor([[0.0, 1.0, 0.0]]).expand(n, 3).contiguous()
    z = torch.linspace(0.05, 1.15, 129)[None, :].expand(n, 129).contiguous()
    mid = (0.5 * (z[:, 1:] + z[:, :-1])).contiguous()
    dists = (z[:, 1:] - z[:, :-1]).contiguous()
    rr = torch.from_numpy(np.random.default_rng(8).normal(size=(n, 3)).astype(np.float32))
    out = m.render(o, d, None, None, zs=(z[:, :128].contiguous(), dists, mid))
    (out["rgb"] * rr).sum().backward()
    ref = m.embeddings.grad.numpy()
    assert np.count_nonzero(ref[14]) > 0 and np.count_nonzero(ref[15]) > 0
    cfg = Fn.RenderConfig(BOX6, m.level_res)
    ed = torch.from_numpy(sc.embeddings).to(DEV).requires_grad_(True)
    s = torch.clamp(torch.exp(torch.tensor(0.3, device=DEV) * 10.0), min=1e-4, max=300.0)
    outs = Fn.RenderFunction.apply(torch.from_numpy(sc.grid).to(DEV), ed, _mlp_flat(sc).to(DEV), s, o.to(DEV),
                                   d.to(DEV), mid.to(DEV), dists.to(DEV), cfg, False, None)
    (outs[0] * rr.to(DEV)).sum().backward()
    got = ed.grad.cpu().numpy()
    assert np.array_equal(got != 0, ref != 0)
    assert rel_err(got, ref) < 1e-5
    for l in (14, 15):
        assert rel_err(got[l], ref[l]) < 1e-5, l


@pytest.mark.parametrize("binned", [False, True], ids=["atomic_grid", "binned_grid"])
@pytest.mark.parametrize("name", ["step_c1.npz", "step_c2.npz"])
def test_fused_train_step_matches_reference(name, binned):
    """The fused step against the reference's own step (golden vectors), with the grid
    gradient from the atomic scatter and from the voxel-binned passes."""
    from voxneus_amd.training import FusedTrainStep, StepConfig
    g = load_golden(name)
    sc = _golden_scene(name)
    st = FusedTrainStep(torch.from_numpy(sc.grid).to(DEV), torch.from_numpy(sc.embeddings).to(DEV),
                        _mlp_flat(sc).to(DEV), torch.tensor(0.3, device=DEV),
                        R.hash_level_resolutions(16, 16, sc.hash_max_resolution),
                        cfg=StepConfig(w_eik=0.1, w_curv=5e-4, binned_grid_gradient=binned))
    out = st.step(torch.from_numpy(g["rays_o"]).to(DEV), torch.from_numpy(g["rays_d"]).to(DEV),
                  torch.from_numpy(g["target"]).to(DEV))
    assert np.array_equal(out["z_vals"].cpu().numpy(), g["z_vals"])
    assert np.array_equal(st.active_linear().cpu().numpy(), g["vr_lin"])
    assert abs(out["loss"].item() - float(g["loss_rgb"])) <= 1e-6 * abs(float(g["loss_rgb"]))
    le, lc = out["reg_losses"].tolist()
    assert abs(le - float(g["loss_eik"])) <= 1e-5 * abs(float(g["loss_eik"]))
    assert abs(lc - float(g["loss_curv"])) <= 1e-5 * abs(float(g["loss_curv"]))
    ref = np.zeros(st.grad_grid.numel(), np.float32)
    ref[g["grid_grad_idx"]] = g["grid_grad_val"]
    assert rel_err(st.grad_grid.cpu().numpy().reshape(-1), ref) < 1e-5
    eg = st.grad_table.cpu().numpy().reshape(-1)
    assert int(np.count_nonzero(eg)) == int(g["emb_grad_nnz"])
    assert rel_err(eg[g["emb_grad_idx"]], g["emb_grad_val"]) < 1e-5
    ref_mlp = np.concatenate([g["dW0"].reshape(-1), g["db0"], g["dW1"].reshape(-1), g["db1"]])
    assert rel_err(st.grad_mlp.cpu().numpy(), ref_mlp) < 1e-5
    assert bool(out["amplified"].item()) == bool(g["amplified"])
    assert abs(st.grad_v.item() - float(g["dv_post"])) <= 1e-5 * abs(float(g["dv_post"]))


def test_vr_compaction_and_regularisers_match_oracle(Fn):
    rng = np.random.default_rng(4)
    res = 48
    data = rng.normal(size=(res, res, res)).astype(np.float32)
    corners = rng.integers(0, res, size=(20000, 8, 3))
    verts, lin = R.extract_active_vertices(torch.from_numpy(corners), torch.tensor([res] * 3))
    bm = torch.zeros(Fn.vr_bitmask_words((res, res, res)), device=DEV, dtype=torch.int32)
    Fn.vr_mark_corners(torch.from_numpy(corners).to(DEV), (res, res, res), bm)
    lst, counts = Fn.vr_compact(bm, (res, res, res), BOX6, relax_radius=1.2)
    n_all, n_in = counts.tolist()
    got = (lst[:n_all] & 0x7FFFFFFF).cpu().numpy()
    assert np.array_equal(got, lin.numpy().astype(np.int32))
    vs = torch.tensor([2.0 / (res - 1)] * 3)
    world = torch.tensor(BOX[0]) + verts.float() * vs
    inside = torch.linalg.norm(world, dim=-1) < 1.2
    assert n_in == int(inside.sum())
    keep = verts[inside] if inside.any() else verts
    le, ge = R.eikonal_loss_and_grad(torch.from_numpy(data), keep, vs)
    lc, gc = R.curvature_loss_and_grad(torch.from_numpy(data), keep, vs)
    gg = torch.zeros(res, res, res, device=DEV)
    losses = Fn.reg_eik_curv(torch.from_numpy(data).to(DEV), vs.tolist(), lst, counts, 0.1, 5e-4, gg)
    ref = (ge * 0.1 + gc * 5e-4).numpy()
    assert rel_err(gg.cpu().numpy(), ref) < 1e-5
    assert abs(losses[0].item() - le.item()) <= 1e-5 * le.item()
    assert abs(losses[1].item() - lc.item()) <= 1e-5 * lc.item()


@pytest.mark.parametrize("relax", [1.2, 0.3])
def test_bitmask_regularisers_equal_list_form(Fn, relax):
    """The tiled pass over the V_R bitmask gives the list form's counts, losses and
    gradient (FP32 sums in another order); relax 0.3 leaves part of V_R outside."""
    rng = np.random.default_rng(6)
    res = 40
    data = torch.from_numpy(rng.normal(size=(res, res, res)).astype(np.float32)).to(DEV)
    corners = torch.from_numpy(rng.integers(0, res, size=(30000, 8, 3))).to(DEV)
    bm = torch.zeros(Fn.vr_bitmask_words((res, res, res)), device=DEV, dtype=torch.int32)
    Fn.vr_mark_corners(corners, (res, res, res), bm)
    lst, counts = Fn.vr_compact(bm, (res, res, res), BOX6, relax_radius=relax)
    vs = Fn.voxel_size_of(BOX6, (res, res, res))
    g_list = torch.full((res, res, res), 0.25, device=DEV)
    l_list = Fn.reg_eik_curv(data, vs, lst, counts, 0.1, 5e-4, g_list)
    g_vr = torch.full((res, res, res), 0.25, device=DEV)
    l_vr, c_vr = Fn.reg_eik_curv_vr(data, vs, BOX6, relax, bm, 0.1, 5e-4, g_vr)
    assert torch.equal(c_vr, counts)
    assert rel_err(l_vr.cpu().numpy(), l_list.cpu().numpy()) < 1e-12
    assert rel_err((g_vr - 0.25).cpu().numpy(), (g_list - 0.25).cpu().numpy()) < 1e-6


def test_reference_shaped_regularisers(Fn):
    """Pattern of test_cuda_parity.py:72-126 (in-place accumulation == separate gradient)."""
    torch.manual_seed(42)
    res = 32
    grid = torch.randn(res, res, res)
    act = torch.randint(1, res - 1, (100, 3))
    vs = torch.tensor([2.0 / (res - 1)] * 3)
    _, ge = R.eikonal_loss_and_grad(grid, act, vs)
    _, gc = R.curvature_loss_and_grad(grid, act, vs)
    a = torch.zeros(res, res, res, device=DEV)
    Fn.eikonal_grad_ref(grid.to(DEV), act.to(DEV), a, vs.tolist(), 1.0)
    assert (a.cpu() - ge).abs().max().item() < 1e-5
    b = torch.zeros(res, res, res, device=DEV)
    Fn.curvature_grad_ref(grid.to(DEV), act.to(DEV), b, vs.tolist(), 1.0)
    assert (b.cpu() - gc).abs().max().item() <= 1e-5 * max(1.0, gc.abs().max().item())


def test_s_amplify_device_side(Fn):
    g = torch.tensor([-0.25], device=DEV)
    f = torch.zeros(1, device=DEV, dtype=torch.int32)
    Fn.s_amplify(g, 5.0, f)
    assert g.item() == -1.25 and f.item() == 1
    Fn.s_amplify(g.neg_(), 5.0, f)
    assert g.item() == 1.25 and f.item() == 0


@pytest.mark.parametrize("shape,lbox", [((17, 17, 17), BOX6), ((64, 64, 64), BOX6),
                                        ((23, 19, 31), (-1.3, -0.7, -1.1, 1.2, 0.9, 1.4))])
def test_lattice_query_equals_point_query(Fn, shape, lbox):
    """The lattice kernel (stencil rows staged in LDS) equals the point query at the same
    points bit for bit: a lattice between the grid's vertices, one on them, and one whose
    box reaches past the grid's (clamped border cells)."""
    sc = scenes.config1_scene(seed=0)
    grid = torch.from_numpy(sc.grid).to(DEV)
    nx, ny, nz = shape
    sdf, nrm = Fn.lattice_query(grid, BOX, nx, ny, nz, lbox)
    ax = []
    for a, n in enumerate(shape):
        lo, hi = np.float32(lbox[a]), np.float32(lbox[a + 3])
        step = np.float32((hi - lo) / np.float32(n - 1))
        ax.append(torch.from_numpy(lo + np.arange(n, dtype=np.float32) * step))
    P = torch.stack(torch.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3).to(DEV)
    o = Fn.grid_query_raw(grid, BOX, P, True, False, True)
    assert torch.equal(sdf.reshape(-1), o["sdf"]) and torch.equal(nrm.reshape(-1, 3), o["normal"])


# ----------------------------------------------------------------------------------------
# BF16 form: bf16 hash table + matrix-core MLP (bf16 operands, FP32 accumulation)
# ----------------------------------------------------------------------------------------

def _bf16_render_pair(Fn, name, n_rays=None):
    """The BF16-form render + backward on the GPU and in the oracle's BF16 emulation,
    same z values and a fixed random upstream gradient."""
    g = load_golden(name)
    sc = _golden_scene(name)
    emb_b = torch.from_numpy(sc.embeddings).to(torch.bfloat16)
    m = R.RefModel(sc.grid, emb_b.float().numpy(), **sc.mlp, hash_max_resolution=sc.hash_max_resolution,
                   mlp_form="bf16")
    sl = slice(0, n_rays)
    o, d = torch.from_numpy(g["rays_o"][sl]), torch.from_numpy(g["rays_d"][sl])
    zs = tuple(torch.from_numpy(g[k][sl]) for k in ("z_vals", "dists", "mid"))
    R_ = o.shape[0]
    rng = np.random.default_rng(5)
    rr = [torch.from_numpy(rng.normal(size=sh).astype(np.float32)) for sh in ((R_, 3), (R_, 1), (R_, 3), (R_, 1))]
    out = m.render(o, d, None, None, zs=zs)
    L = (out["rgb"] * rr[0]).sum() + (out["depth"] * rr[1]).sum() + (out["normal"] * rr[2]).sum() + (
        out["opacity"] * rr[3]).sum()
    L.backward()
    gt = torch.zeros(emb_b.shape, device=DEV, dtype=torch.float32)
    cfg = Fn.RenderConfig(BOX6, m.level_res, mlp_mode=Fn.N.VXN_MLP_BF16_MFMA, table_grad_out=gt)
    grid = torch.from_numpy(sc.grid).to(DEV).requires_grad_(True)
    mlp = _mlp_flat(sc).to(DEV).requires_grad_(True)
    v = torch.tensor(0.3, device=DEV, requires_grad=True)
    s = torch.clamp(torch.exp(v * 10.0), min=1e-4, max=300.0)
    outs = Fn.RenderFunction.apply(grid, emb_b.to(DEV), mlp, s, o.to(DEV), d.to(DEV), zs[2].to(DEV), zs[1].to(DEV),
                                   cfg, False, None)
    Lg = sum((outs[i] * rr[k].to(DEV)).sum() for k, i in enumerate((0, 1, 2, 3)))
    Lg.backward()
    ref_mlp = torch.cat([m.W0.grad.reshape(-1), m.b0.grad, m.W1.grad.reshape(-1), m.b1.grad]).numpy()
    return {"out": out, "outs": [t.detach().cpu() for t in outs[:4]], "m": m, "grid_grad": grid.grad.cpu().numpy(),
            "table_grad": gt.cpu().numpy(), "mlp_grad": mlp.grad.cpu().numpy(), "ref_mlp": ref_mlp,
            "v_grad": v.grad.item(), "sc": sc, "o": o, "d": d, "zs": zs}


@pytest.mark.parametrize("name", ["step_c1.npz", "step_c2.npz"])
def test_bf16_form_matches_emulation(Fn, name):
    """Forward and backward of the BF16 form vs the oracle's emulation of the same
    rounding points (FP32 sums differ only in order inside the matrix cores)."""
    r = _bf16_render_pair(Fn, name)
    out, m = r["out"], r["m"]
    for k, i in (("rgb", 0), ("depth", 1), ("normal", 2), ("opacity", 3)):
        err = float((r["outs"][i] - out[k].detach()).abs().max())
        assert err < 1e-4, (k, err)
    e_grid = rel_err(r["grid_grad"], m.grid.grad.numpy()[0, 0])
    e_tab = rel_err(r["table_grad"], m.embeddings.grad.numpy())
    e_mlp = rel_err(r["mlp_grad"], r["ref_mlp"])
    e_v = abs(r["v_grad"] - m.variance.grad.item()) / max(1e-3, abs(m.variance.grad.item()))
    print(f"bf16 vs emulation: grid {e_grid:.2e} table {e_tab:.2e} mlp {e_mlp:.2e} v {e_v:.2e}")
    assert e_grid < 1e-4 and e_tab < 1e-4 and e_mlp < 1e-4 and e_v < 1e-4


def test_bf16_form_rgb_error_vs_fp64(Fn):
    """Max-abs RGB of the BF16 form against the FP64 oracle with the FP32 parameters."""
    r = _bf16_render_pair(Fn, "step_c2.npz")
    sc = r["sc"]
    m64 = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution,
                     dtype=torch.float64)
    with torch.no_grad():
        ref = m64.render(r["o"].double(), r["d"].double(), None, None, zs=tuple(z.double() for z in r["zs"]))["rgb"]
    err = float((r["outs"][0].double() - ref).abs().max())
    print(f"bf16 form rgb max-abs err vs fp64: {err:.3e}")
    assert err <= 5e-3
