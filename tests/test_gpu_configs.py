"""GPU tests at the five workload configurations, at their real sizes.

* config 1: 64^3 sphere grid, 16 x 2^15 hash, all 512 rays of one step (FP32 form);
* config 2: 128^3 multi-primitive grid, 16 x 2^17 hash, all 2048 rays of one step;
* config 3: 256^3 multi-primitive grid, 16 x 2^19 hash, the 8192 hemisphere rays x 128
  samples of one training step, FP32 and BF16 forms;
* config 4: 320^3 grid with 16 x 2^19 hash: one 800x600 view through the serving loop and
  the SDF + normal query over the 320^3 lattice;
* config 5: 24 resident scenes (grids 96-256^3, tables 2^17-2^19), one render request
  per image size class and three query batches (10^4 - 10^6 points), mixed.

Integer results (z values from the deterministic sampler, V_R) are compared bit for bit
with the CPU oracle on the same inputs; FP32 floats within 1e-5 of the oracle's largest
magnitude (losses 1e-6); the BF16 form states its RGB / depth / normal / gradient error
against the FP64 oracle and bounds it (RGB <= 5e-3).  Where the oracle would take minutes
on the full batch (configs 3-5) it runs on a seeded subset of the rays or points, and the
full batch is checked by properties the oracle can afford (V_R of the same z values,
finite outputs, the two forms' agreement on z).
"""

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R
from voxneus_amd import _native as N
from voxneus_amd import functional as Fn
from voxneus_amd import scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
BOX6 = (-1.0, -1.0, -1.0, 1.0, 1.0, 1.0)
W_EIK, W_CURV = 0.1, 5e-4


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def mlp_flat(mlp):
    return torch.cat([torch.as_tensor(mlp[k]).reshape(-1) for k in ("W0", "b0", "W1", "b1")])


def mlp_split(flat):
    f = flat.detach().cpu().float()
    return {"W0": f[:64 * 51].reshape(64, 51).numpy(), "b0": f[3264:3328].numpy(),
            "W1": f[3328:3520].reshape(3, 64).numpy(), "b1": f[3520:3523].numpy()}


def fused_step(sc, o, d, t, mlp_mode=N.VXN_MLP_FP32):
    from voxneus_amd.training import FusedTrainStep, StepConfig

    table = torch.from_numpy(sc.embeddings).to(DEV)
    if mlp_mode == N.VXN_MLP_BF16_MFMA:
        table = table.to(torch.bfloat16)
    st = FusedTrainStep(torch.from_numpy(sc.grid).to(DEV), table, mlp_flat(sc.mlp).to(DEV),
                        torch.tensor(0.3, device=DEV), R.hash_level_resolutions(16, 16, sc.hash_max_resolution),
                        cfg=StepConfig(w_eik=W_EIK, w_curv=W_CURV, mlp_mode=mlp_mode))
    out = st.step(torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV), torch.from_numpy(t).to(DEV))
    torch.cuda.synchronize()
    return st, out


def oracle_vr(grid_res, o, d, mid, relax=1.2):
    """The oracle's V_R (trainer.py:260-266) of samples at rays_o + rays_d * mid: sorted
    linear indices and the relaxed-sphere filtered list."""
    o, d, mid = (torch.as_tensor(a) for a in (o, d, mid))
    box = R.GridBox((grid_res,) * 3)
    pts = o.unsqueeze(1) + d.unsqueeze(1) * mid.unsqueeze(-1)
    coords = R.world_to_grid_index(pts, box.xyz_min, box.xyz_max, box.res)
    corners, _ = R.get_voxel_corners_and_weights(coords, box.res)
    verts, lin = R.extract_active_vertices(corners, box.res)
    world = box.xyz_min + verts.float() * box.voxel_size
    inside = torch.linalg.norm(world, dim=-1) < relax
    return lin, (lin[inside] if inside.any() else lin), int(inside.sum())


def mlp_block_errors(got, ref):
    """Largest error of each parameter block of the MLP gradient, relative to the largest
    magnitude of the whole gradient."""
    scale = np.abs(ref).max()
    w0g, w0r = got[:3264].reshape(64, 51), ref[:3264].reshape(64, 51)
    out = {"W0 hash": np.abs(w0g[:, :32] - w0r[:, :32]).max(), "W0 sh": np.abs(w0g[:, 32:48] - w0r[:, 32:48]).max(),
           "W0 normal": np.abs(w0g[:, 48:] - w0r[:, 48:]).max(),
           "b0": np.abs(got[3264:3328] - ref[3264:3328]).max(), "W1": np.abs(got[3328:3520] - ref[3328:3520]).max(),
           "b1": np.abs(got[3520:] - ref[3520:]).max()}
    return {k: float(v / scale) for k, v in out.items()}


def mlp_grad_fp64(sc, o, d, t, zs):
    """FP64 gradient of the Huber colour loss w.r.t. the MLP on the same z values."""
    m = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution,
                   dtype=torch.float64)
    dt = torch.float64
    out = m.render(torch.from_numpy(o).to(dt), torch.from_numpy(d).to(dt), None, None,
                   zs=tuple(z.to(dt) for z in zs))
    R.huber(out["rgb"], torch.from_numpy(t).to(dt), 0.1).backward()
    return torch.cat([m.W0.grad.reshape(-1), m.b0.grad, m.W1.grad.reshape(-1), m.b1.grad]).numpy()


def check_step_vs_oracle(st, out, m, ref, lin_tol=1e-5, mlp64=None):
    """The fused step's outputs against the oracle's train_step on the same rays.  With
    ``mlp64`` (the FP64 MLP gradient) an MLP gradient further than ``lin_tol`` from the FP32
    oracle's must be at least as close to FP64 as the FP32 oracle is (both are FP32 sums
    over every sample of the batch; their distance is the two summations' rounding)."""
    assert np.array_equal(out["z_vals"].cpu().numpy(), ref["render"]["z_vals"].numpy())
    assert np.array_equal(st.active_linear().cpu().numpy(), ref["active_linear"].numpy().astype(np.int32))
    lr = float(ref["loss"])
    assert abs(out["loss"].item() - lr) <= 1e-6 * abs(lr)
    le, lc = out["reg_losses"].tolist()
    assert abs(le - float(ref["loss_eik"])) <= 1e-5 * abs(float(ref["loss_eik"]))
    assert abs(lc - float(ref["loss_curv"])) <= 1e-5 * abs(float(ref["loss_curv"]))
    e_grid = rel_err(st.grad_grid.cpu().numpy(), m.grid.grad.numpy()[0, 0])
    gt, rt = st.grad_table.cpu().numpy(), m.embeddings.grad.numpy()
    e_tab = rel_err(gt, rt)
    ref_mlp = torch.cat([m.W0.grad.reshape(-1), m.b0.grad, m.W1.grad.reshape(-1), m.b1.grad]).numpy()
    e_mlp = rel_err(st.grad_mlp.cpu().numpy(), ref_mlp)
    dv = float(m.variance.grad)
    e_v = abs(st.grad_v.item() - dv) / abs(dv)
    print(f"vs oracle: loss {lr:.6e} grid {e_grid:.2e} hash {e_tab:.2e} mlp {e_mlp:.2e} v {e_v:.2e}")
    gm = st.grad_mlp.cpu().numpy()
    print("  mlp gradient blocks vs oracle: " + " ".join(f"{k} {v:.1e}" for k, v in mlp_block_errors(gm, ref_mlp).items()))
    if mlp64 is not None:
        e_gpu64, e_ora64 = rel_err(gm, mlp64), rel_err(ref_mlp, mlp64)
        print(f"  mlp gradient vs fp64: hip {e_gpu64:.2e} oracle {e_ora64:.2e}; hip blocks "
              + " ".join(f"{k} {v:.1e}" for k, v in mlp_block_errors(gm, mlp64).items()))
        if e_mlp >= lin_tol:
            assert e_gpu64 <= max(e_ora64, lin_tol), (e_gpu64, e_ora64)
            e_mlp = 0.0
    assert np.array_equal(gt != 0, rt != 0)
    assert e_grid < lin_tol and e_tab < lin_tol and e_mlp < lin_tol and e_v < lin_tol
    assert bool(out["amplified"].item()) == bool(ref["amplified"])


# ----------------------------------------------------------------------------------------
# Configs 1 and 2: the whole batch of one step against the oracle's train step
# ----------------------------------------------------------------------------------------

@pytest.mark.parametrize("cfg", ["config1", "config2"])
def test_full_batch_step_matches_oracle(cfg):
    """Config 1 (64^3 sphere, 2^15, 512 sphere rays) and config 2 (128^3 multi-primitive,
    2^17, 2048 hemisphere rays), deterministic sampler, every ray of the step."""
    if cfg == "config1":
        sc = scenes.config1_scene(seed=0)
        o, d, t = scenes.config_rays("sphere", 512, seed=11)
    else:
        sc = scenes.make_scene("multi", 128, 17, seed=2)
        o, d, t = scenes.config_rays("multi", 2048, seed=12, shape=sc.shape)
    print(f"{cfg}: grid {sc.grid_res}^3, table 2^{sc.meta['log2_T']}, {o.shape[0]} rays", flush=True)
    st, out = fused_step(sc, o, d, t)
    m = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    ref = m.train_step(torch.from_numpy(o), torch.from_numpy(d), torch.from_numpy(t), w_eik=W_EIK, w_curv=W_CURV)
    r = ref["render"]
    check_step_vs_oracle(st, out, m, ref, mlp64=mlp_grad_fp64(sc, o, d, t, (r["z_vals"], r["dists"],
                                                                             r["mid_z_vals"])))


# ----------------------------------------------------------------------------------------
# Config 3: 256^3, 2^19, 8192 rays x 128 samples
# ----------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def c3():
    sc = scenes.make_scene("multi", 256, 19, seed=3)
    o, d = scenes.hemisphere_ray_batch(8192, seed=100)
    t = scenes.procedural_rgb(sc.shape, o, d)
    print("config 3: 256^3 grid, 16 x 2^19 hash, 8192 rays", flush=True)
    st, out = fused_step(sc, o, d, t)
    full = {"z": out["z_vals"].cpu().numpy(), "mid": out["mid_z_vals"].cpu().numpy(),
            "dists": out["dists"].cpu().numpy(),
            "counts": out["vr_counts"].cpu().numpy(), "loss": out["loss"].item(),
            "active": st.active_linear().cpu().numpy(),
            "finite": bool(torch.isfinite(st.grad_grid).all() and torch.isfinite(st.grad_table).all()
                           and torch.isfinite(st.grad_mlp).all() and torch.isfinite(st.grad_v).all())}
    del st, out
    torch.cuda.empty_cache()
    idx = np.sort(np.random.default_rng(31).choice(8192, size=256, replace=False))
    return {"sc": sc, "o": o, "d": d, "t": t, "full": full, "idx": idx}


def test_config3_fp32_full_step_vr_and_sampler_vs_oracle(c3):
    """The 8192-ray FP32 step: the oracle's sampler gives the same z on 256 seeded rays,
    and the oracle's V_R of the step's own sample positions (all 1,048,576 samples) equals
    the step's, before and after the relaxed-sphere filter."""
    sc, o, d, full, idx = c3["sc"], c3["o"], c3["d"], c3["full"], c3["idx"]
    assert full["finite"] and np.isfinite(full["loss"])
    m = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    os_, ds_ = torch.from_numpy(o[idx]), torch.from_numpy(d[idx])
    near, far = R.near_far_from_sphere(os_, ds_)
    z, dists, mid = m.sample(os_, ds_, near, far)
    assert np.array_equal(full["z"][idx], z.numpy())
    assert np.array_equal(full["mid"][idx], mid.numpy())
    lin, lin_f, n_in = oracle_vr(256, o, d, full["mid"])
    print(f"config 3 V_R: {len(lin)} vertices, {n_in} inside the relaxed sphere", flush=True)
    assert full["counts"].tolist() == [len(lin), n_in]
    assert np.array_equal(full["active"], lin_f.numpy().astype(np.int32))


def test_config3_fp32_step_subset_matches_oracle(c3):
    """The fused FP32 step on 64 of the 8192 rays at the full grid and table: z exact, V_R
    exact, loss within 1e-6, grid / hash / MLP / v gradients within 1e-5 of the oracle's
    train_step on the same rays."""
    sc, idx = c3["sc"], c3["idx"][:64]
    o, d, t = (c3[k][idx] for k in ("o", "d", "t"))
    st, out = fused_step(sc, o, d, t)
    m = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    ref = m.train_step(torch.from_numpy(o), torch.from_numpy(d), torch.from_numpy(t), w_eik=W_EIK, w_curv=W_CURV)
    assert np.array_equal(out["z_vals"].cpu().numpy(), c3["full"]["z"][idx])
    check_step_vs_oracle(st, out, m, ref)


def test_config3_bf16_full_step(c3):
    """The 8192-ray BF16 step: the sampler is FP32 in both forms, so z and V_R equal the
    FP32 step's bit for bit; the loss, every gradient and the losses are finite, and the
    colour loss is within BF16 rounding of the FP32 step's."""
    from voxneus_amd.training import FusedTrainStep  # noqa: F401

    sc = c3["sc"]
    st, out = fused_step(sc, c3["o"], c3["d"], c3["t"], N.VXN_MLP_BF16_MFMA)
    full = c3["full"]
    assert np.array_equal(out["z_vals"].cpu().numpy(), full["z"])
    assert out["vr_counts"].cpu().numpy().tolist() == full["counts"].tolist()
    for g in (st.grad_grid, st.grad_table, st.grad_mlp, st.grad_v, out["reg_losses"]):
        assert bool(torch.isfinite(g).all())
    lb = out["loss"].item()
    print(f"config 3 colour loss: fp32 {full['loss']:.6e} bf16 {lb:.6e}")
    assert abs(lb - full["loss"]) <= 5e-3 * abs(full["loss"])


def test_config3_bf16_error_vs_fp64(c3):
    """The BF16 form (bf16 table, matrix-core MLP) on 256 of the 8192 rays with the step's
    z values: RGB, depth, normal and opacity, and the Huber loss's grid / hash / MLP / v
    gradients, against the FP64 oracle with the FP32 parameters, stated and bounded; and
    against the oracle's BF16 emulation (same rounding points) within 1e-4."""
    sc, idx = c3["sc"], c3["idx"]
    o, d, t = (torch.from_numpy(c3[k][idx]) for k in ("o", "d", "t"))
    zs = tuple(torch.from_numpy(c3["full"][k][idx]) for k in ("z", "mid"))
    dists = torch.from_numpy(c3["full"]["dists"][idx])
    res = R.hash_level_resolutions(16, 16, sc.hash_max_resolution)
    emb_b = torch.from_numpy(sc.embeddings).to(torch.bfloat16)
    gt = torch.zeros(emb_b.shape, device=DEV)
    cfg = Fn.RenderConfig(BOX6, res, mlp_mode=N.VXN_MLP_BF16_MFMA, table_grad_out=gt)
    grid = torch.from_numpy(sc.grid).to(DEV).requires_grad_(True)
    mlp = mlp_flat(sc.mlp).to(DEV).requires_grad_(True)
    v = torch.tensor(0.3, device=DEV, requires_grad=True)
    s = torch.clamp(torch.exp(v * 10.0), min=1e-4, max=300.0)
    outs = Fn.RenderFunction.apply(grid, emb_b.to(DEV), mlp, s, o.to(DEV), d.to(DEV), zs[1].to(DEV),
                                   dists.to(DEV), cfg, False, None)
    R.huber(outs[0], t.to(DEV), 0.1).backward()
    hip = {"rgb": outs[0], "depth": outs[1], "normal": outs[2], "opacity": outs[3]}
    hip = {k: v_.detach().cpu().double() for k, v_ in hip.items()}
    g_hip = [grid.grad.cpu().numpy(), gt.cpu().numpy(), mlp.grad.cpu().numpy(), np.array([v.grad.item()])]

    def oracle(dtype, form, table):
        m = R.RefModel(sc.grid, table, **sc.mlp, hash_max_resolution=sc.hash_max_resolution, dtype=dtype,
                       mlp_form=form)
        out = m.render(o.to(dtype), d.to(dtype), None, None, zs=(zs[0].to(dtype), dists.to(dtype), zs[1].to(dtype)))
        R.huber(out["rgb"], t.to(dtype), 0.1).backward()
        g = [m.grid.grad.numpy()[0, 0], m.embeddings.grad.numpy(),
             torch.cat([m.W0.grad.reshape(-1), m.b0.grad, m.W1.grad.reshape(-1), m.b1.grad]).numpy(),
             np.array([m.variance.grad.item()])]
        return out, g

    o64, g64 = oracle(torch.float64, "fp32", sc.embeddings)
    oem, gem = oracle(torch.float32, "bf16", emb_b.float().numpy())
    err = {k: float((hip[k] - o64[k].detach()).abs().max()) for k in hip}
    # the composite normal is normalize(sum w n); its rounding error grows as 1 / |sum w n|,
    # so it is stated on the rays that hit a surface (|sum w n| >= 0.1)
    nsum = (o64["weights"].detach()[..., None] * o64["normal_samples"].detach()).sum(1).norm(dim=-1)
    hit = nsum >= 0.1
    err["normal_hit"] = float((hip["normal"] - o64["normal"].detach()).abs().max(-1).values[hit].max())
    names = ("grid", "hash", "mlp", "v")
    g_max = {n: rel_err(a, b) for n, a, b in zip(names, g_hip, g64)}
    # bf16 rounding flips a few ReLU masks and Huber branches, which moves single entries by
    # a whole term: the gradient error is bounded as ||g - g64|| / ||g64|| and the largest
    # entry error is stated beside it
    g_err = {n: float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
             for n, a, b in zip(names, g_hip, g64)}
    g_emu = {n: rel_err(a, b) for n, a, b in zip(names, g_hip, gem)}
    print("config 3 bf16 vs fp64: " + " ".join(f"{k} {v_:.2e}" for k, v_ in err.items()) + " | gradients (norm) "
          + " ".join(f"{k} {v_:.2e}" for k, v_ in g_err.items()) + " | (max entry) "
          + " ".join(f"{k} {v_:.2e}" for k, v_ in g_max.items()) + " | vs emulation "
          + " ".join(f"{k} {v_:.2e}" for k, v_ in g_emu.items()))
    assert err["rgb"] <= 5e-3
    assert err["depth"] <= 1e-4 and err["opacity"] <= 1e-4 and err["normal_hit"] <= 1e-3
    # the CPU emulation of this batch: norm grid 4.5e-3, hash 1.2e-2, mlp 1.8e-3, v 8.5e-5
    assert g_err["grid"] < 2e-2 and g_err["hash"] < 5e-2 and g_err["mlp"] < 1e-2 and g_err["v"] < 1e-3
    assert max(g_emu.values()) < 1e-4
    for k in ("rgb", "depth", "opacity"):
        assert float((hip[k] - oem[k].detach().double()).abs().max()) < 1e-4, k


# ----------------------------------------------------------------------------------------
# Config 4: 320^3 grid, 2^19 hash: an 800x600 view and the 320^3 lattice query
# ----------------------------------------------------------------------------------------

def oracle_model(rs, dtype=torch.float32):
    """The oracle's model of a resident serving scene (its grid, table, MLP and s)."""
    mlp = mlp_split(rs.mlp)
    m = R.RefModel(rs.grid.cpu().numpy(), rs.table.float().cpu().numpy(), **mlp, dtype=dtype,
                   occupancy_mask=rs.occ_mask.cpu(),
                   gradient_mode="analytical" if rs.gradient_mode == N.gradient_mode_code("analytical")
                   else "interpolated")
    m.level_res = list(rs.level_res)
    if dtype == torch.float32:
        assert float(np.float32(rs.s)) == float(m.s.detach())
    return m


def oracle_render_rows(rs, o, d, rows, fp64=False):
    """render.py:35-54 on the chosen rows: near/far from the unit sphere, deterministic
    sampler in FP32, then the render in FP32 or in FP64 on the same z."""
    o, d = o[rows].cpu(), d[rows].cpu()
    m = oracle_model(rs)
    near, far = R.near_far_from_sphere(o, d)
    zs = m.sample(o, d, near, far)
    if fp64:
        m = oracle_model(rs, torch.float64)
        o, d, zs = o.double(), d.double(), tuple(z.double() for z in zs)
    with torch.no_grad():
        out = m.render(o, d, None, None, zs=zs)
    return out


def check_render(rs, req, res, mlp_mode, n_pix, seed):
    from voxneus_amd.serving import pinhole_rays

    o, d = pinhole_rays(req, DEV)
    rows = np.sort(np.random.default_rng(seed).choice(req.rows, size=n_pix, replace=False))
    rgb, depth = res.outputs["rgb"][rows].cpu().double(), res.outputs["depth"][rows].cpu().double()
    if mlp_mode == N.VXN_MLP_FP32:
        ref = oracle_render_rows(rs, o, d, rows)
        e_rgb = rel_err(rgb, ref["rgb"].double())
        e_d = rel_err(depth, ref["depth"].double())
        print(f"  render {req.W}x{req.H} on {rs.grid.shape[0]}^3: rgb {e_rgb:.2e} depth {e_d:.2e} vs FP32 oracle")
        assert e_rgb < 1e-5 and e_d < 1e-5
    else:
        ref = oracle_render_rows(rs, o, d, rows, fp64=True)
        e_rgb = float((rgb - ref["rgb"]).abs().max())
        e_d = float((depth - ref["depth"]).abs().max())
        print(f"  render {req.W}x{req.H} on {rs.grid.shape[0]}^3 (bf16): rgb max-abs {e_rgb:.2e} depth {e_d:.2e} "
              f"vs FP64 oracle")
        assert e_rgb <= 5e-3 and e_d <= 1e-4


def check_query(rs, pts, res, n, seed):
    rows = np.sort(np.random.default_rng(seed).choice(pts.shape[0], size=min(n, pts.shape[0]), replace=False))
    p = pts[rows].cpu()
    m = oracle_model(rs)
    with torch.no_grad():
        sdf = R.query_sdf(m.grid, m.box, p)
        _, nrm, _, _ = R.query_normals(m.grid, m.box, p, m.gradient_mode)
    assert torch.equal(res.outputs["sdf"][rows].cpu(), sdf.reshape(-1, 1))
    e = float((res.outputs["normal"][rows].cpu() - nrm).abs().max())
    print(f"  query of {pts.shape[0]} points on {rs.grid.shape[0]}^3: sdf exact, normal max-abs {e:.2e}")
    assert e <= 1e-6


@pytest.mark.parametrize("mlp_mode", [N.VXN_MLP_FP32, N.VXN_MLP_BF16_MFMA])
def test_config4_full_view_render(mlp_mode):
    """One 800x600 hemisphere view (f = 1200) of a 320^3 / 16 x 2^19 resident scene through
    the serving loop; 256 sampled pixels against the oracle's render (FP32 form within 1e-5,
    BF16 form within 5e-3 of FP64)."""
    from voxneus_amd.serving import RenderRequest, Server, make_resident_scene

    bf16 = mlp_mode == N.VXN_MLP_BF16_MFMA
    rs = make_resident_scene(320, 19, 4, DEV, table_dtype=torch.bfloat16 if bf16 else torch.float32)
    cams = scenes.hemisphere_cameras(49)
    req = RenderRequest(0, cams[17], 800, 600, 1200.0, 1200.0, 400.0, 300.0)
    srv = Server([rs], mlp_mode=mlp_mode, ray_capacity=1 << 17, point_capacity=1 << 10, ray_arena=800 * 600,
                 point_arena=1 << 10)
    res = srv.serve([req], realtime=False)[0]
    assert bool(torch.isfinite(res.outputs["rgb"]).all()) and bool(torch.isfinite(res.outputs["depth"]).all())
    print(f"config 4: 800x600 view, pixels with depth > 0: {float((res.outputs['depth'] > 0).float().mean()):.2f}")
    check_render(rs, req, res, mlp_mode, 256, seed=41)


def test_config4_lattice_query_vs_oracle():
    """SDF and Eq. 7 normal at every point of the 320^3 lattice over the grid's box; 10^4
    seeded lattice points against the oracle's query_sdf / query_normals (SDF bit-exact,
    normal within 1e-6)."""
    from voxneus_amd.serving import make_resident_scene

    rs = make_resident_scene(320, 19, 4, DEV)
    n = 320
    sdf, nrm = Fn.lattice_query(rs.grid, ((-1.0,) * 3, (1.0,) * 3), n, n, n, BOX6)
    assert bool(torch.isfinite(sdf).all()) and bool(torch.isfinite(nrm).all())
    rng = np.random.default_rng(43)
    ijk = rng.integers(0, n, size=(10000, 3))
    lo, hi = np.float32(-1.0), np.float32(1.0)
    step = np.float32((hi - lo) / np.float32(n - 1))
    pts = torch.from_numpy(lo + ijk.astype(np.float32) * step)
    m = oracle_model(rs)
    with torch.no_grad():
        rsdf = R.query_sdf(m.grid, m.box, pts).reshape(-1)
        _, rnrm, _, _ = R.query_normals(m.grid, m.box, pts)
    i, j, k = (torch.from_numpy(ijk[:, a]) for a in range(3))
    gs, gn = sdf.cpu()[i, j, k], nrm.cpu()[i, j, k]
    assert torch.equal(gs, rsdf)
    e = float((gn - rnrm).abs().max())
    print(f"config 4 lattice: 320^3 points, 10^4 checked: sdf exact, normal max-abs {e:.2e}")
    assert e <= 1e-6


# ----------------------------------------------------------------------------------------
# Config 5: 24 resident scenes
# ----------------------------------------------------------------------------------------

@pytest.mark.parametrize("mlp_mode", [N.VXN_MLP_FP32, N.VXN_MLP_BF16_MFMA])
def test_config5_serving_mix_vs_oracle_and_alone(mlp_mode):
    """24 resident scenes (grids {96,128,192,256}^3, tables 2^17-2^19), one render per image
    size class {64^2, 200x150, 400x300, 800x600} and queries of 10^4, 10^5 and 10^6 points,
    all admitted together: sampled pixels and points against the oracle, and every result
    bit-identical to the same request served alone."""
    from voxneus_amd.serving import QueryRequest, RenderRequest, Server, make_resident_scene

    bf16 = mlp_mode == N.VXN_MLP_BF16_MFMA
    grids, logs = (96, 128, 192, 256), (17, 18, 19)
    tdt = torch.bfloat16 if bf16 else torch.float32
    scs = [make_resident_scene(grids[i % 4], logs[i % 3], 1000 + i, DEV, table_dtype=tdt) for i in range(24)]
    cams = scenes.hemisphere_cameras(49)
    reqs = []
    for k, (W, H) in enumerate(((64, 64), (200, 150), (400, 300), (800, 600))):
        reqs.append(RenderRequest(3 + 5 * k, cams[7 * k + 3], W, H, 1.5 * W, 1.5 * W, W / 2, H / 2, 0.001 * k))
    gen = torch.Generator(device=DEV)
    for k, n in enumerate((10 ** 4, 10 ** 5, 10 ** 6)):
        gen.manual_seed(500 + k)
        pts = torch.rand((n, 3), generator=gen, device=DEV) * 2.0 - 1.0
        reqs.append(QueryRequest(2 + 7 * k, pts, 0.0015 * k))
    srv = Server(scs, mlp_mode=mlp_mode, ray_capacity=1 << 16, point_capacity=1 << 20, ray_arena=1 << 23,
                 point_arena=1 << 24)
    mixed = srv.serve(reqs, realtime=False)
    print(f"config 5 ({'bf16' if bf16 else 'fp32'}): 24 scenes, {len(reqs)} requests", flush=True)
    for i, (r, q) in enumerate(zip(mixed, reqs)):
        rs = scs[q.scene]
        if isinstance(q, RenderRequest):
            check_render(rs, q, r, mlp_mode, 64, seed=50 + i)
        else:
            check_query(rs, q.points, r, 1000, seed=60 + i)
        alone = srv.serve([q], realtime=False)[0]
        for k in r.outputs:
            assert torch.equal(alone.outputs[k], r.outputs[k]), (i, k)
