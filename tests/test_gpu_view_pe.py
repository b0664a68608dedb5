"""The positional-encoding view embedding (VoxNeuS view_embed_type="pe", models/voxneus.py:
78-82; embedder.py:12-44) through the fused render on the GPU, against the oracle on the
same z values and a fixed random upstream gradient:

* FP32 form: forward outputs and the autograd gradients of the grid, hash table, MLP
  (W0 [64, 62]) and variance within 1e-5 of the oracle's PyTorch autograd (the device's
  sin / cos and the CPU's differ by an ulp);
* BF16 form: within 1e-4 of the oracle's emulation of the same rounding points;
* the fused training step with view_embed_type="pe": Huber loss and MLP gradient equal
  the oracle's on the step's own z values;
* the SH form is unchanged (the existing golden-vector tests cover it)."""

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R
from tests.conftest import load_golden
from voxneus_amd import scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
BOX6 = (-1.0, -1.0, -1.0, 1.0, 1.0, 1.0)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _pe_mlp(sc, seed=7):
    """The scene's SH MLP with 27 random view columns in place of the 16 SH ones."""
    g = torch.Generator().manual_seed(seed)
    W0 = torch.from_numpy(sc.mlp["W0"])
    Wv = (torch.rand(64, 27, generator=g) - 0.5) * 0.6
    W0p = torch.cat([W0[:, :32], Wv, W0[:, 48:51]], 1).contiguous()
    return {"W0": W0p.numpy(), "b0": sc.mlp["b0"], "W1": sc.mlp["W1"], "b1": sc.mlp["b1"]}


def _flat(mlp):
    return torch.cat([torch.from_numpy(np.ascontiguousarray(mlp[k])).reshape(-1) for k in ("W0", "b0", "W1", "b1")])


def _pair(form, name="step_c1.npz", table_dtype=torch.float32):
    from voxneus_amd import functional as Fn
    from voxneus_amd import _native as N
    g = load_golden(name)
    sc = scenes.config1_scene(seed=0) if name == "step_c1.npz" else scenes.make_scene("multi", 128, 17, seed=2)
    mlp = _pe_mlp(sc)
    emb = torch.from_numpy(sc.embeddings).to(table_dtype)
    m = R.RefModel(sc.grid, emb.float().numpy(), **mlp, hash_max_resolution=sc.hash_max_resolution,
                   mlp_form=form, view_embed_type="pe")
    o, d = torch.from_numpy(g["rays_o"]), torch.from_numpy(g["rays_d"])
    zs = tuple(torch.from_numpy(g[k]) for k in ("z_vals", "dists", "mid"))
    rng = np.random.default_rng(5)
    Rn = o.shape[0]
    rr = [torch.from_numpy(rng.normal(size=sh).astype(np.float32)) for sh in ((Rn, 3), (Rn, 1), (Rn, 3), (Rn, 1))]
    out = m.render(o, d, None, None, zs=zs)
    L = sum((out[k] * rr[i]).sum() for i, k in enumerate(("rgb", "depth", "normal", "opacity")))
    L.backward()
    mode = N.VXN_MLP_BF16_MFMA if form == "bf16" else N.VXN_MLP_FP32
    gt = torch.zeros(emb.shape, device=DEV, dtype=torch.float32) if form == "bf16" else None
    cfg = Fn.RenderConfig(BOX6, m.level_res, mlp_mode=mode, table_grad_out=gt, view_embed=N.VXN_VIEW_PE)
    grid = torch.from_numpy(sc.grid).to(DEV).requires_grad_(True)
    ed = emb.to(DEV).requires_grad_(form != "bf16")
    mf = _flat(mlp).to(DEV).requires_grad_(True)
    v = torch.tensor(0.3, device=DEV, requires_grad=True)
    s = torch.clamp(torch.exp(v * 10.0), min=1e-4, max=300.0)
    outs = Fn.RenderFunction.apply(grid, ed, mf, s, o.to(DEV), d.to(DEV), zs[2].to(DEV), zs[1].to(DEV), cfg, False,
                                   None)
    Lg = sum((outs[i] * rr[i].to(DEV)).sum() for i in range(4))
    Lg.backward()
    ref_mlp = torch.cat([m.W0.grad.reshape(-1), m.b0.grad, m.W1.grad.reshape(-1), m.b1.grad]).numpy()
    table_grad = gt.cpu().numpy() if gt is not None else ed.grad.cpu().numpy()
    return {"out": out, "outs": [t.detach().cpu() for t in outs[:4]], "m": m, "grid": grid.grad.cpu().numpy(),
            "table": table_grad, "mlp": mf.grad.cpu().numpy(), "ref_mlp": ref_mlp, "v": v.grad.item()}


def test_pe_fp32_form_matches_oracle():
    r = _pair("fp32")
    m = r["m"]
    assert r["mlp"].shape == (64 * 62 + 64 + 3 * 64 + 3,)
    for k, i in (("rgb", 0), ("depth", 1), ("normal", 2), ("opacity", 3)):
        assert float((r["outs"][i] - r["out"][k].detach()).abs().max()) < 1e-5, k
    assert rel_err(r["grid"], m.grid.grad.numpy()[0, 0]) < 1e-5
    assert rel_err(r["table"], m.embeddings.grad.numpy()) < 1e-5
    assert rel_err(r["mlp"], r["ref_mlp"]) < 1e-5
    # the 27 view columns of W0 receive their own gradients
    g0 = r["mlp"][:64 * 62].reshape(64, 62)
    assert rel_err(g0[:, 32:59], r["ref_mlp"][:64 * 62].reshape(64, 62)[:, 32:59]) < 1e-5
    assert abs(r["v"] - m.variance.grad.item()) <= 1e-5 * max(1e-3, abs(m.variance.grad.item()))


@pytest.mark.parametrize("name", ["step_c1.npz", "step_c2.npz"])
def test_pe_bf16_form_matches_emulation(name):
    r = _pair("bf16", name, table_dtype=torch.bfloat16)
    m = r["m"]
    for k, i in (("rgb", 0), ("depth", 1), ("normal", 2), ("opacity", 3)):
        assert float((r["outs"][i] - r["out"][k].detach()).abs().max()) < 1e-4, k
    assert rel_err(r["grid"], m.grid.grad.numpy()[0, 0]) < 1e-4
    assert rel_err(r["table"], m.embeddings.grad.numpy()) < 1e-4
    assert rel_err(r["mlp"], r["ref_mlp"]) < 1e-4
    assert abs(r["v"] - m.variance.grad.item()) <= 1e-4 * max(1e-3, abs(m.variance.grad.item()))


def test_pe_fused_train_step_matches_oracle():
    from voxneus_amd.training import FusedTrainStep, StepConfig
    g = load_golden("step_c1.npz")
    sc = scenes.config1_scene(seed=0)
    mlp = _pe_mlp(sc)
    st = FusedTrainStep(torch.from_numpy(sc.grid).to(DEV), torch.from_numpy(sc.embeddings).to(DEV),
                        _flat(mlp).to(DEV), torch.tensor(0.3, device=DEV),
                        R.hash_level_resolutions(16, 16, sc.hash_max_resolution),
                        cfg=StepConfig(view_embed_type="pe"))
    o, d, tgt = (torch.from_numpy(g[k]).to(DEV) for k in ("rays_o", "rays_d", "target"))
    out = st.step(o, d, tgt)
    assert np.array_equal(out["z_vals"].cpu().numpy(), g["z_vals"])  # the sampler is independent of the MLP
    m = R.RefModel(sc.grid, sc.embeddings, **mlp, hash_max_resolution=sc.hash_max_resolution, view_embed_type="pe")
    zs = tuple(x.cpu() for x in (out["z_vals"], out["dists"], out["mid_z_vals"]))
    ref = m.render(o.cpu(), d.cpu(), None, None, zs=zs)
    loss = R.huber(ref["rgb"], tgt.cpu())
    loss.backward()
    lv = float(loss.detach())
    assert abs(float(out["loss"]) - lv) <= 1e-6 * abs(lv)
    ref_mlp = torch.cat([m.W0.grad.reshape(-1), m.b0.grad, m.W1.grad.reshape(-1), m.b1.grad]).numpy()
    assert rel_err(st.grad_mlp.cpu().numpy(), ref_mlp) < 1e-5
    # a PE step refuses an SH-sized MLP
    with pytest.raises(ValueError):
        FusedTrainStep(torch.from_numpy(sc.grid).to(DEV), torch.from_numpy(sc.embeddings).to(DEV),
                       torch.zeros(3523, device=DEV), torch.tensor(0.3, device=DEV),
                       R.hash_level_resolutions(16, 16, sc.hash_max_resolution), cfg=StepConfig(view_embed_type="pe"))


def test_pe_model_renders_through_the_fused_path():
    from voxneus_amd.voxneus import VoxNeuS
    torch.manual_seed(3)
    m = VoxNeuS(grid_res=32, hash_log2_size=14, view_embed_type="pe").to(DEV)
    assert tuple(m.radiance_mlp.net[0].weight.shape) == (64, 62)
    g = load_golden("step_c1.npz")
    o, d = torch.from_numpy(g["rays_o"][:16]).to(DEV), torch.from_numpy(g["rays_d"][:16]).to(DEV)
    from voxneus_amd.rendering import RaySampler
    out = m.render(o, d, RaySampler(64, 64), None, None)
    assert out["rgb"].shape == (16, 3) and bool(torch.isfinite(out["rgb"]).all())
    out["rgb"].sum().backward()
    assert m.radiance_mlp.net[0].weight.grad is not None
    assert float(m.radiance_mlp.net[0].weight.grad[:, 32:59].abs().sum()) > 0.0
