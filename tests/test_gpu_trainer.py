"""The trainer's per-step settings on the fused path, on the GPU, against the oracle:

* a fractional NeuS2 lambda (radiance_hash.py:90-102) in the fused render: forward
  outputs and the autograd gradients of grid, hash table (inactive levels get none),
  MLP and variance within 1e-5 of the oracle's PyTorch autograd;
* the NeuS iter_cos alpha (neus_alpha.py:46-57) at anneal ratios 0, 0.3 and 1, forward
  and backward, same bar;
* the fused training step with both and the regularisers, against the oracle's
  train_step on the step's own z values (loss 1e-6, gradients 1e-5);
* ``FusedTrainer`` (trainer.py:166-353): its first step equals the oracle's step at the
  scheduler's weights and lambda; over steps crossing two super-resolution milestones it
  keeps the model's parameters as the trained storage, applies the schedules, the
  amplification gate, the cooldown and the variance floors."""

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


def _flat(mlp):
    return torch.cat([torch.from_numpy(np.ascontiguousarray(mlp[k])).reshape(-1) for k in ("W0", "b0", "W1", "b1")])


def _render_pair(active_levels=None, cos_anneal_ratio=0.0, use_iter_cos=False):
    from voxneus_amd import functional as Fn
    from voxneus_amd.hash_grid import level_weights
    g = load_golden("step_c1.npz")
    sc = scenes.config1_scene(seed=0)
    m = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    o, d = torch.from_numpy(g["rays_o"]), torch.from_numpy(g["rays_d"])
    zs = tuple(torch.from_numpy(g[k]) for k in ("z_vals", "dists", "mid"))
    rng = np.random.default_rng(11)
    Rn = o.shape[0]
    rr = [torch.from_numpy(rng.normal(size=sh).astype(np.float32)) for sh in ((Rn, 3), (Rn, 1), (Rn, 3), (Rn, 1))]
    out = m.render(o, d, None, None, zs=zs, active_levels=active_levels, cos_anneal_ratio=cos_anneal_ratio,
                   use_iter_cos=use_iter_cos)
    L = sum((out[k] * rr[i]).sum() for i, k in enumerate(("rgb", "depth", "normal", "opacity")))
    L.backward()
    cfg = Fn.RenderConfig(BOX6, m.level_res, level_weights(len(m.level_res), active_levels),
                          cos_anneal_ratio=cos_anneal_ratio, use_iter_cos=use_iter_cos)
    grid = torch.from_numpy(sc.grid).to(DEV).requires_grad_(True)
    ed = torch.from_numpy(sc.embeddings).to(DEV).requires_grad_(True)
    mf = _flat(sc.mlp).to(DEV).requires_grad_(True)
    v = torch.tensor(0.3, device=DEV, requires_grad=True)
    s = torch.clamp(torch.exp(v * 10.0), min=1e-4, max=300.0)
    outs = Fn.RenderFunction.apply(grid, ed, mf, s, o.to(DEV), d.to(DEV), zs[2].to(DEV), zs[1].to(DEV), cfg, False,
                                   None)
    Lg = sum((outs[i] * rr[i].to(DEV)).sum() for i in range(4))
    Lg.backward()
    ref_mlp = torch.cat([m.W0.grad.reshape(-1), m.b0.grad, m.W1.grad.reshape(-1), m.b1.grad]).numpy()
    for k, i in (("rgb", 0), ("depth", 1), ("normal", 2), ("opacity", 3)):
        assert float((outs[i].detach().cpu() - out[k].detach()).abs().max()) < 1e-5, k
    assert rel_err(grid.grad.cpu().numpy(), m.grid.grad.numpy()[0, 0]) < 1e-5
    assert rel_err(ed.grad.cpu().numpy(), m.embeddings.grad.numpy()) < 1e-5
    assert rel_err(mf.grad.cpu().numpy(), ref_mlp) < 1e-5
    assert abs(v.grad.item() - m.variance.grad.item()) <= 1e-5 * max(1e-3, abs(m.variance.grad.item()))
    return ed.grad.cpu().numpy(), out


@pytest.mark.parametrize("lam", [5.5, 2.0, 15.25])
def test_fractional_active_levels_render_matches_oracle(lam):
    eg, _ = _render_pair(active_levels=lam)
    # levels past ceil(lambda) receive no gradient, the partial level a scaled one
    full = int(np.ceil(lam))
    assert not np.any(eg[full:])
    assert np.any(eg[full - 1])


@pytest.mark.parametrize("ratio", [0.0, 0.3, 1.0])
def test_iter_cos_render_matches_oracle(ratio):
    _, out = _render_pair(cos_anneal_ratio=ratio, use_iter_cos=True)
    assert bool(torch.isfinite(out["rgb"]).all())


def test_iter_cos_changes_the_render():
    """The iter_cos render differs from the plain one (the switch reaches the kernel)."""
    from voxneus_amd import functional as Fn
    g = load_golden("step_c1.npz")
    sc = scenes.config1_scene(seed=0)
    lr = R.hash_level_resolutions(16, 16, sc.hash_max_resolution)
    args = (torch.from_numpy(sc.grid).to(DEV), torch.from_numpy(sc.embeddings).to(DEV), _flat(sc.mlp).to(DEV),
            torch.tensor(20.0, device=DEV), torch.from_numpy(g["rays_o"]).to(DEV),
            torch.from_numpy(g["rays_d"]).to(DEV), torch.from_numpy(g["mid"]).to(DEV),
            torch.from_numpy(g["dists"]).to(DEV))
    a = Fn.RenderFunction.apply(*args, Fn.RenderConfig(BOX6, lr), False, None)
    b = Fn.RenderFunction.apply(*args, Fn.RenderConfig(BOX6, lr, cos_anneal_ratio=0.3, use_iter_cos=True), False,
                                None)
    assert float((a[3] - b[3]).abs().max()) > 1e-3


@pytest.mark.parametrize("lam,ratio", [(7.25, 0.4), (16.0, 1.0)])
def test_fused_step_with_lambda_and_iter_cos_matches_oracle(lam, ratio):
    from voxneus_amd.training import FusedTrainStep, StepConfig
    g = load_golden("step_c1.npz")
    sc = scenes.config1_scene(seed=0)
    st = FusedTrainStep(torch.from_numpy(sc.grid).to(DEV), torch.from_numpy(sc.embeddings).to(DEV),
                        _flat(sc.mlp).to(DEV), torch.tensor(0.3, device=DEV),
                        R.hash_level_resolutions(16, 16, sc.hash_max_resolution),
                        cfg=StepConfig(w_eik=0.01, w_curv=1e-6, active_levels=lam, use_iter_cos=True,
                                       cos_anneal_ratio=ratio))
    o, d, tgt = (torch.from_numpy(g[k]).to(DEV) for k in ("rays_o", "rays_d", "target"))
    out = st.step(o, d, tgt)
    m = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    zs = tuple(x.cpu() for x in (out["z_vals"], out["dists"], out["mid_z_vals"]))
    ref = m.train_step(o.cpu(), d.cpu(), tgt.cpu(), w_eik=0.01, w_curv=1e-6, zs=zs, active_levels=lam,
                       cos_anneal_ratio=ratio, use_iter_cos=True)
    lv = float(ref["loss"])
    assert abs(float(out["loss"]) - lv) <= 1e-6 * abs(lv)
    ref_mlp = torch.cat([m.W0.grad.reshape(-1), m.b0.grad, m.W1.grad.reshape(-1), m.b1.grad]).numpy()
    assert rel_err(st.grad_mlp.cpu().numpy(), ref_mlp) < 1e-5
    assert rel_err(st.grad_grid.cpu().numpy(), m.grid.grad.numpy()[0, 0]) < 1e-5
    assert rel_err(st.grad_table.cpu().numpy(), m.embeddings.grad.numpy()) < 1e-5
    assert abs(float(st.grad_v.reshape(())) - float(m.variance.grad)) <= 1e-5 * max(1e-3, abs(float(m.variance.grad)))
    # the per-step settings change without rebuilding the step
    st.set_active_levels(3.5)
    st.cfg.use_iter_cos = False
    out2 = st.step(o, d, tgt, zs=tuple(x for x in (out["z_vals"], out["dists"], out["mid_z_vals"])))
    m2 = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    ref2 = m2.train_step(o.cpu(), d.cpu(), tgt.cpu(), w_eik=0.01, w_curv=1e-6, zs=zs, active_levels=3.5)
    assert abs(float(out2["loss"]) - float(ref2["loss"])) <= 1e-6 * abs(float(ref2["loss"]))
    assert not np.any(st.grad_table.cpu().numpy()[4:])


def _trainer_config():
    return {"training": {"total_steps": 400, "sr_milestones": [3, 6], "sr_resolutions": [32, 40, 48],
                         "occupancy_update_interval": 4},
            "sampler": {"perturb": False},
            "scheduler": {"stage1_end": 2, "stage2_end": 5, "s_warmup_steps": 2},
            "optimizer": {"lr_sdf": 0.1, "lr_sdf_stage2": 0.05, "lr_sdf_stage3": 0.02}}


def _batch(n=512):
    g = load_golden("step_c1.npz")
    return {"rays_o": torch.from_numpy(g["rays_o"][:n]), "rays_d": torch.from_numpy(g["rays_d"][:n]),
            "rgb": torch.from_numpy(g["target"][:n])}


def test_fused_trainer_first_step_matches_oracle():
    from voxneus_amd.trainer import FusedTrainer
    from voxneus_amd.voxneus import VoxNeuS
    torch.manual_seed(5)
    model = VoxNeuS(grid_res=32, hash_log2_size=14)
    tr = FusedTrainer(model, _trainer_config(), device=DEV)
    st = tr.step_engine
    mlp = {k: t.detach().cpu().numpy().copy() for k, t in zip(
        ("W0", "b0", "W1", "b1"), (model.radiance_mlp.net[0].weight, model.radiance_mlp.net[0].bias,
                                   model.radiance_mlp.net[2].weight, model.radiance_mlp.net[2].bias))}
    grid0 = st.grid.cpu().numpy().copy()
    emb0 = st.table.cpu().numpy().copy()
    b = _batch()
    met = tr.train_step(0, b)
    out = tr.last_out
    ref = R.RefModel(grid0, emb0, **mlp)
    zs = tuple(x.cpu() for x in (out["z_vals"], out["dists"], out["mid_z_vals"]))
    w_eik, w_curv = tr.scheduler.get_weights(0)
    r = ref.train_step(b["rays_o"], b["rays_d"], b["rgb"], w_eik=w_eik, w_curv=w_curv, amplify=False, zs=zs,
                       active_levels=tr.scheduler.get_lambda(0))
    h = tr.read_metrics(met)
    assert abs(h["loss_total"] - float(r["loss"])) <= 1e-6 * abs(float(r["loss"]))
    assert abs(h["loss_eik"] - float(r["loss_eik"])) <= 1e-5 * abs(float(r["loss_eik"]))
    ref_mlp = torch.cat([ref.W0.grad.reshape(-1), ref.b0.grad, ref.W1.grad.reshape(-1), ref.b1.grad]).numpy()
    assert rel_err(st.grad_mlp.cpu().numpy(), ref_mlp) < 1e-5
    assert rel_err(st.grad_grid.cpu().numpy(), ref.grid.grad.numpy()[0, 0]) < 1e-5
    assert rel_err(st.grad_table.cpu().numpy(), ref.embeddings.grad.numpy()) < 1e-5
    assert abs(float(st.grad_v.reshape(())) - float(ref.variance.grad)) <= 1e-5 * max(1e-3, abs(float(ref.variance.grad)))
    assert h["s_amplified"] is False and h["active_lambda"] == 2.0 and h["w_eik"] == 1e-2
    # PSNR from the colour MSE (trainer.py:234-236) under the Huber loss too
    mse = float(((r["render"]["rgb"].detach().double() - b["rgb"].double()) ** 2).mean())
    assert abs(h["psnr"] - (-10.0 * np.log10(max(mse, 1e-8)))) < 1e-4
    # the optimizer moved the model's own parameters
    assert not np.array_equal(model.sdf_grid.grid.detach().cpu().numpy()[0, 0], grid0)
    assert np.array_equal(model.radiance_mlp.flat().detach().cpu().numpy(), st.mlp.cpu().numpy())


def test_fused_trainer_schedules_over_milestones():
    from voxneus_amd.rendering import RaySampler
    from voxneus_amd.trainer import FusedTrainer
    from voxneus_amd.voxneus import VoxNeuS
    torch.manual_seed(6)
    model = VoxNeuS(grid_res=32, hash_log2_size=14)
    tr = FusedTrainer(model, _trainer_config(), device=DEV)
    st = tr.step_engine
    b = _batch()
    for step in range(9):
        h = tr.read_metrics(tr.train_step(step, b))
        assert np.isfinite(h["loss_total"]) and np.isfinite(h["loss_eik"]), step
        w_eik, w_curv = tr.scheduler.get_weights(step)
        assert (h["w_eik"], h["w_curv"]) == (w_eik, w_curv) and (st.cfg.w_eik, st.cfg.w_curv) == (w_eik, w_curv)
        assert h["active_lambda"] == 2.0 + step / 10.0
        assert st.cfg.cos_anneal_ratio == step / 10000.0
        assert st.cfg.amplify == (step >= 2)
        if step < 2:
            assert h["s_amplified"] is False
        assert h["grid_res"] == (32 if step < 3 else 40 if step < 6 else 48)
        assert h["lr_sdf"] == (0.1 if step < 3 else 0.05 if step < 6 else 0.02)
        assert tr.optimizer.lr_sdf == h["lr_sdf"]
        v = float(st.variance)
        if step >= 6:
            assert v >= 0.5010 - 1e-7
        elif step >= 3:
            assert 0.4382 - 1e-7 <= v < 0.5010
        else:
            assert v < 0.4382
        assert h["sparse_updated"] > 0
    # the cooldown restarted at the second milestone and ran three steps
    assert st.post_upsample_cooldown == 497
    assert tuple(model.sdf_grid.grid.shape) == (1, 1, 48, 48, 48)
    assert model.sdf_grid.grid.data_ptr() == st.grid.data_ptr()
    assert model.radiance_mlp.net[0].weight.data_ptr() == st.mlp.data_ptr()
    assert model.radiance_hash.encoder.embeddings.data_ptr() == st.table.data_ptr()
    assert model.variance_net.variance.data_ptr() == st.variance.data_ptr()
    assert int(model.sdf_grid.res[0]) == 48
    # the model renders the trained state
    with torch.no_grad():
        out = model.render(b["rays_o"][:64].to(DEV), b["rays_d"][:64].to(DEV), RaySampler(64, 64, perturb=False),
                           None, None, return_corner_indices=False)
    assert bool(torch.isfinite(out["rgb"]).all())
    with pytest.raises(ValueError):
        tr.train_step(9)  # no batch and no dataset


def test_concurrent_grid_scatter_equals_serial():
    """The grid scatter on its own stream (vxn_render_args.scatter_stream), beside the
    hash-table gradient, gives the serial step's outputs: loss, MLP gradient, regulariser
    losses and the hash-table gradient bit for bit (its records and sums follow a fixed
    order), the grid gradient up to the order of its atomic adds."""
    from voxneus_amd.training import FusedTrainStep, StepConfig
    g = load_golden("step_c1.npz")
    sc = scenes.config1_scene(seed=0)
    o, d, tgt = (torch.from_numpy(g[k]).to(DEV) for k in ("rays_o", "rays_d", "target"))
    res = []
    for conc in (False, True):
        st = FusedTrainStep(torch.from_numpy(sc.grid).to(DEV), torch.from_numpy(sc.embeddings).to(DEV),
                            _flat(sc.mlp).to(DEV), torch.tensor(0.3, device=DEV),
                            R.hash_level_resolutions(16, 16, sc.hash_max_resolution),
                            cfg=StepConfig(concurrent_scatter=conc))
        assert (st.scatter is not None) == conc
        for _ in range(2):
            out = st.step(o, d, tgt)
        res.append((float(out["loss"]), st.grad_grid.cpu().numpy(), st.grad_table.cpu().numpy(),
                    st.grad_mlp.cpu().numpy(), out["reg_losses"].cpu().numpy()))
    a, b = res
    assert a[0] == b[0]
    assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])
    assert np.array_equal(a[2], b[2])
    assert rel_err(b[1], a[1]) < 1e-6


@pytest.mark.parametrize("form", ["bf16", "pe"])
def test_fused_trainer_other_forms(form):
    """The trainer on the BF16 form (bf16 hash table with an FP32 master in the optimizer,
    BF16 MLP on the matrix cores) and on the positional-encoding view embedding: steps run,
    losses stay finite, and the model's own parameters (the bf16 table included) move."""
    from voxneus_amd import _native as N
    from voxneus_amd.trainer import FusedTrainer
    from voxneus_amd.voxneus import VoxNeuS
    torch.manual_seed(7)
    if form == "bf16":
        model = VoxNeuS(grid_res=32, hash_log2_size=14, hash_dtype=torch.bfloat16, mlp_mode=N.VXN_MLP_BF16_MFMA)
    else:
        model = VoxNeuS(grid_res=32, hash_log2_size=14, view_embed_type="pe")
    tr = FusedTrainer(model, _trainer_config(), device=DEV)
    emb0 = model.radiance_hash.encoder.embeddings.detach().float().cpu().clone()
    w0 = model.radiance_mlp.net[0].weight.detach().cpu().clone()
    b = _batch(256)
    for step in range(4):
        h = tr.read_metrics(tr.train_step(step, b))
        assert np.isfinite(h["loss_total"]) and np.isfinite(h["loss_eik"]), (form, step)
    assert model.radiance_hash.encoder.embeddings.data_ptr() == tr.step_engine.table.data_ptr()
    assert not torch.equal(model.radiance_hash.encoder.embeddings.detach().float().cpu(), emb0)
    assert not torch.equal(model.radiance_mlp.net[0].weight.detach().cpu(), w0)
    assert tuple(model.sdf_grid.grid.shape) == (1, 1, 40, 40, 40)
