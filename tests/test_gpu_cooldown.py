"""Progressive super-resolution on the fused training step and the post-upsample eikonal
clamp (trainer.py:131-164, 296-318) on the GPU:

* ``vxn_eik_clamp_add`` equals the oracle bit for bit (same float32 operations), on the
  vector and the unaligned scalar paths, with and without clamping, and with a NaN;
* a cooldown step's grid gradient equals the plain step's (eikonal weight 0) plus the
  oracle's clamped eikonal gradient over the step's V_R (tolerance: the scatters add in
  a different order, 1e-5 relative to the largest entry);
* ``SuperResolutionSchedule.apply`` rescales the grid as the oracle's F.interpolate
  (bit-exact), restarts the SDF Adam at the stage rate, and arms the cooldown; the next
  step and optimizer step run at the new shape and match a step built on that grid."""

import math

import pytest
import torch

from oracle import ref_cpu as R

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


@pytest.fixture(scope="module")
def Fn():
    from voxneus_amd import functional
    return functional


@pytest.mark.parametrize("n,off,w,clamp,nan", [(1000003, 0, 0.1, 0.1, False), (4096, 1, 0.1, 1e9, False),
                                                (777, 3, 2.0, 1e-3, False), (5, 0, 1.0, 0.5, False),
                                                (20000, 0, 0.1, 0.05, True)])
def test_clamp_add_matches_oracle(Fn, n, off, w, clamp, nan):
    g = torch.Generator().manual_seed(n + off)
    grad = torch.randn(n + off, generator=g)
    src = torch.randn(n + off, generator=g) * 3.0
    if nan:
        src[off + n // 3] = math.nan
    want = R.eik_clamp_add(grad[off:], src[off:], w, clamp)
    dd, sd = grad.to(DEV), src.to(DEV)
    dst = dd[off:]  # an offset of off floats: off % 4 != 0 takes the scalar path
    Fn.eik_clamp_add(dst, sd[off:], w, clamp)
    torch.testing.assert_close(dst.cpu(), want, rtol=0, atol=0, equal_nan=True)
    assert torch.equal(dd[:off].cpu(), grad[:off])


def test_clamp_add_empty(Fn):
    e = torch.empty(0, device=DEV)
    Fn.eik_clamp_add(e, torch.empty(0, device=DEV), 0.1, 0.1)


def _step(cfg, like=None):
    """A step on the sphere scene, or on copies of another step's parameters."""
    from voxneus_amd import scenes
    from voxneus_amd.training import FusedTrainStep
    sc = scenes.config1_scene(seed=0)
    if like is None:
        mlp = torch.cat([torch.from_numpy(sc.mlp[k]).reshape(-1) for k in ("W0", "b0", "W1", "b1")])
        p = (torch.from_numpy(sc.grid).to(DEV), torch.from_numpy(sc.embeddings).to(DEV), mlp.to(DEV),
             torch.tensor(0.3, device=DEV))
    else:
        p = tuple(t.detach().clone() for t in (like.grid, like.table, like.mlp, like.variance))
    return FusedTrainStep(*p, R.hash_level_resolutions(16, 16, sc.hash_max_resolution), cfg=cfg)


def _rays():
    from voxneus_amd import scenes
    o, d, tgt = scenes.config_rays("sphere", 512, seed=3)
    return tuple(torch.from_numpy(x).to(DEV) for x in (o, d, tgt))


def _eikonal_oracle(st):
    lin = st.active_linear().cpu().to(torch.int64)
    X, Y, Z = st.grid.shape
    act = torch.stack([lin // (Y * Z), (lin // Z) % Y, lin % Z], -1)
    return R.eikonal_loss_and_grad(st.grid.cpu(), act, st.eps3)[1]


def _close(a, b):
    tol = 1e-5 * float(b.abs().max())
    err = float((a - b).abs().max())
    assert err <= tol, (err, tol)


@pytest.mark.parametrize("concurrent", [True, False])
def test_cooldown_step_adds_the_clamped_eikonal_gradient(concurrent):
    from voxneus_amd.training import StepConfig
    o, d, tgt = _rays()
    base = _step(StepConfig(w_eik=0.0, concurrent_regularisers=concurrent))
    base.step(o, d, tgt)
    g0 = base.grad_grid.cpu()
    e = _eikonal_oracle(base)
    m = float((e * 0.1).abs().max())
    assert m > 0.0
    for clamp, should_scale in ((0.25 * m, True), (4.0 * m, False)):
        st = _step(StepConfig(w_eik=0.1, eik_grad_clamp=clamp, concurrent_regularisers=concurrent))
        st.post_upsample_cooldown = 3
        st.step(o, d, tgt)
        assert st.post_upsample_cooldown == 2
        want = R.eik_clamp_add(g0, e, 0.1, clamp)
        _close(st.grad_grid.cpu(), want)
        added = want - g0
        assert (float(added.abs().max()) < 0.3 * m) == should_scale
        assert torch.equal(st.reg_losses.cpu(), base.reg_losses.cpu())
    # cooldown over: the fused path (eikonal and curvature in one pass), no clamp
    st = _step(StepConfig(w_eik=0.1, eik_grad_clamp=0.25 * m, concurrent_regularisers=concurrent))
    st.step(o, d, tgt)
    _close(st.grad_grid.cpu(), R.eik_clamp_add(g0, e, 0.1, 0.0, cooldown=False))


def test_super_resolution_schedule_on_the_fused_step():
    from voxneus_amd.training import FusedOptimizer, StepConfig, SuperResolutionSchedule
    o, d, tgt = _rays()
    cfg = StepConfig(eik_grad_clamp=1e-4)
    st = _step(cfg)
    opt = FusedOptimizer(st, total_steps=100)
    st.step(o, d, tgt)
    opt.step(0)
    old = st.grid.cpu()
    sch = SuperResolutionSchedule(milestones=(1,), resolutions=(64, 96), lr_sdf_stages=(0.1, 0.05), cooldown_steps=2)
    assert not sch.apply(0, st, opt)
    assert sch.apply(1, st, opt)
    assert tuple(st.grid.shape) == (96, 96, 96)
    assert torch.equal(st.grid.cpu(), R.upsample_sdf_grid(old[None, None], (96, 96, 96))[0, 0])
    assert opt.lr_sdf == 0.05 and opt.sdf_step == 0 and float(opt.grid_m.abs().max()) == 0.0
    assert st.post_upsample_cooldown == 2
    # the next step at the new shape equals a step object built on the rescaled grid and
    # the once-updated hash table, MLP and variance
    ref = _step(cfg, like=st)
    ref.post_upsample_cooldown = 2
    out = st.step(o, d, tgt)
    out_ref = ref.step(o, d, tgt)
    assert st.post_upsample_cooldown == 1 and ref.post_upsample_cooldown == 1
    assert torch.equal(out["mid_z_vals"], out_ref["mid_z_vals"])
    _close(st.grad_grid.cpu(), ref.grad_grid.cpu())
    assert abs(float(out["loss"]) - float(out_ref["loss"])) <= 1e-6 * abs(float(out_ref["loss"]))
    assert torch.equal(st.vr_counts.cpu(), ref.vr_counts.cpu())
    opt.step(1)
    assert opt.sdf_step == 1 and int(opt.n_updated.item()) == int((st.grad_grid != 0).sum())


def test_optimizer_refuses_a_rescaled_grid_without_reset():
    from voxneus_amd.training import FusedOptimizer, StepConfig
    st = _step(StepConfig())
    opt = FusedOptimizer(st, total_steps=10)
    st.rescale_grid(80)
    with pytest.raises(ValueError, match="reset_sdf"):
        opt.step(0)
