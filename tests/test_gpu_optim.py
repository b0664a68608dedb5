"""GPU parity of the optimizer kernels with the CPU oracle on the same inputs, through
the C ABI: the native dense and masked Adam, the masked PyTorch formula with weight
decay, and the multi-tensor torch.optim.Adam with its device-side options.  All
bit-exact (same operation order, one rounding per operation)."""

import ctypes
import math

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R
from voxneus_amd import _native as N

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


def _data(n, steps, seed, zero_frac=0.0):
    g = torch.Generator().manual_seed(seed)
    p = torch.randn(n, generator=g)
    grads = []
    for _ in range(steps):
        gr = torch.randn(n, generator=g) * torch.exp(torch.randn(n, generator=g) * 2.0)
        if zero_frac:
            gr[torch.rand(n, generator=g) < zero_frac] = 0.0
        grads.append(gr)
    return p, grads


def _stream():
    return N.stream_ptr(DEV)


@pytest.mark.parametrize("masked", [False, True])
def test_native_adam_matches_oracle(masked):
    n = 100003  # not a multiple of 4: the vector tail
    p0, grads = _data(n, 5, seed=1, zero_frac=0.3 if masked else 0.0)
    p, m, v = p0.clone(), torch.zeros(n), torch.zeros(n)
    pd, md, vd = p0.to(DEV), torch.zeros(n, device=DEV), torch.zeros(n, device=DEV)
    cnt = torch.zeros(1, device=DEV, dtype=torch.int32)
    L = N.lib()
    for t, g in enumerate(grads, 1):
        k = R.adam_native_step(p, g, m, v, t, 0.9, 0.99, 0.1, 1e-8, masked)
        gd = g.to(DEV)
        if masked:
            cnt.zero_()
            N.check(L.vxn_masked_adam_upd(pd.data_ptr(), gd.data_ptr(), md.data_ptr(), vd.data_ptr(), n, t, 0.9,
                                          0.99, 0.1, 1e-8, cnt.data_ptr(), _stream()))
            assert int(cnt.item()) == k
        else:
            N.check(L.vxn_adam_upd(pd.data_ptr(), gd.data_ptr(), md.data_ptr(), vd.data_ptr(), n, t, 0.9, 0.99,
                                   0.1, 1e-8, _stream()))
        assert torch.equal(pd.cpu(), p) and torch.equal(md.cpu(), m) and torch.equal(vd.cpu(), v), t


def test_sparse_adam_with_weight_decay_matches_oracle():
    n = 65537
    p0, grads = _data(n, 5, seed=2, zero_frac=0.5)
    p, m, v = p0.clone(), torch.zeros(n), torch.zeros(n)
    pd, md, vd = p0.to(DEV), torch.zeros(n, device=DEV), torch.zeros(n, device=DEV)
    cnt = torch.zeros(1, device=DEV, dtype=torch.int32)
    L = N.lib()
    for t, g in enumerate(grads, 1):
        k = R.sparse_adam_masked_step(p, g, m, v, t, 0.05, 0.9, 0.99, 1e-8, 0.01)
        gd = g.to(DEV)
        cnt.zero_()
        N.check(L.vxn_sparse_adam_decay(pd.data_ptr(), gd.data_ptr(), md.data_ptr(), vd.data_ptr(), n, t, 0.9, 0.99,
                                        0.05, 1e-8, 0.01, cnt.data_ptr(), _stream()))
        assert int(cnt.item()) == k
        assert torch.equal(pd.cpu(), p) and torch.equal(md.cpu(), m) and torch.equal(vd.cpu(), v), t


def test_multi_tensor_adam_matches_oracle():
    """Hash-table-, MLP- and variance-shaped tensors in one launch; the variance's rate is
    scaled on the device when its flag is set, and it is floored; a bf16 parameter keeps
    an FP32 master and receives its rounding."""
    sizes = [16 * 1024 * 2 + 2, 3523, 1]
    lrs = [1e-2, 1e-3, 5e-4]
    data = [_data(n, 6, seed=10 + i) for i, n in enumerate(sizes)]
    host = [[p.clone(), torch.zeros_like(p), torch.zeros_like(p)] for p, _ in data]
    dev = [[p.to(DEV), torch.zeros(p.shape, device=DEV), torch.zeros(p.shape, device=DEV)] for p, _ in data]
    bf = torch.zeros(sizes[0], device=DEV, dtype=torch.bfloat16)
    flag = torch.zeros(1, device=DEV, dtype=torch.int32)
    floor = -0.5
    L = N.lib()
    for t in range(1, 7):
        amplified = t % 2 == 0
        flag.fill_(1 if amplified else 0)
        grads_d = [data[i][1][t - 1].to(DEV) for i in range(3)]
        ts = (N.AdamTensor * 3)()
        for i in range(3):
            g = data[i][1][t - 1]
            lr = lrs[i] * (5.0 if (i == 2 and amplified) else 1.0)
            R.torch_adam_step(*host[i][:1], g, host[i][1], host[i][2], t, lr)
            if i == 2:
                host[i][0].clamp_min_(floor)
            d = ts[i]
            d.param, d.grad = dev[i][0].data_ptr(), grads_d[i].data_ptr()
            d.exp_avg, d.exp_avg_sq = dev[i][1].data_ptr(), dev[i][2].data_ptr()
            d.n, d.lr = sizes[i], lrs[i]
            d.lr_flag, d.lr_flag_scale = (flag.data_ptr(), 5.0) if i == 2 else (None, 1.0)
            d.param_min = floor if i == 2 else -math.inf
            d.param_bf16 = bf.data_ptr() if i == 0 else None
        N.check(L.vxn_adam_step(ctypes.cast(ts, ctypes.c_void_p), 3, t, 0.9, 0.999, 1e-8, _stream()))
        for i in range(3):
            for a, b in zip(dev[i], host[i]):
                assert torch.equal(a.cpu(), b), (t, i)
        assert torch.equal(bf.cpu(), host[0][0].to(torch.bfloat16))


def test_optimizer_classes_on_the_device():
    """SparseGridAdam / NetAdam (create_optimizer's groups) step parameters exactly as the
    oracle steps CPU copies with the same gradients; last_updated_count is the number of
    nonzero grid gradients."""
    from voxneus_amd.optim import NetAdam, SparseGridAdam
    g = torch.Generator().manual_seed(5)
    grid = torch.randn(1, 1, 20, 20, 20, generator=g)
    table = torch.randn(16, 256, 2, generator=g).to(torch.bfloat16)
    w = torch.randn(64, 51, generator=g)
    var = torch.tensor(0.3)
    P = {k: torch.nn.Parameter(x.to(DEV)) for k, x in (("grid", grid), ("table", table), ("w", w), ("var", var))}
    sdf = SparseGridAdam([P["grid"]], lr=0.1, betas=(0.9, 0.99))
    net = NetAdam([{"params": [P["table"]], "lr": 1e-2}, {"params": [P["w"]], "lr": 1e-3},
                   {"params": [P["var"]], "lr": 5e-4}], betas=(0.9, 0.999), eps=1e-8)
    host = {"grid": [grid.clone(), torch.zeros_like(grid), torch.zeros_like(grid)],
            "table": [table.float(), torch.zeros(table.shape), torch.zeros(table.shape)],
            "w": [w.clone(), torch.zeros_like(w), torch.zeros_like(w)],
            "var": [var.clone(), torch.zeros_like(var), torch.zeros_like(var)]}
    lrs = {"table": 1e-2, "w": 1e-3, "var": 5e-4}
    for t in range(1, 4):
        grads = {k: torch.randn(x[0].shape, generator=g) for k, x in host.items()}
        grads["grid"][torch.rand(grid.shape, generator=g) < 0.7] = 0.0
        for k in P:
            P[k].grad = grads[k].to(DEV).to(P[k].dtype) if k != "table" else grads[k].to(DEV).to(torch.bfloat16)
        sdf.step()
        net.step()
        R.adam_native_step(*host["grid"][:1], grads["grid"], host["grid"][1], host["grid"][2], t, 0.9, 0.99, 0.1,
                           1e-8, True)
        for k in ("table", "w", "var"):
            gk = grads[k].to(torch.bfloat16).float() if k == "table" else grads[k]
            R.torch_adam_step(host[k][0], gk, host[k][1], host[k][2], t, lrs[k])
        assert sdf.last_updated_count == int((grads["grid"] != 0).sum())
        assert torch.equal(P["grid"].detach().cpu(), host["grid"][0])
        assert torch.equal(net.state[P["table"]]["master"].cpu(), host["table"][0])
        assert torch.equal(P["table"].detach().cpu(), host["table"][0].to(torch.bfloat16))
        assert torch.equal(P["w"].detach().cpu(), host["w"][0])
        assert torch.equal(P["var"].detach().cpu(), host["var"][0])


def test_fused_optimizer_step_matches_oracle():
    """FusedOptimizer after a fused training step: every parameter equals the oracle's
    optimizer applied to CPU copies of the same parameters and device gradients."""
    from oracle.ref_cpu import hash_level_resolutions
    from voxneus_amd import scenes
    from voxneus_amd.optim import cosine_lr_factor
    from voxneus_amd.training import FusedOptimizer, FusedTrainStep, StepConfig
    sc = scenes.config1_scene(seed=0)
    mlp = torch.cat([torch.from_numpy(sc.mlp[k]).reshape(-1) for k in ("W0", "b0", "W1", "b1")])
    st = FusedTrainStep(torch.from_numpy(sc.grid).to(DEV), torch.from_numpy(sc.embeddings).to(DEV), mlp.to(DEV),
                        torch.tensor(0.3, device=DEV), hash_level_resolutions(16, 16, sc.hash_max_resolution),
                        cfg=StepConfig())
    opt = FusedOptimizer(st, total_steps=100)
    o, d, tgt = scenes.config_rays("sphere", 256, seed=3)
    o, d, tgt = (torch.from_numpy(x).to(DEV) for x in (o, d, tgt))
    host = {"grid": st.grid.cpu(), "table": st.table.cpu(), "mlp": st.mlp.cpu(), "var": st.variance.cpu()}
    mom = {k: (torch.zeros_like(x), torch.zeros_like(x)) for k, x in host.items()}
    for it in range(3):
        st.step(o, d, tgt)
        gr = {"grid": st.grad_grid.cpu(), "table": st.grad_table.cpu(), "mlp": st.grad_mlp.cpu(),
              "var": st.grad_v.cpu().reshape(())}
        amp = bool(st.amplified.item())
        opt.step(it, variance_floor=0.25)
        t = it + 1
        R.adam_native_step(host["grid"], gr["grid"], *mom["grid"], t, 0.9, 0.99, 0.1, 1e-8, True)
        f = cosine_lr_factor(it, 100)
        R.torch_adam_step(host["table"], gr["table"], *mom["table"], t, 0.01 * f)
        R.torch_adam_step(host["mlp"], gr["mlp"], *mom["mlp"], t, 1e-3 * f)
        R.torch_adam_step(host["var"], gr["var"], *mom["var"], t, 5e-4 * (5.0 if amp else 1.0))
        host["var"].clamp_min_(0.25)
        assert int(opt.n_updated.item()) == int((gr["grid"] != 0).sum())
        for k, dv in (("grid", st.grid), ("table", st.table), ("mlp", st.mlp), ("var", st.variance)):
            assert torch.equal(dv.cpu(), host[k]), (it, k)


def test_sdf_optimizer_after_a_milestone_takes_the_default_betas():
    """At a super-resolution milestone the reference replaces the grid's optimizer with
    SparseGridAdam(sdf_params, lr=new_lr) (trainer.py:157), i.e. the class default betas
    (0.9, 0.999) (optimizer.py:24), not create_optimizer's (0.9, 0.99): after
    SuperResolutionSchedule.apply the grid's updates equal the oracle's native Adam with
    beta2 = 0.999 at the stage rate, from step 1 with fresh moments."""
    from oracle.ref_cpu import hash_level_resolutions
    from voxneus_amd import scenes
    from voxneus_amd.training import FusedOptimizer, FusedTrainStep, StepConfig, SuperResolutionSchedule
    sc = scenes.config1_scene(seed=0)
    mlp = torch.cat([torch.from_numpy(sc.mlp[k]).reshape(-1) for k in ("W0", "b0", "W1", "b1")])
    st = FusedTrainStep(torch.from_numpy(sc.grid).to(DEV), torch.from_numpy(sc.embeddings).to(DEV), mlp.to(DEV),
                        torch.tensor(0.3, device=DEV), hash_level_resolutions(16, 16, sc.hash_max_resolution),
                        cfg=StepConfig())
    opt = FusedOptimizer(st, total_steps=100)
    o, d, tgt = (torch.from_numpy(x).to(DEV) for x in scenes.config_rays("sphere", 256, seed=3))
    st.step(o, d, tgt)
    opt.step(0)
    sched = SuperResolutionSchedule(milestones=(1,), resolutions=(64, 80), lr_sdf_stages=(0.1, 0.05),
                                    cooldown_steps=0)
    assert sched.apply(1, st, opt) and tuple(st.grid.shape) == (80, 80, 80)
    host = st.grid.cpu()
    m, v = torch.zeros_like(host), torch.zeros_like(host)
    for it in range(1, 4):
        st.step(o, d, tgt)
        g = st.grad_grid.cpu()
        opt.step(it)
        R.adam_native_step(host, g, m, v, it, 0.9, 0.999, 0.05, 1e-8, True)
        assert torch.equal(st.grid.cpu(), host), it


@pytest.mark.parametrize("form", ["fp32", "bf16"])
def test_fused_optimizer_resumes_from_a_checkpoint(form, tmp_path):
    """Three steps, then FusedOptimizer.state_dict() through torch.save / torch.load
    (weights_only=True) into a fresh step and optimizer holding copies of the parameters:
    after the next step the hash table, MLP, variance and their moments equal the
    uninterrupted run's bit for bit (the bf16 form's FP32 master included), the grid and
    its moments to float rounding.  The saved layout is the reference's
    (checkpoint.py:30-31): SparseGridAdam's and torch.optim.Adam's state dicts, which a
    torch.optim.Adam over create_optimizer's parameter groups (optimizer.py:142-147)
    accepts."""
    from oracle.ref_cpu import hash_level_resolutions
    from voxneus_amd import scenes
    from voxneus_amd.training import FusedOptimizer, FusedTrainStep, StepConfig

    sc = scenes.config1_scene(seed=0)
    mlp = torch.cat([torch.from_numpy(sc.mlp[k]).reshape(-1) for k in ("W0", "b0", "W1", "b1")]).to(DEV)
    table = torch.from_numpy(sc.embeddings).to(DEV)
    mode = N.VXN_MLP_FP32
    if form == "bf16":
        table, mode = table.to(torch.bfloat16), N.VXN_MLP_BF16_MFMA
    res = hash_level_resolutions(16, 16, sc.hash_max_resolution)

    def make(grid, tab, w, var):
        st = FusedTrainStep(grid, tab, w, var, res, cfg=StepConfig(mlp_mode=mode))
        return st, FusedOptimizer(st, total_steps=100)

    st, opt = make(torch.from_numpy(sc.grid).to(DEV), table, mlp, torch.tensor(0.3, device=DEV))
    o, d, tgt = (torch.from_numpy(x).to(DEV) for x in scenes.config_rays("sphere", 256, seed=3))
    for it in range(3):
        st.step(o, d, tgt)
        opt.step(it)
    torch.save(opt.state_dict(), tmp_path / "opt.pt")
    mlp_m_saved = opt.mlp_m.clone()
    st2, opt2 = make(st.grid.clone(), st.table.clone(), st.mlp.clone(), st.variance.clone())
    opt2.load_state_dict(torch.load(tmp_path / "opt.pt", weights_only=True))
    for s_, o_ in ((st, opt), (st2, opt2)):
        s_.step(o, d, tgt)
        o_.step(3)
    torch.cuda.synchronize()
    for name in ("table", "mlp", "variance"):
        assert torch.equal(getattr(st, name), getattr(st2, name)), name
    for name in ("table_m", "table_v", "mlp_m", "mlp_v", "var_m", "var_v"):
        assert torch.equal(getattr(opt, name), getattr(opt2, name)), name
    # the grid gradient's device atomics add in another order on every run (the step's
    # run-to-run test states it), so the grid and its moments agree to float rounding
    for a_, b_ in ((st.grid, st2.grid), (opt.grid_m, opt2.grid_m), (opt.grid_v, opt2.grid_v)):
        assert float((a_ - b_).abs().max()) <= 1e-5 * float(a_.abs().max())
    if form == "bf16":
        assert torch.equal(opt.table_master, opt2.table_master)
    # the reference's torch.optim.Adam takes the saved network state as it is
    sd = torch.load(tmp_path / "opt.pt", weights_only=True)["net_optimizer"]
    params = [torch.zeros(tuple(st.table.shape)), torch.zeros(64, 51), torch.zeros(64), torch.zeros(3, 64),
              torch.zeros(3), torch.zeros(())]
    ref = torch.optim.Adam([{"params": params[:1], "lr": 1e-2}, {"params": params[1:5], "lr": 1e-3},
                            {"params": params[5:], "lr": 5e-4}], betas=(0.9, 0.999), eps=1e-8)
    ref.load_state_dict({"state": {k: {kk: vv.cpu() for kk, vv in v.items()} for k, v in sd["state"].items()},
                         "param_groups": sd["param_groups"]})
    assert float(ref.state[params[1]]["step"]) == 3.0
    assert torch.equal(ref.state[params[1]]["exp_avg"], mlp_m_saved[:64 * 51].reshape(64, 51).cpu())
    assert tuple(ref.state[params[0]]["exp_avg_sq"].shape) == tuple(st.table.shape)
