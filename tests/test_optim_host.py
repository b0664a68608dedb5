"""Pin the oracle's optimizer restatements (CPU only): torch.optim.Adam itself (the
library routine the reference calls), the cases where the masked and native formulas
reduce to it, Adam's first-step closed form, and the optimizer groups of
create_optimizer (training/optimizer.py:119-154)."""

import math

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R


def _data(n=5000, steps=6, seed=0, zero_frac=0.0):
    g = torch.Generator().manual_seed(seed)
    p = torch.randn(n, generator=g)
    grads = []
    for _ in range(steps):
        gr = torch.randn(n, generator=g) * torch.exp(torch.randn(n, generator=g) * 2.0)
        if zero_frac:
            gr[torch.rand(n, generator=g) < zero_frac] = 0.0
        grads.append(gr)
    return p, grads


def _ulp_close(a, b, ulps=4, scale=0.0):
    """|a - b| within ``ulps`` float32 ulps of max(|a|, |b|), plus ulps of ``scale`` (the
    size of an update: parameters near zero carry the update's rounding)."""
    a, b = a.double(), b.double()
    eps = torch.finfo(torch.float32).eps
    tol = ulps * eps * (torch.maximum(a.abs(), b.abs()) + scale)
    return bool(((a - b).abs() <= tol).all())


def test_torch_adam_restatement_matches_torch_optim():
    p0, grads = _data()
    lr = 3e-3
    ref = p0.clone().requires_grad_(True)
    opt = torch.optim.Adam([ref], lr=lr, betas=(0.9, 0.999), eps=1e-8, foreach=False)
    p, m, v = p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0)
    for t, g in enumerate(grads, 1):
        ref.grad = g.clone()
        opt.step()
        R.torch_adam_step(p, g, m, v, t, lr)
        assert _ulp_close(p, ref.detach(), scale=10 * lr), t
    st = opt.state[ref]
    assert _ulp_close(m, st["exp_avg"], scale=1e-6) and _ulp_close(v, st["exp_avg_sq"])


@pytest.mark.parametrize("wd", [0.0, 0.01])
def test_masked_formula_is_adam_where_every_gradient_is_nonzero(wd):
    """With no zero gradients the masked PyTorch formula (optimizer.py:92-113) is
    torch.optim.Adam with L2 weight decay; with zeros, those elements and their moments
    never change."""
    p0, grads = _data(seed=1)
    ref = p0.clone().requires_grad_(True)
    opt = torch.optim.Adam([ref], lr=0.1, betas=(0.9, 0.99), eps=1e-8, weight_decay=wd, foreach=False)
    p, m, v = p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0)
    for t, g in enumerate(grads, 1):
        ref.grad = g.clone()
        opt.step()
        assert R.sparse_adam_masked_step(p, g, m, v, t, 0.1, 0.9, 0.99, 1e-8, wd) == p.numel()
        assert _ulp_close(p, ref.detach(), 8, scale=1.0), t
    p0, grads = _data(seed=2, zero_frac=0.4)
    p, m, v = p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0)
    never = torch.ones_like(p0, dtype=torch.bool)
    for t, g in enumerate(grads, 1):
        never &= g == 0.0
        R.sparse_adam_masked_step(p, g, m, v, t, 0.1, 0.9, 0.99, 1e-8, wd)
    assert never.any()
    assert torch.equal(p[never], p0[never]) and not m[never].any() and not v[never].any()


def test_native_formula_first_step_and_eps_free_limit():
    """adam_upd_kernel.cu: step 1 moves each element by lr * g / (|g| + eps / sqrt(1 - b2));
    with eps = 0 the native and PyTorch formulas are the same update."""
    p0, grads = _data(seed=3)
    lr, b1, b2, eps = 0.1, 0.9, 0.99, 1e-8
    p, m, v = p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0)
    R.adam_native_step(p, grads[0], m, v, 1, b1, b2, lr, eps, masked=False)
    g64 = grads[0].double()
    want = p0.double() - lr * g64 / (g64.abs() + eps / math.sqrt(1.0 - b2))
    assert torch.allclose(p.double(), want, rtol=0, atol=1e-6)
    ref = p0.clone().requires_grad_(True)
    opt = torch.optim.Adam([ref], lr=lr, betas=(b1, b2), eps=0.0, foreach=False)
    p, m, v = p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0)
    for t, g in enumerate(grads, 1):
        ref.grad = g.clone()
        opt.step()
        R.adam_native_step(p, g, m, v, t, b1, b2, lr, 0.0, masked=False)
        # step sizes formed in float (native) and double (torch) differ by a few ulps
        assert torch.allclose(p, ref.detach(), rtol=0, atol=2e-6), t


def test_native_masked_leaves_zero_gradients_untouched():
    p0, grads = _data(seed=4, zero_frac=0.5)
    p, m, v = p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0)
    pd, md, vd = p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0)
    g = grads[0]
    n = R.adam_native_step(p, g, m, v, 1, 0.9, 0.99, 0.1, 1e-8, masked=True)
    R.adam_native_step(pd, g, md, vd, 1, 0.9, 0.99, 0.1, 1e-8, masked=False)
    nz = g != 0
    assert n == int(nz.sum())
    assert torch.equal(p[nz], pd[nz]) and torch.equal(p[~nz], p0[~nz])


def test_create_optimizer_groups():
    """training/optimizer.py:119-154: sparse Adam on the grid with betas (0.9, 0.99); Adam
    with betas (0.9, 0.999), eps 1e-8 and one group each for hash, MLP, variance."""
    from voxneus_amd.optim import NetAdam, SparseGridAdam, cosine_lr_factor, create_optimizer
    from voxneus_amd.voxneus import VoxNeuS
    model = VoxNeuS(grid_res=16, hash_log2_size=10)
    opts = create_optimizer(model, lr_sdf=0.1, lr_hash=1e-2, lr_mlp=1e-3, lr_variance=5e-4)
    sdf, net = opts["sdf_optimizer"], opts["net_optimizer"]
    assert isinstance(sdf, SparseGridAdam) and isinstance(net, NetAdam)
    (g,) = sdf.param_groups
    assert g["params"] == [model.sdf_grid.grid] and g["lr"] == 0.1 and g["betas"] == (0.9, 0.99)
    assert g["eps"] == 1e-8 and g["weight_decay"] == 0.0
    groups = net.param_groups
    assert [x["lr"] for x in groups] == [1e-2, 1e-3, 5e-4]
    assert all(x["betas"] == (0.9, 0.999) and x["eps"] == 1e-8 for x in groups)
    assert [len(x["params"]) for x in groups] == [len(list(m.parameters())) for m in
                                                 (model.radiance_hash, model.radiance_mlp, model.variance_net)]
    assert isinstance(create_optimizer(model, use_sparse_sdf=False)["sdf_optimizer"], NetAdam)
    assert cosine_lr_factor(0, 100) == 1.0
    assert abs(cosine_lr_factor(50, 100) - 0.525) < 1e-12 and abs(cosine_lr_factor(100, 100) - 0.05) < 1e-12
    with pytest.raises(NotImplementedError):
        NetAdam([model.sdf_grid.grid], weight_decay=0.1)
