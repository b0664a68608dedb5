"""The post-upsample eikonal clamp of the oracle (trainer.py:309-316), pinned by what the
clamp must do rather than by its formula: the added term never exceeds the clamp in
max-abs, it is a positive multiple of the weighted gradient, the threshold is strict,
NaN leaves it unscaled as torch's max and comparison do, and outside the cooldown it is
the plain weighted add.  Also the host side of the progressive super-resolution
schedule (trainer.py:131-164)."""

import math

import pytest
import torch

from oracle import ref_cpu as R


def _g(n, seed, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(n, generator=g) * scale


def test_clamp_bounds_the_added_term():
    g = _g(5000, 1, 3.0)
    for w, clamp in ((0.1, 0.1), (1.0, 0.5), (0.37, 1e-3)):
        out = R.eik_clamp_add(torch.zeros(5000), g, w, clamp)
        m = float(out.abs().max())
        assert m <= clamp * (1 + 2 ** -22) and m >= clamp * (1 - 2 ** -22)
        # direction and sign kept: a positive multiple of w * g, in float64
        ratio = out.double() / (g.double() * w)
        c = clamp / float((g.double() * w).abs().max())
        assert torch.allclose(ratio, torch.full_like(ratio, c), rtol=4e-7, atol=0)


def test_unclamped_is_the_weighted_add():
    g = _g(3000, 2, 1e-3)
    grad = _g(3000, 3)
    w = 0.25  # a power of two: g * w is exact, so the sum is one rounding of grad + w g
    out = R.eik_clamp_add(grad, g, w, clamp=1.0)
    ref = (grad.double() + g.double() * 0.25).to(torch.float32)
    assert torch.equal(out, ref)
    # outside the cooldown the clamp does nothing, however large the gradient
    big = g * 1e6
    assert torch.equal(R.eik_clamp_add(grad, big, w, 1e-3, cooldown=False), R.eik_clamp_add(grad, big, w, 1e9))


def test_threshold_is_strict_in_float32():
    g = torch.tensor([0.5, -1.0, 0.25])
    # max |g w| equals float32(0.1) exactly: not above the float32 clamp, so unscaled
    w = float(torch.tensor(0.1, dtype=torch.float32))
    out = R.eik_clamp_add(torch.zeros(3), g, w, 0.1)
    assert torch.equal(out, g * w)
    assert float((g * w).abs().max()) == float(torch.tensor(0.1, dtype=torch.float32))


def test_nan_maximum_leaves_the_gradient_unscaled():
    g = _g(100, 4, 10.0)
    g[17] = math.nan
    out = R.eik_clamp_add(torch.zeros(100), g, 1.0, 0.1)
    keep = torch.ones(100, dtype=torch.bool)
    keep[17] = False
    assert torch.equal(out[keep], g[keep]) and math.isnan(float(out[17]))


def test_super_resolution_schedule_milestones():
    from voxneus_amd.training import SuperResolutionSchedule

    class _St:
        def __init__(self):
            self.res, self.post_upsample_cooldown = 96, 0

        def rescale_grid(self, r):
            self.res = r

    class _Opt:
        lr = None

        def reset_sdf(self, lr):
            self.lr = lr

    sch = SuperResolutionSchedule()
    st, opt = _St(), _Opt()
    assert not sch.apply(9999, st, opt) and st.res == 96 and opt.lr is None
    assert sch.apply(10000, st, opt) and st.res == 160 and opt.lr == 0.05 and st.post_upsample_cooldown == 500
    assert sch.apply(30000, st, opt) and st.res == 320 and opt.lr == 0.02
    # more resolutions than stage rates: the last rate holds (trainer.py:138)
    sch2 = SuperResolutionSchedule(milestones=(5, 6, 7), resolutions=(8, 16, 24, 32), lr_sdf_stages=(0.1, 0.05))
    st2, opt2 = _St(), _Opt()
    for k in (5, 6, 7):
        assert sch2.apply(k, st2, opt2)
    assert st2.res == 32 and opt2.lr == 0.05
    with pytest.raises(IndexError):
        SuperResolutionSchedule(milestones=(1,), resolutions=(8,)).apply(1, _St())


def test_scalar_over_tensor_is_reciprocal_times_scalar():
    # the reference writes clamp / norm with a Python float over a 0-d float32 tensor; the
    # oracle restates it as norm.reciprocal() * clamp, which is what PyTorch evaluates (it
    # differs from a correctly rounded division for some inputs)
    x = torch.rand(4000, generator=torch.Generator().manual_seed(5)) * 10 + 1e-3
    assert all(torch.equal(0.1 / x[i], x[i].reciprocal() * 0.1) for i in range(4000))
    assert not torch.equal(0.1 / x, (torch.tensor(0.1, dtype=torch.float64) / x.double()).float())
