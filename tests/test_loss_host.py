"""The training colour loss variants of the oracle (trainer.py:213-245), pinned by what
they must equal: the masked loss with an all-ones mask is the plain mean, with a mask
selecting some rays it is the plain mean over those rays, with an all-zero mask the plain
mean; MSE is the mean squared error; the mask BCE has its closed form and its weight."""

import math

import torch

from oracle import ref_cpu as R


def _data(n=50, seed=0):
    g = torch.Generator().manual_seed(seed)
    return torch.rand(n, 3, generator=g), torch.rand(n, 3, generator=g), torch.rand(n, 1, generator=g)


def test_mask_of_ones_and_zeros_is_the_plain_mean():
    pred, tgt, op = _data()
    for lt in ("Huber", "mse"):
        plain = R.color_loss(pred, tgt, lt)[1]
        for m in (torch.ones(50, 1), torch.zeros(50, 1)):
            assert torch.allclose(R.color_loss(pred, tgt, lt, True, m, op)[1], plain, rtol=1e-6, atol=0)


def test_mask_selects_rays():
    pred, tgt, op = _data(seed=1)
    m = torch.zeros(50, 1)
    sel = [3, 7, 8, 20, 41]
    m[sel] = 1.0
    for lt in ("Huber", "mse"):
        ref = R.color_loss(pred[sel], tgt[sel], lt)[1]
        assert torch.allclose(R.color_loss(pred, tgt, lt, True, m, op)[1], ref, rtol=1e-6, atol=0)


def test_mse_and_huber_closed_forms():
    pred, tgt, _ = _data(seed=2)
    assert torch.allclose(R.color_loss(pred, tgt, "MSE")[0], ((pred - tgt) ** 2).mean())
    d = (pred - tgt).abs().double()
    ref = torch.where(d < 0.1, 0.5 * d * d / 0.1, d - 0.05).mean()
    assert abs(float(R.color_loss(pred, tgt, "Huber")[0]) - float(ref)) < 1e-7


def test_mask_bce_closed_form_and_weight():
    pred, tgt, _ = _data(3)
    op = torch.tensor([[0.5], [0.25], [1.0]])
    m = torch.tensor([[1.0], [0.0], [1.0]])
    total, rgb, lm = R.color_loss(pred, tgt, "Huber", True, m, op, mask_weight=0.2)
    x = [0.5, 0.25, 1.0 - 1e-6]
    ref = (-math.log(x[0]) - math.log(1.0 - x[1]) - math.log(float(torch.tensor(x[2], dtype=torch.float32)))) / 3
    assert abs(float(lm) - ref) < 1e-6
    assert abs(float(total) - (float(rgb) + 0.2 * float(lm))) < 1e-7
    assert R.color_loss(pred, tgt, "Huber", False, m, op)[2] is None
