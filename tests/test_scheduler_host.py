"""The trainer's schedules (training/scheduler.py:36-127) on the CPU, pinned by the values
the paper's training schedule fixes (Sec. 5.1): the eikonal weight 1e-2 until 11k, linear
to 1e-3 at 21k, then constant; the curvature weight 1e-8 until 11k, linear to 5e-6 at
21k, then geometric to 5e-7 at 40k (the midpoint of the decay is the geometric mean);
lambda = 2 + one level per 2.5 % of training, capped at 16 (NeuS2 Sec. 3.3)."""

import math

import pytest

from voxneus_amd.scheduler import VoxNeuSScheduler, amplify_s_gradient, compute_neus2_lambda


def close(a, b, rel=1e-12):
    return abs(a - b) <= rel * max(abs(a), abs(b))


def test_eikonal_weight_stages():
    s = VoxNeuSScheduler()
    for step in (0, 5000, 11000):
        assert s.get_w_eik(step) == 1e-2
    assert close(s.get_w_eik(16000), 5.5e-3)
    assert close(s.get_w_eik(13500), 1e-2 - 0.25 * 9e-3)
    assert close(s.get_w_eik(21000), 1e-3)
    for step in (21001, 30000, 40000, 50000):
        assert s.get_w_eik(step) == 1e-3
    # non-increasing over training, up to the rounding of the linear segment's end
    w = [s.get_w_eik(t) for t in range(0, 40001, 250)]
    assert all(a >= b * (1 - 1e-15) for a, b in zip(w, w[1:]))


def test_curvature_weight_stages():
    s = VoxNeuSScheduler()
    assert s.get_w_curv(0) == 1e-8 and s.get_w_curv(11000) == 1e-8
    assert close(s.get_w_curv(16000), 0.5 * (1e-8 + 5e-6))
    assert close(s.get_w_curv(21000), 5e-6)
    assert close(s.get_w_curv(30500), math.sqrt(5e-6 * 5e-7))
    assert close(s.get_w_curv(40000), 5e-7)
    # geometric decay: equal ratios over equal step intervals
    r1 = s.get_w_curv(25000) / s.get_w_curv(23000)
    r2 = s.get_w_curv(37000) / s.get_w_curv(35000)
    assert close(r1, r2, 1e-9) and r1 < 1.0
    assert s.get_weights(16000) == (s.get_w_eik(16000), s.get_w_curv(16000))


def test_custom_stages_and_degenerate_total():
    s = VoxNeuSScheduler(total_steps=100, stage1_end=10, stage2_end=100, w_curv_peak=1e-4, w_curv_final=1e-6)
    assert close(s.get_w_curv(100), 1e-4)
    assert close(s.get_w_curv(101), 1e-4 * (1e-2 ** 1.0))  # the remaining span is clamped to one step


def test_neus2_lambda():
    s = VoxNeuSScheduler()
    assert s.get_lambda(0) == 2.0
    assert s.get_lambda(500) == 2.5
    assert s.get_lambda(1000) == 3.0
    assert s.get_lambda(14000) == 16.0
    assert s.get_lambda(20000) == 16.0
    assert s.get_lambda(1000, start_lambda=1.0, max_lambda=2.5) == 2.0
    assert compute_neus2_lambda(10, total_steps=100) == 2.0 + 10 / 2.0
    assert compute_neus2_lambda(3, total_steps=10) == 5.0  # an interval below one step is one step
    assert compute_neus2_lambda(7000) == s.get_lambda(7000)


def test_amplify_is_the_training_module_entry():
    from voxneus_amd import training
    assert amplify_s_gradient is training.amplify_s_gradient


def test_trainer_refuses_unsupported_options():
    pytest.importorskip("torch")
    from voxneus_amd import trainer
    assert trainer._VARIANCE_FLOORS == (0.4382, 0.5010)
    # s >= 80 and s >= 150 at the floors (trainer.py:338-345)
    assert math.exp(10 * 0.4382) >= 80.0 - 0.1 and math.exp(10 * 0.5010) >= 150.0 - 0.1
