"""The trainer's schedules (training/scheduler.py): the 3-stage eikonal and curvature
weights, the NeuS2 progressive hash bandwidth, and the s-gradient amplification.

The weights and bandwidth are host arithmetic evaluated once per step; the amplification
of the fused step runs on the device (``vxn_s_amplify``), and ``amplify_s_gradient`` here
is the module-level entry of the reference for a parameter's ``.grad``.
"""

from __future__ import annotations

from typing import Tuple

from .training import amplify_s_gradient  # noqa: F401  (scheduler.py:14-33)


class VoxNeuSScheduler:
    """scheduler.py:36-114: the regulariser weights of step ``step`` and the NeuS2 lambda."""

    def __init__(self, total_steps: int = 40000, stage1_end: int = 11000, stage2_end: int = 21000,
                 w_eik_init: float = 1e-2, w_eik_final: float = 1e-3, w_curv_init: float = 1e-8,
                 w_curv_peak: float = 5e-6, w_curv_final: float = 5e-7):
        self.total_steps = total_steps
        self.stage1_end = stage1_end
        self.stage2_end = stage2_end
        self.w_eik_init = w_eik_init
        self.w_eik_final = w_eik_final
        self.w_curv_init = w_curv_init
        self.w_curv_peak = w_curv_peak
        self.w_curv_final = w_curv_final

    def get_w_eik(self, step: int) -> float:
        """Constant, then linear to the final weight over stage 2, then constant."""
        if step <= self.stage1_end:
            return self.w_eik_init
        if step <= self.stage2_end:
            t = (step - self.stage1_end) / (self.stage2_end - self.stage1_end)
            return self.w_eik_init + t * (self.w_eik_final - self.w_eik_init)
        return self.w_eik_final

    def get_w_curv(self, step: int) -> float:
        """Constant, then linear to the peak over stage 2, then geometric to the final
        weight over the remaining steps."""
        if step <= self.stage1_end:
            return self.w_curv_init
        if step <= self.stage2_end:
            t = (step - self.stage1_end) / (self.stage2_end - self.stage1_end)
            return self.w_curv_init + t * (self.w_curv_peak - self.w_curv_init)
        t = (step - self.stage2_end) / max(self.total_steps - self.stage2_end, 1)
        return self.w_curv_peak * ((self.w_curv_final / self.w_curv_peak) ** t)

    def get_weights(self, step: int) -> Tuple[float, float]:
        return self.get_w_eik(step), self.get_w_curv(step)

    def get_lambda(self, step: int, start_lambda: float = 2.0, step_ratio: float = 0.025,
                   max_lambda: float = 16.0) -> float:
        return compute_neus2_lambda(step, self.total_steps, start_lambda, step_ratio, max_lambda)


def compute_neus2_lambda(step: int, total_steps: int = 40000, start_lambda: float = 2.0, step_ratio: float = 0.025,
                         max_lambda: float = 16.0) -> float:
    """scheduler.py:117-127: lambda grows by one level every ``step_ratio`` of training."""
    interval = max(int(total_steps * step_ratio), 1)
    return min(start_lambda + step / float(interval), max_lambda)
