"""ctypes binding of libvoxneus_hip.so (the C ABI declared in include/voxneus_hip.h).

The library is loaded from ``voxneus_amd/lib`` (built in-tree by ``voxneus_amd.build``).
There is no fallback: if the library is missing or a tensor is not on a HIP device the
call raises, so a GPU run can never silently take another code path.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libvoxneus_hip.so")

VXN_DTYPE_F32 = 0
VXN_DTYPE_BF16 = 1
VXN_MLP_FP32 = 0
VXN_MLP_BF16_MFMA = 1
VXN_GRADIENT_INTERPOLATED = 0
VXN_GRADIENT_ANALYTICAL = 1
GRADIENT_MODES = {"interpolated": VXN_GRADIENT_INTERPOLATED, "analytical": VXN_GRADIENT_ANALYTICAL}


def gradient_mode_code(mode: str) -> int:
    """models/sdf_grid.py:142-160: the mode name, case-insensitive; anything else is a ValueError."""
    try:
        return GRADIENT_MODES[mode.lower()]
    except (KeyError, AttributeError):
        raise ValueError(f"Unknown gradient mode: {mode}") from None
MAX_LEVELS = 32
MLP_PARAMS = 64 * 51 + 64 + 3 * 64 + 3
VXN_VIEW_SH = 0
VXN_VIEW_PE = 1
VXN_LOSS_HUBER = 0
VXN_LOSS_MSE = 1


def loss_type_code(loss_type: str) -> int:
    """The trainer's loss_type ("Huber" or anything else = MSE, trainer.py:59,216,233)."""
    return VXN_LOSS_HUBER if str(loss_type).lower() == "huber" else VXN_LOSS_MSE
MLP_PARAMS_PE = 64 * 62 + 64 + 3 * 64 + 3


def view_embed_code(view_embed_type: str) -> int:
    """VoxNeuS view_embed_type ("sh" or "pe", models/voxneus.py:78-82) -> VXN_VIEW_*."""
    t = str(view_embed_type).lower()
    if t == "sh":
        return VXN_VIEW_SH
    if t == "pe":
        return VXN_VIEW_PE
    raise ValueError(f"unknown view_embed_type {view_embed_type!r} (expected 'sh' or 'pe')")


def mlp_params(view_embed: int) -> int:
    return MLP_PARAMS_PE if view_embed == VXN_VIEW_PE else MLP_PARAMS

c_f = ctypes.c_float
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_vp = ctypes.c_void_p


class SamplerArgs(ctypes.Structure):
    _fields_ = [
        ("rays_o", c_vp), ("rays_d", c_vp), ("near", c_vp), ("far", c_vp), ("n_rays", c_i64),
        ("grid", c_vp), ("X", c_i32), ("Y", c_i32), ("Z", c_i32), ("box", c_f * 6),
        ("occ_mask", c_vp), ("occ_res", c_i32), ("occ_box", c_f * 6),
        ("n_samples", c_i32), ("n_importance", c_i32), ("up_sample_steps", c_i32), ("perturb", c_i32),
        ("t_lin", c_vp), ("u_det", c_vp), ("t_rand", c_vp), ("u_rand", c_vp),
        ("z_vals", c_vp), ("dists", c_vp), ("mid_z", c_vp), ("inds", c_vp), ("bound_radius", c_f),
        ("vr_bitmask", c_vp), ("vr_cells", c_vp),
    ]


class RenderArgs(ctypes.Structure):
    _fields_ = [
        ("rays_o", c_vp), ("rays_d", c_vp), ("mid_z", c_vp), ("dists", c_vp), ("n_rays", c_i64),
        ("n_samples", c_i32), ("grid", c_vp), ("X", c_i32), ("Y", c_i32), ("Z", c_i32), ("box", c_f * 6),
        ("table", c_vp), ("table_dtype", c_i32), ("n_levels", c_i32), ("log2_T", c_i32), ("n_features", c_i32),
        ("level_res", c_i32 * MAX_LEVELS), ("level_w", c_f * MAX_LEVELS),
        ("mlp", c_vp), ("mlp_mode", c_i32), ("s", c_f), ("sphere_radius", c_f), ("cos_anneal_ratio", c_f),
        ("use_iter_cos", c_i32),
        ("rgb", c_vp), ("depth", c_vp), ("normal", c_vp), ("opacity", c_vp), ("weights", c_vp), ("alpha", c_vp),
        ("transmittance", c_vp), ("sdf_samples", c_vp), ("grad_samples", c_vp), ("normal_samples", c_vp),
        ("corner_indices", c_vp),
        ("d_rgb", c_vp), ("d_depth", c_vp), ("d_normal", c_vp), ("d_opacity", c_vp), ("d_weights", c_vp),
        ("target_rgb", c_vp), ("huber_delta", c_f),
        ("grad_grid", c_vp), ("grad_table", c_vp), ("grad_mlp", c_vp), ("grad_s", c_vp), ("loss", c_vp),
        ("vr_bitmask", c_vp), ("workspace", c_vp), ("workspace_bytes", c_i64),
        ("kernel_events", c_vp), ("gradient_mode", c_i32), ("fork_event", c_vp), ("s_dev", c_vp),
        ("view_embed", c_i32), ("loss_type", c_i32), ("target_mask", c_vp), ("mask_weight", c_f),
        ("loss_parts", c_vp), ("scatter_stream", c_vp), ("scatter_events", c_vp),
        ("variance", c_vp), ("s_max", c_f), ("s_out", c_vp), ("d_variance", c_vp), ("s_amp_k", c_f),
        ("amplified", c_vp), ("grad_table_overwrite", c_i32), ("grid_binned", c_i32), ("grid_wait_event", c_vp),
    ]


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("grid", c_vp), ("X", c_i32), ("Y", c_i32), ("Z", c_i32), ("box", c_f * 6),
        ("occ_mask", c_vp), ("occ_res", c_i32), ("occ_box", c_f * 6),
        ("table", c_vp), ("table_dtype", c_i32), ("n_levels", c_i32), ("log2_T", c_i32), ("n_features", c_i32),
        ("level_res", c_i32 * MAX_LEVELS), ("level_w", c_f * MAX_LEVELS),
        ("mlp", c_vp), ("s", c_f), ("sphere_radius", c_f), ("gradient_mode", c_i32),
    ]


class ServeRenderArgs(ctypes.Structure):
    _fields_ = [
        ("scenes", c_vp), ("n_scenes", c_i32), ("ray_scene", c_vp), ("n_rays", c_vp), ("capacity", c_i64),
        ("rays_o", c_vp), ("rays_d", c_vp),
        ("n_samples", c_i32), ("n_importance", c_i32), ("up_sample_steps", c_i32),
        ("t_lin", c_vp), ("u_det", c_vp), ("bound_radius", c_f), ("mlp_mode", c_i32),
        ("cos_anneal_ratio", c_f), ("use_iter_cos", c_i32),
        ("rgb", c_vp), ("depth", c_vp), ("opacity", c_vp), ("normal", c_vp),
        ("workspace", c_vp), ("workspace_bytes", c_i64),
    ]


class CompositeArgs(ctypes.Structure):
    _fields_ = [
        ("n_rays", c_i64), ("n_samples", c_i32), ("colors", c_vp), ("sdf", c_vp), ("normals", c_vp),
        ("ray_dirs", c_vp), ("dirs_per_sample", c_i32), ("z_vals", c_vp), ("dists", c_vp), ("s", c_vp),
        ("pts", c_vp), ("alpha_in", c_vp), ("sphere_radius", c_f), ("cos_anneal_ratio", c_f),
        ("use_iter_cos", c_i32),
        ("rgb", c_vp), ("depth", c_vp), ("normal", c_vp), ("opacity", c_vp), ("weights", c_vp), ("alpha", c_vp),
        ("transmittance", c_vp),
        ("d_rgb", c_vp), ("d_depth", c_vp), ("d_normal", c_vp), ("d_opacity", c_vp), ("d_weights", c_vp),
        ("d_alpha", c_vp), ("d_transmittance", c_vp),
        ("d_colors", c_vp), ("d_sdf", c_vp), ("d_normals", c_vp), ("d_alpha_in", c_vp), ("d_s", c_vp),
    ]


class AdamTensor(ctypes.Structure):
    _fields_ = [
        ("param", c_vp), ("grad", c_vp), ("exp_avg", c_vp), ("exp_avg_sq", c_vp), ("n", c_i64),
        ("lr", ctypes.c_double), ("lr_flag", c_vp), ("lr_flag_scale", ctypes.c_double), ("param_min", c_f),
        ("param_bf16", c_vp),
    ]


# name -> (restype, argtypes); every entry point returns int except the two strings
_SIGS = {
    "vxn_version": (ctypes.c_char_p, []),
    "vxn_last_error": (ctypes.c_char_p, []),
    "vxn_grid_query_fwd": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32,
                                   c_vp]),
    "vxn_grid_query_bwd": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_i32, c_vp]),
    "vxn_gradient_interp_fwd": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_i64, c_vp, c_vp, c_vp]),
    "vxn_gradient_interp_bwd": (c_i32, [c_vp, c_vp, c_i64, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp]),
    "vxn_analytical_gradient_fwd": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "vxn_analytical_gradient_bwd": (c_i32, [c_vp, c_vp, c_i64, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp]),
    "vxn_voxel_corners": (c_i32, [c_vp, c_i64, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp]),
    "vxn_lattice_query": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32,
                                  c_vp]),
    "vxn_mc_case_table": (c_i32, [c_vp, c_vp]),
    "vxn_mc_workspace_bytes": (c_i32, [c_i32, c_i32, c_i32, c_vp]),
    "vxn_mc_count": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_f, c_vp, c_i64, c_vp, c_vp]),
    "vxn_mc_emit": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_f, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "vxn_upsample_grid": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_i32, c_i32, c_i32, c_vp]),
    "vxn_gaussian_filter_workspace_bytes": (c_i32, [c_i32, c_i32, c_i32, ctypes.POINTER(c_i64)]),
    "vxn_gaussian_filter": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_i32, c_vp, c_vp, c_i64, c_vp]),
    "vxn_occupancy_update_workspace_bytes": (c_i32, [c_i64, c_vp]),
    "vxn_occupancy_update": (c_i32, [c_vp, c_vp, c_i32, c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_f, c_f, c_f, c_i32,
                                     c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "vxn_adam_upd": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i32, c_f, c_f, c_f, c_f, c_vp]),
    "vxn_masked_adam_upd": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i32, c_f, c_f, c_f, c_f, c_vp, c_vp]),
    "vxn_sparse_adam_decay": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double, c_vp, c_vp]),
    "vxn_adam_step": (c_i32, [c_vp, c_i32, c_i64, ctypes.c_double, ctypes.c_double, ctypes.c_double, c_vp]),
    "vxn_hash_fwd": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp]),
    "vxn_hash_bwd": (c_i32, [c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp]),
    "vxn_upsample": (c_i32, [ctypes.POINTER(SamplerArgs), c_vp]),
    "vxn_render_workspace_bytes": (c_i32, [ctypes.POINTER(RenderArgs), ctypes.POINTER(c_i64)]),
    "vxn_render_forward": (c_i32, [ctypes.POINTER(RenderArgs), c_vp]),
    "vxn_render_backward": (c_i32, [ctypes.POINTER(RenderArgs), c_vp]),
    "vxn_render_train": (c_i32, [ctypes.POINTER(RenderArgs), c_vp]),
    "vxn_vr_mark_corners": (c_i32, [c_vp, c_i64, c_i32, c_i32, c_i32, c_vp, c_vp]),
    "vxn_vr_workspace_bytes": (c_i32, [c_i32, c_i32, c_i32, ctypes.POINTER(c_i64)]),
    "vxn_vr_compact": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_f, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "vxn_reg_workspace_bytes": (c_i32, [c_i32, c_i32, c_i32, ctypes.POINTER(c_i64)]),
    "vxn_reg_eik_curv": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_f, c_f, c_vp, c_vp, c_vp, c_i64,
                                 c_vp]),
    "vxn_eikonal_grad": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_i64, c_vp, c_vp, c_f, c_vp]),
    "vxn_curvature_grad": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_i64, c_vp, c_vp, c_f, c_vp]),
    "vxn_s_amplify": (c_i32, [c_vp, c_f, c_vp, c_vp]),
    "vxn_eik_clamp_add_workspace_bytes": (c_i32, [ctypes.POINTER(c_i64)]),
    "vxn_eik_clamp_add": (c_i32, [c_vp, c_vp, c_i64, c_f, c_f, c_vp, c_i64, c_vp]),
    "vxn_reg_vr_workspace_bytes": (c_i32, [c_i32, c_i32, c_i32, ctypes.POINTER(c_i64)]),
    "vxn_reg_eik_curv_vr": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_f, c_vp, c_f, c_f, c_vp, c_vp, c_vp,
                                    c_vp, c_i64, c_vp]),
    "vxn_composite_fwd": (c_i32, [ctypes.POINTER(CompositeArgs), c_vp]),
    "vxn_composite_bwd_workspace_bytes": (c_i32, [c_i64, ctypes.POINTER(c_i64)]),
    "vxn_composite_bwd": (c_i32, [ctypes.POINTER(CompositeArgs), c_vp, c_i64, c_vp]),
    "vxn_neus_alpha_fwd": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_f, c_i32, c_vp, c_vp]),
    "vxn_neus_alpha_workspace_bytes": (c_i32, [c_i64, ctypes.POINTER(c_i64)]),
    "vxn_neus_alpha_bwd": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_f, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp,
                                   c_i64, c_vp]),
    "vxn_point_mlp_fwd": (c_i32, [c_vp, c_vp, c_i32, c_vp, c_vp, c_i64, c_i32, c_vp, c_vp]),
    "vxn_point_mlp_bwd_workspace_bytes": (c_i32, [c_i64, c_i32, ctypes.POINTER(c_i64)]),
    "vxn_point_mlp_bwd": (c_i32, [c_vp, c_vp, c_i32, c_vp, c_vp, c_i64, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                  c_i64, c_vp]),
    "vxn_reg_eik_curv_vr_noalloc": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_f, c_vp, c_f, c_f, c_vp, c_vp,
                                            c_vp, c_vp]),
    "vxn_serve_scene_bytes": (c_i32, [ctypes.POINTER(c_i64)]),
    "vxn_serve_pack_scene": (c_i32, [ctypes.POINTER(SceneDesc), c_vp, c_i64]),
    "vxn_serve_render_workspace_bytes": (c_i32, [ctypes.POINTER(ServeRenderArgs), ctypes.POINTER(c_i64)]),
    "vxn_serve_render": (c_i32, [ctypes.POINTER(ServeRenderArgs), c_vp]),
    "vxn_serve_query": (c_i32, [c_vp, c_i32, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "vxn_serve_move": (c_i32, [c_vp, c_vp, c_i32, c_i64, c_i64, c_i32, c_vp, c_vp, c_i32, c_vp, c_vp, c_i32, c_vp,
                               c_vp, c_vp]),
}

_lib = None


def exported_symbols() -> Sequence[str]:
    return list(_SIGS)


def lib():
    """Load (once) and return the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m voxneus_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class HipError(RuntimeError):
    pass


class SideStream:
    """A second stream (default priority) on ``device`` and ``n_events`` (at least 2)
    untimed HIP events: ``fork(stream)`` starts side work after what ``stream`` has queued
    (``events[0]`` can also be recorded by a render launch as its ``fork_event``),
    ``finish()`` marks its end and ``join(stream)`` makes ``stream`` wait for that mark."""

    _hip = None

    def __init__(self, device, n_events: int = 2):
        if SideStream._hip is None:
            h = ctypes.CDLL("libamdhip64.so.7")
            h.hipEventCreateWithFlags.argtypes = [ctypes.POINTER(c_vp), ctypes.c_uint]
            h.hipStreamWaitEvent.argtypes = [c_vp, c_vp, ctypes.c_uint]
            h.hipEventDestroy.argtypes = [c_vp]
            h.hipEventRecord.argtypes = [c_vp, c_vp]
            SideStream._hip = h
        self.stream = torch.cuda.Stream(device=device)
        self.events = (c_vp * max(2, n_events))()
        for i in range(max(2, n_events)):
            ev = c_vp()
            if SideStream._hip.hipEventCreateWithFlags(ctypes.byref(ev), 2) != 0:  # hipEventDisableTiming
                raise HipError("hipEventCreateWithFlags failed")
            self.events[i] = ev

    @property
    def handle(self) -> int:
        return self.stream.cuda_stream

    def fork(self, stream: Optional[int] = None):
        """Make the side stream wait for the work queued on ``stream`` so far (or, with
        None, for the last record of ``events[0]``, e.g. by a render launch)."""
        h = SideStream._hip
        if stream is not None and h.hipEventRecord(self.events[0], c_vp(stream)) != 0:
            raise HipError("hipEventRecord failed")
        if h.hipStreamWaitEvent(c_vp(self.handle), self.events[0], 0) != 0:
            raise HipError("hipStreamWaitEvent failed")

    def finish(self):
        """Mark the end of the side work queued so far (``join`` waits for this mark)."""
        if SideStream._hip.hipEventRecord(self.events[1], c_vp(self.handle)) != 0:
            raise HipError("hipEventRecord failed")

    def join(self, stream: int):
        if SideStream._hip.hipStreamWaitEvent(c_vp(stream), self.events[1], 0) != 0:
            raise HipError("hipStreamWaitEvent failed")


class KernelEvents:
    """``n`` HIP timing events for ``vxn_render_args.kernel_events``.

    The events come from the HIP runtime torch has loaded (the library and torch share it
    through the ``libamdhip64.so.7`` soname), so they are recorded on torch's streams.
    """

    _hip = None

    def __init__(self, n: int = 6):
        if KernelEvents._hip is None:
            h = ctypes.CDLL("libamdhip64.so.7")
            h.hipEventCreate.argtypes = [ctypes.POINTER(c_vp)]
            h.hipEventElapsedTime.argtypes = [ctypes.POINTER(c_f), c_vp, c_vp]
            h.hipEventSynchronize.argtypes = [c_vp]
            h.hipEventDestroy.argtypes = [c_vp]
            h.hipEventRecord.argtypes = [c_vp, c_vp]
            KernelEvents._hip = h
        self.n = n
        self.array = (c_vp * n)()
        for i in range(n):
            ev = c_vp()
            if KernelEvents._hip.hipEventCreate(ctypes.byref(ev)) != 0:
                raise HipError("hipEventCreate failed")
            self.array[i] = ev

    def record(self, i: int, stream: int):
        """Record event i on a raw hipStream_t (torch's ``stream.cuda_stream``)."""
        if KernelEvents._hip.hipEventRecord(self.array[i], c_vp(stream)) != 0:
            raise HipError("hipEventRecord failed")

    def elapsed_ms(self, i: int, j: int) -> float:
        """Time between events i and j (waits for j)."""
        h = KernelEvents._hip
        if h.hipEventSynchronize(self.array[j]) != 0:
            raise HipError("hipEventSynchronize failed")
        ms = c_f(0.0)
        if h.hipEventElapsedTime(ctypes.byref(ms), self.array[i], self.array[j]) != 0:
            raise HipError("hipEventElapsedTime failed")
        return float(ms.value)

    def __del__(self):
        # at interpreter shutdown the class attribute may already be gone
        h = getattr(KernelEvents, "_hip", None) if KernelEvents is not None else None
        if h is not None:
            for ev in self.array:
                if ev:
                    h.hipEventDestroy(ev)


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().vxn_last_error().decode(errors="replace")
        raise HipError(f"{what or 'voxneus_hip'} failed ({rc}): {msg}")


def stream_ptr(device: Optional[torch.device] = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def dptr(t: Optional[torch.Tensor], dtype: Optional[torch.dtype] = None, name: str = "tensor") -> Optional[int]:
    """Device pointer of a contiguous HIP tensor (None -> null)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError(f"{name} must be on a HIP device (got {t.device})")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t.data_ptr()


def box6(*box):
    """Host float[6] {xyz_min, xyz_max} from (min3, max3), a flat 6-sequence or 6 floats."""
    if len(box) == 2:
        vals = list(box[0]) + list(box[1])
    elif len(box) == 1:
        vals = list(box[0])
    else:
        vals = list(box)
    if len(vals) != 6:
        raise ValueError(f"a box needs 6 values (got {len(vals)})")
    return (c_f * 6)(*[float(v) for v in vals])


def f3(vals):
    return (c_f * 3)(*[float(v) for v in vals])
