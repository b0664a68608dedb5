"""Checkpoints (training/checkpoint.py:11-63) and the checkpoint -> resident scene path of
the serving loop.

``save_checkpoint`` / ``load_checkpoint`` keep the reference's file layout (``step``,
``model_state`` with the reference's state-dict keys, ``config``, ``metrics``,
``grid_res``, ``<name>_state`` per optimizer) and its grid-rescale-on-load rule.  Files
are read with ``torch.load(weights_only=True)``: tensors, numbers, strings and containers
load, nothing in the file is executed.

``load_resident_scene`` turns a checkpoint straight into a ``serving.ResidentScene``: the
grid, hash table, MLP, sharpness and occupancy mask are memory-mapped from the file and
copied once into HBM, without building a model object.
"""

from __future__ import annotations

import os
from typing import Any, Dict, Optional, Union

import torch

KEY_GRID = "sdf_grid.grid"
KEY_TABLE = "radiance_hash.encoder.embeddings"
KEY_MLP = ("radiance_mlp.net.0.weight", "radiance_mlp.net.0.bias", "radiance_mlp.net.2.weight",
           "radiance_mlp.net.2.bias")
KEY_VARIANCE = "variance_net.variance"
KEY_OCC = "occupancy_grid.occupancy_mask"


def save_checkpoint(filepath: str, step: int, model: torch.nn.Module, optimizers: Dict[str, Any],
                    config: Dict[str, Any], metrics: Optional[Dict[str, float]] = None) -> None:
    """checkpoint.py:11-33."""
    d = os.path.dirname(filepath)
    if d:
        os.makedirs(d, exist_ok=True)
    ck = {"step": step, "model_state": model.state_dict(), "config": config, "metrics": metrics or {},
          "grid_res": [int(x) for x in model.sdf_grid.res]}
    for name, opt in optimizers.items():
        ck[f"{name}_state"] = opt.state_dict()
    torch.save(ck, filepath)


def _read(filepath: str, device=None) -> Dict[str, Any]:
    if not os.path.exists(filepath):
        raise FileNotFoundError(f"Checkpoint not found at: {filepath}")
    return torch.load(filepath, map_location=device or "cpu", weights_only=True)


def load_checkpoint(filepath: str, model: torch.nn.Module, optimizers: Optional[Dict[str, Any]] = None,
                    device=None) -> Dict[str, Any]:
    """checkpoint.py:36-63: rescale the model's grid to the saved resolution if it differs
    (on the device), then load the model and optimizer states."""
    ck = _read(filepath, device)
    saved = ck.get("grid_res")
    if saved is not None and [int(x) for x in saved] != [int(x) for x in model.sdf_grid.res]:
        model.sdf_grid.scale_volume_grid([int(x) for x in saved])
    model.load_state_dict(ck["model_state"])
    if optimizers is not None:
        for name, opt in optimizers.items():
            key = f"{name}_state"
            if key in ck:
                opt.load_state_dict(ck[key])
    return ck


def load_resident_scene(source: Union[str, Dict[str, Any]], device, table_dtype: torch.dtype = torch.float32,
                        hash_base_resolution: int = 16, hash_max_resolution: float = 2048.0,
                        s_max: Optional[float] = None, gradient_mode: Optional[str] = None,
                        sphere_radius: float = 1.0, use_occupancy: bool = True):
    """A serving scene from a checkpoint file (or an already-loaded checkpoint / state dict).

    The reference's VoxNeuS builds its hash grid with base resolution 16 and finest 2048
    (models/radiance_hash.py:175-199, not stored in the state dict); ``s_max`` and
    ``gradient_mode`` default to the checkpoint's ``config["model"]`` entries, else 300 and
    "interpolated" (scripts/train.py:78-93)."""
    from .hash_grid import level_resolutions
    from .serving import ResidentScene

    if isinstance(source, str):
        if not os.path.exists(source):
            raise FileNotFoundError(f"Checkpoint not found at: {source}")
        ck = torch.load(source, map_location="cpu", weights_only=True, mmap=True)
    else:
        ck = source
    st = ck.get("model_state", ck)
    mcfg = (ck.get("config") or {}).get("model", {}) if isinstance(ck.get("config"), dict) else {}
    missing = [k for k in (KEY_GRID, KEY_TABLE, KEY_VARIANCE, *KEY_MLP) if k not in st]
    if missing:
        raise KeyError(f"checkpoint lacks {missing}")
    grid = st[KEY_GRID]
    if grid.dim() != 5 or grid.shape[:2] != (1, 1):
        raise ValueError(f"{KEY_GRID} must be [1,1,X,Y,Z] (got {tuple(grid.shape)})")
    if st.get("sdf_grid.feat_grid") is not None:
        raise NotImplementedError("geometry feature grids (NeuS2 feat_dim > 0) are outside this path")
    table = st[KEY_TABLE]
    L, T, F = table.shape
    W0, b0, W1, b1 = (st[k] for k in KEY_MLP)
    if tuple(W0.shape) == (64, 62):
        raise NotImplementedError("serving scenes use the SH view embedding (51 -> 64 -> 3 MLP); this "
                                  "checkpoint's MLP takes the positional encoding")
    if tuple(W0.shape) != (64, 51) or tuple(W1.shape) != (3, 64):
        raise NotImplementedError("the fused render implements the 51 -> 64 -> 3 radiance MLP")
    mn = st.get("sdf_grid.xyz_min", torch.tensor([-1.0, -1.0, -1.0]))
    mx = st.get("sdf_grid.xyz_max", torch.tensor([1.0, 1.0, 1.0]))
    box = tuple(float(v) for v in torch.cat([mn.reshape(3), mx.reshape(3)]).tolist())
    smax = float(mcfg.get("s_max", 300.0) if s_max is None else s_max)
    # variance.py:31, evaluated on the device as the model's own s is
    v = st[KEY_VARIANCE].detach().to(device=device, dtype=torch.float32).reshape(())
    s = float(torch.clamp(torch.exp(v * 10.0), min=1e-4, max=smax))
    mode = gradient_mode if gradient_mode is not None else mcfg.get("gradient_mode", "interpolated")
    occ = occ_box = None
    if use_occupancy and KEY_OCC in st:
        occ = st[KEY_OCC].to(device=device, dtype=torch.bool).contiguous()
        omn = st.get("occupancy_grid.xyz_min", mn)
        omx = st.get("occupancy_grid.xyz_max", mx)
        occ_box = tuple(float(x) for x in torch.cat([omn.reshape(3), omx.reshape(3)]).tolist())
    flat = torch.cat([W0.reshape(-1), b0.reshape(-1), W1.reshape(-1), b1.reshape(-1)]).to(torch.float32)
    res = level_resolutions(L, hash_base_resolution, hash_max_resolution)
    scene = ResidentScene(grid.reshape(grid.shape[2:]).to(device=device, dtype=torch.float32).contiguous(),
                          table.to(device=device).to(table_dtype).contiguous(), flat.to(device), res, s, box,
                          occ_mask=occ, occ_box=occ_box, sphere_radius=sphere_radius, gradient_mode=mode)
    return scene
