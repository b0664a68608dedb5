"""Continuous-batching render and SDF/normal-query serving of many resident scenes.

The reference renders one scene at a time in 2048-ray chunks (``scripts/render.py:35-54``:
``near_far_from_sphere`` + ``VoxNeuS.render`` under ``no_grad`` with the deterministic
sampler) and answers SDF/normal queries in 65536-point batches
(``geometry/mesh.py:80-114``); it has no request loop, batching or latency accounting.

Here every scene's grid, hash table, MLP and occupancy grid stay resident in HBM and are
described to the kernels by one packed record each in a device table.  A request's rays
(or points) live in a contiguous range of a device arena, which is the admission pool:
a request is admitted when a range of its size is free.  Each step packs up to
``ray_capacity`` rays of the admitted render requests (oldest first, any mix of scenes)
into one batch and runs the serving kernels over it: segment move -> sampler -> fused
render forward -> move back; query requests get the same treatment with the point
query kernel.  The step's sizes live in device memory, so each step kind is captured
once as a HIP graph and replayed.  Requests that arrive while others run join at the
next step.  Per-ray work does not depend on the other rays of a batch, so a request's
result is bit-identical alone and in a mix.
"""

from __future__ import annotations

import ctypes
import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _native as N


# ----------------------------------------------------------------------------------------
# Scenes
# ----------------------------------------------------------------------------------------

class ResidentScene:
    """One trained scene kept in HBM: grid [X,Y,Z] f32, hash table [16,T,2] f32/bf16,
    flat MLP [3523] f32, NeuS sharpness s, occupancy mask (bool [r,r,r] or None) and the
    model's gradient mode ("interpolated" Eq. 7 or "analytical" Eq. 4)."""

    def __init__(self, grid: torch.Tensor, table: torch.Tensor, mlp: torch.Tensor, level_res: Sequence[int],
                 s: float, box=(-1.0, -1.0, -1.0, 1.0, 1.0, 1.0), occ_mask: Optional[torch.Tensor] = None,
                 occ_box=None, level_w: Optional[Sequence[float]] = None, sphere_radius: float = 1.0,
                 gradient_mode: str = "interpolated"):
        if grid.dim() != 3 or grid.dtype != torch.float32 or not grid.is_cuda:
            raise ValueError("grid must be a float32 [X,Y,Z] HIP tensor")
        if table.dim() != 3 or table.shape[0] != 16 or table.shape[2] != 2 or not table.is_cuda:
            raise ValueError("table must be a [16,T,2] HIP tensor")
        if table.shape[1] < 2 or table.shape[1] & (table.shape[1] - 1):
            raise ValueError(f"the hash table size T must be a power of two (got {table.shape[1]})")
        if table.dtype not in (torch.float32, torch.bfloat16):
            raise ValueError(f"table must be float32 or bfloat16 (got {table.dtype})")
        if len(level_res) != 16 or (level_w is not None and len(level_w) != 16):
            raise ValueError("level_res (and level_w when given) must hold one value per level (16)")
        self.grid = grid.contiguous()
        self.table = table.contiguous()
        if mlp.numel() != N.MLP_PARAMS:
            raise ValueError(f"a serving scene's MLP is the SH-embedding 51 -> 64 -> 3 net ({N.MLP_PARAMS} "
                             f"parameters; got {mlp.numel()})")
        self.mlp = mlp.contiguous().to(torch.float32)
        self.level_res = [int(r) for r in level_res]
        self.level_w = None if level_w is None else [float(w) for w in level_w]
        self.s = float(s)
        self.box = tuple(float(v) for v in box)
        self.occ_mask = None if occ_mask is None else occ_mask.contiguous().to(torch.bool)
        self.occ_box = self.box if occ_box is None else tuple(float(v) for v in occ_box)
        self.sphere_radius = float(sphere_radius)
        self.gradient_mode = N.gradient_mode_code(gradient_mode)

    @property
    def nbytes(self) -> int:
        n = self.grid.numel() * 4 + self.table.numel() * self.table.element_size() + self.mlp.numel() * 4
        return n + (self.occ_mask.numel() if self.occ_mask is not None else 0)

    def desc(self) -> "N.SceneDesc":
        d = N.SceneDesc()
        d.grid = self.grid.data_ptr()
        d.X, d.Y, d.Z = self.grid.shape
        d.box = N.box6(self.box)
        if self.occ_mask is not None:
            d.occ_mask, d.occ_res = self.occ_mask.data_ptr(), int(self.occ_mask.shape[0])
        d.occ_box = N.box6(self.occ_box)
        d.table = self.table.data_ptr()
        d.table_dtype = N.VXN_DTYPE_BF16 if self.table.dtype == torch.bfloat16 else N.VXN_DTYPE_F32
        d.n_levels, d.n_features = 16, 2
        d.log2_T = int(self.table.shape[1]).bit_length() - 1
        for l, r in enumerate(self.level_res):
            d.level_res[l] = r
            d.level_w[l] = 1.0 if self.level_w is None else self.level_w[l]
        d.mlp = self.mlp.data_ptr()
        d.s, d.sphere_radius = self.s, self.sphere_radius
        d.gradient_mode = self.gradient_mode
        return d


def make_resident_scene(grid_res: int, log2_T: int, seed: int, device, table_dtype=torch.float32,
                        occ_res: int = 128, per_level_scale: float = 1.5, variance: float = 0.3,
                        gradient_mode: str = "interpolated") -> ResidentScene:
    """A seeded multi-primitive scene built on the device: the SDF of
    ``scenes.MultiPrimitive.seeded(seed)`` at the grid vertices (float64, stored float32),
    a U(-1e-4, 1e-4) hash table, a kaiming-uniform MLP, an all-True occupancy grid."""
    from . import scenes as S
    from .hash_grid import level_resolutions

    shape = S.MultiPrimitive.seeded(seed)
    ax = torch.linspace(-1.0, 1.0, grid_res, dtype=torch.float32, device=device).double()
    p = torch.stack(torch.meshgrid(ax, ax, ax, indexing="ij"), -1)
    d = None
    for c, r in zip(shape.centres, shape.radii):
        ds = torch.linalg.norm(p - torch.as_tensor(c, device=device), dim=-1) - float(r)
        d = ds if d is None else _smooth_min(d, ds, shape.k)
    q = torch.stack([torch.hypot(p[..., 0], p[..., 1]) - shape.torus_R, p[..., 2]], -1)
    d = _smooth_min(d, torch.linalg.norm(q, dim=-1) - shape.torus_r, shape.k)
    grid = d.to(torch.float32).contiguous()
    del p, q, d
    g = torch.Generator(device=device)
    g.manual_seed(seed + 2)
    table = ((torch.rand((16, 1 << log2_T, 2), generator=g, device=device) * 2.0 - 1.0) * 1e-4).to(table_dtype)
    mlp = S.mlp_params(51, 64, seed + 3)
    flat = torch.cat([torch.from_numpy(mlp[k]).reshape(-1) for k in ("W0", "b0", "W1", "b1")]).to(device)
    s = min(max(math.exp(variance * 10.0), 1e-4), 300.0)
    occ = torch.ones((occ_res,) * 3, dtype=torch.bool, device=device)
    res = level_resolutions(16, 16, S.hash_max_resolution(16, per_level_scale, 16))
    return ResidentScene(grid, table, flat, res, s, occ_mask=occ, gradient_mode=gradient_mode)


def _smooth_min(a, b, k):
    h = torch.clamp(k - (a - b).abs(), min=0.0) / k
    return torch.minimum(a, b) - h * h * k * 0.25


def pack_scene_table(scenes: Sequence[ResidentScene], device) -> torch.Tensor:
    """The device table of packed scene records (one per scene, in order)."""
    nb = ctypes.c_int64(0)
    N.check(N.lib().vxn_serve_scene_bytes(ctypes.byref(nb)), "serve_scene_bytes")
    host = np.zeros((len(scenes), nb.value), np.uint8)
    for i, sc in enumerate(scenes):
        d = sc.desc()
        N.check(N.lib().vxn_serve_pack_scene(ctypes.byref(d), host[i].ctypes.data, nb.value), f"pack_scene {i}")
    return torch.from_numpy(host.reshape(-1)).to(device)


# ----------------------------------------------------------------------------------------
# Requests
# ----------------------------------------------------------------------------------------

@dataclass
class RenderRequest:
    """A full image of ``scene`` from a pinhole camera (OpenCV convention, rays through
    pixel centres, normalised directions: ``rendering/rays.py:11-69``)."""

    scene: int
    c2w: np.ndarray
    W: int
    H: int
    fx: float
    fy: float
    cx: float
    cy: float
    arrival: float = 0.0

    @property
    def rows(self) -> int:
        return self.W * self.H


@dataclass
class QueryRequest:
    """SDF and unit normal of ``scene`` at ``points`` [n, 3]."""

    scene: int
    points: torch.Tensor
    arrival: float = 0.0

    @property
    def rows(self) -> int:
        return int(self.points.shape[0])


@dataclass
class Result:
    index: int
    kind: str
    scene: int
    rows: int
    arrival: float
    admitted: float = math.nan
    done: float = math.nan
    outputs: Dict[str, torch.Tensor] = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def pinhole_rays(req: RenderRequest, device) -> (torch.Tensor, torch.Tensor):
    """Rays of a render request, pixel-major (row y, then x), on the device."""
    j, i = torch.meshgrid(torch.arange(req.H, dtype=torch.float32, device=device) + 0.5,
                          torch.arange(req.W, dtype=torch.float32, device=device) + 0.5, indexing="ij")
    dirs = torch.stack([(i - req.cx) / req.fx, (j - req.cy) / req.fy, torch.ones_like(i)], -1).reshape(-1, 3)
    c2w = torch.as_tensor(np.asarray(req.c2w, np.float32), device=device)
    d = dirs @ c2w[:3, :3].T
    d = d / torch.linalg.norm(d, dim=-1, keepdim=True)
    o = c2w[:3, 3].expand(d.shape)
    return o.contiguous(), d.contiguous()


class Arena:
    """First-fit allocator of row ranges (the admission pool of one arena)."""

    def __init__(self, rows: int):
        self.rows = int(rows)
        self.free = [(0, self.rows)]  # sorted (offset, length)

    def alloc(self, n: int) -> Optional[int]:
        for k, (off, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (off + n, ln - n)
                return off
        return None

    def release(self, off: int, n: int):
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((o, ln))
        self.free = merged

    @property
    def free_rows(self) -> int:
        return sum(ln for _, ln in self.free)


def plan_step(live, cap: int, max_segs: int):
    """Chunks of the live requests, oldest first, that fill one step of ``cap`` rows and at
    most ``max_segs`` segments: [(live, first row, rows)]; advances each request's cursor."""
    out, room = [], cap
    for lv in live:
        if room == 0 or len(out) == max_segs:
            break
        left = lv.res.rows - lv.next
        if left <= 0:
            continue
        n = min(left, room)
        out.append((lv, lv.next, n))
        lv.next += n
        room -= n
    return out


class _Live:
    """An admitted request: its arena range and progress."""

    def __init__(self, res: Result, off: int):
        self.res = res
        self.off = off
        self.next = 0  # rows handed to steps
        self.done = 0  # rows whose step has completed


# ----------------------------------------------------------------------------------------
# Server
# ----------------------------------------------------------------------------------------

class Server:
    """Serves render and query requests over ``scenes`` on the current HIP device.

    ``mlp_mode``: ``N.VXN_MLP_FP32`` or ``N.VXN_MLP_BF16_MFMA``.  ``ray_arena`` /
    ``point_arena`` rows bound the requests admitted at once (the serving memory pool,
    allocated up front); ``ray_capacity`` / ``point_capacity`` are the rows of one step.
    """

    def __init__(self, scenes: Sequence[ResidentScene], mlp_mode: int = N.VXN_MLP_FP32, ray_capacity: int = 1 << 16,
                 point_capacity: int = 1 << 20, ray_arena: int = 1 << 23, point_arena: int = 1 << 24,
                 max_segs: int = 256, n_samples: int = 64, n_importance: int = 64, up_sample_steps: int = 4,
                 use_graphs: bool = True, max_in_flight: int = 2):
        self.dev = scenes[0].grid.device
        self.scenes = list(scenes)
        self.mlp_mode = mlp_mode
        self.table = pack_scene_table(self.scenes, self.dev)
        self.C, self.Q = int(ray_capacity), int(point_capacity)
        self.max_segs = int(max_segs)
        self.max_in_flight = int(max_in_flight)
        f = lambda *sh: torch.zeros(*sh, device=self.dev, dtype=torch.float32)
        i32 = lambda *sh: torch.zeros(*sh, device=self.dev, dtype=torch.int32)
        # arenas (the admission pools)
        self.ray_pool, self.pt_pool = Arena(ray_arena), Arena(point_arena)
        self.a_o, self.a_d, self.a_rgb, self.a_depth = f(ray_arena, 3), f(ray_arena, 3), f(ray_arena, 3), f(ray_arena, 1)
        self.p_xyz, self.p_sdf, self.p_nrm = f(point_arena, 3), f(point_arena, 1), f(point_arena, 3)
        # step buffers
        self.b_o, self.b_d, self.b_rgb, self.b_depth = f(self.C, 3), f(self.C, 3), f(self.C, 3), f(self.C, 1)
        self.b_scene, self.b_n = i32(self.C), i32(1)
        self.q_xyz, self.q_sdf, self.q_nrm = f(self.Q, 3), f(self.Q, 1), f(self.Q, 3)
        self.q_scene, self.q_n = i32(self.Q), i32(1)
        self.r_segs, self.r_nseg = i32(self.max_segs, 3), i32(1)
        self.s_segs, self.s_nseg = i32(self.max_segs, 3), i32(1)
        self._pinned = [torch.zeros(2 * self.max_segs * 3 + 2, dtype=torch.int32, pin_memory=True) for _ in range(2)]
        self._pin_k = 0
        steps = up_sample_steps if n_importance > 0 else 0
        per = max(n_importance // steps, 1) if steps else 0
        from .functional import _linspace_dev

        # the sampler's tables, built exactly as the single-scene sampler builds them
        self.t_lin = _linspace_dev(0.0, 1.0, n_samples, self.dev)
        self.u_det = _linspace_dev(0.5 / per, 1.0 - 0.5 / per, per, self.dev) if per else self.t_lin
        a = N.ServeRenderArgs()
        a.scenes, a.n_scenes = self.table.data_ptr(), len(self.scenes)
        a.ray_scene, a.n_rays, a.capacity = self.b_scene.data_ptr(), self.b_n.data_ptr(), self.C
        a.rays_o, a.rays_d = self.b_o.data_ptr(), self.b_d.data_ptr()
        a.n_samples, a.n_importance, a.up_sample_steps = n_samples, n_importance, up_sample_steps
        a.t_lin, a.u_det, a.bound_radius, a.mlp_mode = self.t_lin.data_ptr(), self.u_det.data_ptr(), 1.0, mlp_mode
        a.rgb, a.depth = self.b_rgb.data_ptr(), self.b_depth.data_ptr()
        nb = ctypes.c_int64(0)
        N.check(N.lib().vxn_serve_render_workspace_bytes(ctypes.byref(a), ctypes.byref(nb)), "serve_workspace")
        self.ws = torch.empty(int(nb.value), device=self.dev, dtype=torch.uint8)
        a.workspace, a.workspace_bytes = self.ws.data_ptr(), self.ws.numel()
        self.args = a
        self.graphs = {}
        if use_graphs:
            self._capture()

    # -- the two step kinds (each a fixed launch sequence over device-side sizes) --------
    def _render_step(self, stream: int):
        L = N.lib()
        N.check(L.vxn_serve_move(self.r_segs.data_ptr(), self.r_nseg.data_ptr(), self.max_segs, self.C,
                                 self.a_o.shape[0], 1, self.a_o.data_ptr(), self.b_o.data_ptr(), 3,
                                 self.a_d.data_ptr(), self.b_d.data_ptr(), 3, self.b_scene.data_ptr(),
                                 self.b_n.data_ptr(), stream), "serve_move")
        N.check(L.vxn_serve_render(ctypes.byref(self.args), stream), "serve_render")
        N.check(L.vxn_serve_move(self.r_segs.data_ptr(), self.r_nseg.data_ptr(), self.max_segs, self.C,
                                 self.a_o.shape[0], 0, self.a_rgb.data_ptr(), self.b_rgb.data_ptr(), 3,
                                 self.a_depth.data_ptr(), self.b_depth.data_ptr(), 1, None, None, stream),
                "serve_move")

    def _query_step(self, stream: int):
        L = N.lib()
        N.check(L.vxn_serve_move(self.s_segs.data_ptr(), self.s_nseg.data_ptr(), self.max_segs, self.Q,
                                 self.p_xyz.shape[0], 1, self.p_xyz.data_ptr(), self.q_xyz.data_ptr(), 3, None, None,
                                 0, self.q_scene.data_ptr(), self.q_n.data_ptr(), stream), "serve_move")
        N.check(L.vxn_serve_query(self.table.data_ptr(), len(self.scenes), self.q_scene.data_ptr(),
                                  self.q_n.data_ptr(), self.Q, self.q_xyz.data_ptr(), self.q_sdf.data_ptr(),
                                  self.q_nrm.data_ptr(), stream), "serve_query")
        N.check(L.vxn_serve_move(self.s_segs.data_ptr(), self.s_nseg.data_ptr(), self.max_segs, self.Q,
                                 self.p_xyz.shape[0], 0, self.p_sdf.data_ptr(), self.q_sdf.data_ptr(), 1,
                                 self.p_nrm.data_ptr(), self.q_nrm.data_ptr(), 3, None, None, stream), "serve_move")

    def _capture(self):
        # empty steps (no segments) to warm up, then capture each step kind once
        self.r_nseg.zero_()
        self.s_nseg.zero_()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self._render_step(side.cuda_stream)
            self._query_step(side.cuda_stream)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        for name, fn in (("render", self._render_step), ("query", self._query_step)):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                fn(torch.cuda.current_stream().cuda_stream)
            self.graphs[name] = g
        torch.cuda.synchronize()

    def _run_step(self, name: str):
        if name in self.graphs:
            self.graphs[name].replay()
        elif name == "render":
            self._render_step(torch.cuda.current_stream().cuda_stream)
        else:
            self._query_step(torch.cuda.current_stream().cuda_stream)

    # -- request handling ------------------------------------------------------------
    def _admit(self, idx: int, req) -> Optional[_Live]:
        kind = "render" if isinstance(req, RenderRequest) else "query"
        if not 0 <= req.scene < len(self.scenes):
            raise ValueError(f"request {idx}: no scene {req.scene}")
        pool = self.ray_pool if kind == "render" else self.pt_pool
        if req.rows > pool.rows:
            raise ValueError(f"request {idx}: {req.rows} rows exceed the {kind} arena ({pool.rows})")
        off = pool.alloc(req.rows)
        if off is None:
            return None
        if kind == "render":
            o, d = pinhole_rays(req, self.dev)
            self.a_o[off:off + req.rows].copy_(o)
            self.a_d[off:off + req.rows].copy_(d)
        else:
            self.p_xyz[off:off + req.rows].copy_(req.points.to(self.dev, torch.float32).reshape(-1, 3))
        return _Live(Result(idx, kind, req.scene, req.rows, req.arrival), off)

    def serve(self, requests: Sequence, realtime: bool = True) -> List[Result]:
        """Runs ``requests`` (each with an ``arrival`` time in seconds from the start) to
        completion and returns one Result per request, in order.  ``realtime=False``
        admits every request at once (arrival times then only order the queue)."""
        order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
        pending = deque(order)
        results: List[Optional[Result]] = [None] * len(requests)
        live_r, live_q = deque(), deque()
        in_flight = deque()  # (event, [(live, rows)])
        stream = torch.cuda.current_stream()
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        t0 = time.perf_counter()

        def retire(block: bool):
            while in_flight and (block or in_flight[0][0].query()):
                ev, work = in_flight.popleft()
                ev.synchronize()
                t_done = start.elapsed_time(ev) * 1e-3
                block = False
                for lv, n in work:
                    lv.done += n
                    if lv.done == lv.res.rows:
                        r = lv.res
                        a, b = lv.off, lv.off + r.rows
                        if r.kind == "render":
                            r.outputs = {"rgb": self.a_rgb[a:b].clone(), "depth": self.a_depth[a:b].clone()}
                            self.ray_pool.release(lv.off, r.rows)
                            live_r.remove(lv)
                        else:
                            r.outputs = {"sdf": self.p_sdf[a:b].clone(), "normal": self.p_nrm[a:b].clone()}
                            self.pt_pool.release(lv.off, r.rows)
                            live_q.remove(lv)
                        r.done = t_done
                        results[r.index] = r

        while pending or live_r or live_q or in_flight:
            now = time.perf_counter() - t0
            blocked = False  # the oldest arrived request did not fit in its arena
            while pending and (not realtime or requests[pending[0]].arrival <= now):
                i = pending[0]
                lv = self._admit(i, requests[i])
                if lv is None:
                    blocked = True
                    break  # first come, first served: wait for memory
                lv.res.admitted = now
                (live_r if lv.res.kind == "render" else live_q).append(lv)
                pending.popleft()
            rseg = plan_step(live_r, self.C, self.max_segs)
            qseg = plan_step(live_q, self.Q, self.max_segs)
            if not rseg and not qseg:
                if in_flight:
                    retire(True)
                elif pending:
                    # with nothing live or in flight the arenas are empty: a request that did
                    # not fit then never will (one that arrived after the admission pass is
                    # admitted on the next pass)
                    if blocked and not live_r and not live_q:
                        raise RuntimeError("a pending request cannot be admitted: the arena is too small")
                    wait = requests[pending[0]].arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                continue
            if len(in_flight) >= self.max_in_flight:
                retire(True)
            # stage this step's segment tables (pinned buffer k was last read two steps ago)
            buf = self._pinned[self._pin_k]
            self._pin_k ^= 1
            hb = buf.numpy()
            hb[0], hb[1] = len(rseg), len(qseg)
            tab = hb[2:].reshape(2, self.max_segs, 3)
            for k, (lv, first, n) in enumerate(rseg):
                tab[0, k] = (lv.off + first, n, lv.res.scene)
            for k, (lv, first, n) in enumerate(qseg):
                tab[1, k] = (lv.off + first, n, lv.res.scene)
            ns = 3 * self.max_segs
            if rseg:
                self.r_nseg.copy_(buf[0:1], non_blocking=True)
                self.r_segs.view(-1)[:3 * len(rseg)].copy_(buf[2:2 + 3 * len(rseg)], non_blocking=True)
                self._run_step("render")
            if qseg:
                self.s_nseg.copy_(buf[1:2], non_blocking=True)
                self.s_segs.view(-1)[:3 * len(qseg)].copy_(buf[2 + ns:2 + ns + 3 * len(qseg)], non_blocking=True)
                self._run_step("query")
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(stream)
            in_flight.append((ev, [(lv, n) for lv, _, n in rseg] + [(lv, n) for lv, _, n in qseg]))
            retire(False)
        torch.cuda.synchronize()
        return results  # type: ignore[return-value]


def latency_stats(results: Sequence[Result], kind: Optional[str] = None) -> Dict[str, float]:
    """p50 / p99 / mean latency (ms) and rows/s over the results of one kind."""
    rs = [r for r in results if kind is None or r.kind == kind]
    if not rs:
        return {}
    lat = np.array([r.latency for r in rs]) * 1e3
    span = max(r.done for r in rs) - min(r.arrival for r in rs)
    rows = sum(r.rows for r in rs)
    return {"count": len(rs), "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "rows": rows, "rows_per_s": rows / span if span > 0 else float("nan")}
