"""The serving loop on the GPU: every request's result equals the standalone
single-scene call on the same rays / points, and is the same alone and in a mix
(bit-exact: per-ray work never depends on the rest of a batch); served renders and
queries against the CPU oracle's render.py / mesh.py arithmetic (render.py:35-54:
near/far from the unit sphere, deterministic sampler; mesh.py:80-114: query_sdf and
query_normals)."""

import ctypes

import numpy as np
import pytest
import torch

from voxneus_amd import _native as N
from voxneus_amd import functional as Fn
from voxneus_amd import scenes as S
from voxneus_amd.serving import QueryRequest, RenderRequest, Server, make_resident_scene, pinhole_rays

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


@pytest.fixture(scope="module")
def scenes():
    # the third scene's model uses the Eq. 4 analytical gradient
    spec = [(64, 15, 1, torch.float32, "interpolated"), (96, 16, 2, torch.bfloat16, "interpolated"),
            (48, 15, 3, torch.float32, "analytical")]
    return [make_resident_scene(r, t, seed, DEV, table_dtype=dt, gradient_mode=gm) for r, t, seed, dt, gm in spec]


def _requests(n, n_scenes, seed, sizes=((24, 16), (40, 30), (7, 5)), arrival_gap=0.0):
    rng = np.random.default_rng(seed)
    cams = S.hemisphere_cameras(49)
    reqs = []
    for i in range(n):
        sc = int(rng.integers(n_scenes))
        if rng.random() < 0.7:
            W, H = sizes[int(rng.integers(len(sizes)))]
            c2w = cams[int(rng.integers(49))]
            reqs.append(RenderRequest(sc, c2w, W, H, 1.5 * W, 1.5 * W, W / 2, H / 2, arrival=i * arrival_gap))
        else:
            g = torch.Generator(device=DEV)
            g.manual_seed(seed * 1000 + i)
            pts = torch.rand((int(rng.integers(100, 3000)), 3), generator=g, device=DEV) * 2.2 - 1.1
            reqs.append(QueryRequest(sc, pts, arrival=i * arrival_gap))
    return reqs


def _standalone(sc, req, mlp_mode):
    if isinstance(req, QueryRequest):
        out = Fn.grid_query_raw(sc.grid, sc.box, req.points, want_sdf=True, want_grad=False, want_normal=True,
                                gradient_mode=sc.gradient_mode)
        return {"sdf": out["sdf"].reshape(-1, 1), "normal": out["normal"]}
    o, d = pinhole_rays(req, DEV)
    z, dists, mid = Fn.upsample(o, d, None, None, sc.grid, sc.box, 64, 64, 4, False, sc.occ_mask, sc.occ_box,
                                bound_radius=1.0)
    cfg = Fn.RenderConfig(sc.box, sc.level_res, mlp_mode=mlp_mode, gradient_mode=sc.gradient_mode)
    out = Fn.RenderFunction.apply(sc.grid, sc.table, sc.mlp, torch.tensor(sc.s, device=DEV), o, d, mid, dists, cfg,
                                  False, None)
    return {"rgb": out[0], "depth": out[1]}


@pytest.mark.parametrize("mode", [N.VXN_MLP_FP32, N.VXN_MLP_BF16_MFMA])
def test_serving_matches_standalone_calls(scenes, mode):
    reqs = _requests(14, len(scenes), seed=5)
    srv = Server(scenes, mlp_mode=mode, ray_capacity=1500, point_capacity=2048, ray_arena=20000, point_arena=16000)
    res = srv.serve(reqs, realtime=False)
    for r, q in zip(res, reqs):
        ref = _standalone(scenes[q.scene], q, mode)
        for k, v in ref.items():
            assert torch.equal(r.outputs[k], v.reshape(r.outputs[k].shape)), (r.index, r.kind, k)


def test_serving_alone_equals_mixed_and_graph_equals_eager(scenes):
    reqs = _requests(20, len(scenes), seed=7)
    mixed = Server(scenes, ray_capacity=777, point_capacity=999, ray_arena=9000, point_arena=9000).serve(
        reqs, realtime=False)
    eager = Server(scenes, ray_capacity=777, point_capacity=999, ray_arena=9000, point_arena=9000,
                   use_graphs=False).serve(reqs, realtime=False)
    for i in (0, 3, 11):
        alone = Server(scenes, ray_capacity=1 << 12, point_capacity=1 << 12, ray_arena=4096,
                       point_arena=4096).serve([reqs[i]], realtime=False)[0]
        for k in alone.outputs:
            assert torch.equal(alone.outputs[k], mixed[i].outputs[k]), (i, k)
    for a, b in zip(mixed, eager):
        for k in a.outputs:
            assert torch.equal(a.outputs[k], b.outputs[k]), (a.index, k)


def test_serving_realtime_admission_and_latency(scenes):
    reqs = _requests(24, len(scenes), seed=9, arrival_gap=0.002)
    srv = Server(scenes, ray_capacity=1024, point_capacity=1024, ray_arena=3000, point_arena=3000)
    res = srv.serve(reqs, realtime=True)
    for r in res:
        assert r.done >= r.arrival and r.admitted >= r.arrival - 1e-3
        assert all(torch.isfinite(v).all() for v in r.outputs.values())
    assert srv.ray_pool.free_rows == 3000 and srv.pt_pool.free_rows == 3000


@pytest.mark.parametrize("mode", [N.VXN_MLP_FP32, N.VXN_MLP_BF16_MFMA])
def test_serving_matches_oracle(scenes, mode):
    """Served renders on 48 sampled pixels per request against the oracle's render of the
    same rays (FP32 form within 1e-5; BF16 form within 5e-3 RGB of the FP64 oracle) and
    served queries on 300 points per request against query_sdf (bit-exact) and
    query_normals (1e-6); the scenes cover FP32 and bf16 tables and the Eq. 4 gradient."""
    from tests.test_gpu_configs import check_query, check_render

    reqs = _requests(12, len(scenes), seed=13, sizes=((40, 30), (64, 48)))
    srv = Server(scenes, mlp_mode=mode, ray_capacity=2048, point_capacity=2048, ray_arena=40000,
                 point_arena=40000)
    res = srv.serve(reqs, realtime=False)
    kinds = set()
    for i, (r, q) in enumerate(zip(res, reqs)):
        sc = scenes[q.scene]
        kinds.add((r.kind, q.scene))
        if isinstance(q, RenderRequest):
            check_render(sc, q, r, mode, 48, seed=70 + i)
        else:
            check_query(sc, q.points, r, 300, seed=80 + i)
    assert {k for k, _ in kinds} == {"render", "query"}


def test_resident_scene_rejects_tables_the_kernels_cannot_index(scenes):
    """The hash index is masked with T - 1, so T must be a power of two; the fused MLP
    takes 16 levels of 2 features; a misaligned BF16 table is refused by the C-ABI."""
    sc = scenes[0]
    with pytest.raises(ValueError, match="power of two"):
        type(sc)(sc.grid, torch.zeros(16, 3000, 2, device=DEV), sc.mlp, sc.level_res, sc.s)
    with pytest.raises(ValueError, match="one value per level"):
        type(sc)(sc.grid, sc.table, sc.mlp, sc.level_res[:8], sc.s)
    with pytest.raises(ValueError, match="float32 or bfloat16"):
        type(sc)(sc.grid, sc.table.half(), sc.mlp, sc.level_res, sc.s)
    raw = torch.zeros(16 * 1024 * 2 + 8, dtype=torch.bfloat16, device=DEV)
    odd = raw[2:2 + 16 * 1024 * 2].view(16, 1024, 2)  # 4 bytes past a 16-byte boundary
    bad = type(sc)(sc.grid, odd, sc.mlp, sc.level_res, sc.s)
    assert bad.table.data_ptr() % 16 == 4
    n = ctypes.c_int64(0)
    assert N.lib().vxn_serve_scene_bytes(ctypes.byref(n)) == 0
    buf = (ctypes.c_uint8 * n.value)()
    d = bad.desc()
    assert N.lib().vxn_serve_pack_scene(ctypes.byref(d), buf, n.value) == -3
    d = sc.desc()
    assert N.lib().vxn_serve_pack_scene(ctypes.byref(d), buf, n.value) == 0
