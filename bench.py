"""Benchmark: fused VoxNeuS fwd+bwd training step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form both|bf16|fp32] [--no-cpu-baseline]

Workload (the precision/throughput configuration): FP32 SDF grid 256^3 of a seeded
multi-primitive scene, 16-level x 2-feature hash with 2^19 entries per level (per-level
scale 1.5), 51->64->3 radiance MLP, 8192 rays per step drawn from 49 hemisphere pinhole
cameras (400x300, f=600), 64 stratified + 4x16 upsampled samples per ray.  One step =
ray bounds, 4-step upsampler, fused render forward + Huber loss + backward, V_R build,
closed-form eikonal (w=0.1) + curvature (w=5e-4) gradients, k=5 s-amplification
(training/trainer.py:188-326).  Inputs are resident in HBM before timing starts.

Forms: BF16 (bf16 hash table, bf16 MLP on matrix cores with FP32 accumulation; the
headline ``value``) and FP32 (every operand FP32), each with its max-abs RGB error against
the FP64 oracle on the same z values.

Prints one JSON line.  ``roofline`` covers the dominant kernel of the render launch (by
time), timed live with HIP events recorded between the kernels on the stream they run on;
``traffic`` is read from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary.
``grid_gather`` is the SDF + normal query over a 320^3 lattice.  ``cpu_baseline`` times the
CPU oracle (an op-for-op restatement of the reference's PyTorch path) on a bounded sample
of the same workload.
"""

from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
RAYS = 8192
SAMPLES = 128
LATTICE = 320
# Algorithmic bytes per sample (no cache credit), per kernel of the fused render launch:
#   hash_features  (BF16 form) the geometry pass, grid stencil 32 words = 128 B in, SDF +
#                  gradient 16 B and a flag byte out, then the sample lists and the gather,
#                  16 levels x 8 corners x 2 features = 256 bf16 table entries = 512 B
#                  gathered, 16 feature pairs written (64 B), counted for every sample as the
#                  reference gathers them (the kernel skips the samples whose colour reaches
#                  no output: q < 0 or outside the sphere);
#   render_kernel  grid forward 32 words (8 corners + 24 face neighbours) = 128 B (BF16: the
#                  geometry pass's 16 B read back), the hash features (BF16: 64 B read back;
#                  FP32: 1024 B gathered in the kernel), per-ray streaming ~9 B;
#   hash_gradient  the hash-table gradient stage (binned: count, scan, record write, per-bin
#                  sums between the same two events; in the BF16 form the count and scans run
#                  before the render kernel, beside the gather, so the span holds the record
#                  writes and sums), accounted as the 256 FP32 read-modify-writes it replaces
#                  = 2048 B;
#   grid_gradient  the grid gradient (binned: count, scan, brick write, per-tile sums; or the
#                  atomic scatter), accounted as the 32 words read-modify-write it replaces
#                  = 256 B.
KERNEL_BYTES = {
    # FP32 form: the render kernel gathers the hash features itself (no hash_features launch)
    "fp32": {"render_kernel": 128 + 1024 + 9, "hash_gradient": 2048, "grid_gradient": 256},
    "bf16": {"hash_features": 128 + 17 + 512 + 64, "render_kernel": 16 + 64 + 9, "hash_gradient": 2048,
             "grid_gradient": 256},
}
# the sampler: 112 SDF queries x 8 corners x 4 B per 128-sample ray (no cache credit)
SIDE_BYTES_PER_SAMPLE = {"upsample": 28}
# names of the kernels in the committed rocprofv3 traffic summaries
PROFILE_NAMES = {"upsample": "upsample_kernel", "reg_vr": "reg_vr_kernel"}
# event index pairs bracketing each kernel in vxn_render_args.kernel_events
KERNEL_EVENTS = {"hash_features": (0, 1), "render_kernel": (1, 2), "grid_gradient": (2, 3), "hash_gradient": (3, 4),
                 "reductions": (4, 5)}
# with the grid gradient on its own stream (StepConfig.concurrent_scatter, the default) the
# hash-table gradient starts with it, and the reductions' span includes the wait for it
KERNEL_EVENTS_CONCURRENT = dict(KERNEL_EVENTS, hash_gradient=(2, 4))
# lattice query: 32-word stencil (128 B) in, SDF + normal (16 B) out per lattice point
LATTICE_BYTES_PER_POINT = 128 + 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(seed=3, n_batches=4):
    from voxneus_amd import scenes

    t0 = time.time()
    sc = scenes.make_scene("multi", 256, 19, seed=seed)
    batches = []
    for b in range(n_batches):
        o, d = scenes.hemisphere_ray_batch(RAYS, seed=100 + b)
        batches.append((o, d, scenes.procedural_rgb(sc.shape, o, d)))
    log(f"workload built in {time.time() - t0:.1f}s")
    return sc, batches


def mlp_flat(sc):
    import torch

    return torch.cat([torch.from_numpy(sc.mlp[k]).reshape(-1) for k in ("W0", "b0", "W1", "b1")])


def run_gpu(sc, batches, form, steps, warmup, serial_grid=False):
    import torch

    from voxneus_amd import _native as N
    from voxneus_amd.hash_grid import level_resolutions
    from voxneus_amd.training import FusedTrainStep, StepConfig

    dev = torch.device("cuda:0")
    table = torch.from_numpy(sc.embeddings).to(dev)
    mlp_mode = N.VXN_MLP_FP32
    if form == "bf16":
        table = table.to(torch.bfloat16)
        mlp_mode = N.VXN_MLP_BF16_MFMA
    st = FusedTrainStep(torch.from_numpy(sc.grid).to(dev), table, mlp_flat(sc).to(dev),
                        torch.tensor(0.3, device=dev), level_resolutions(16, 16, sc.hash_max_resolution),
                        cfg=StepConfig(perturb=True, mlp_mode=mlp_mode, concurrent_scatter=not serial_grid))
    gb = [tuple(torch.from_numpy(a).to(dev) for a in b) for b in batches]
    ev = []

    def one(i, timed):
        o, d, t = gb[i % len(gb)]
        if timed:
            st.kernel_events = N.KernelEvents(6)
            ev.append(st.kernel_events)
        else:
            st.kernel_events = None
        return st.step(o, d, t)

    for i in range(warmup):
        one(i, False)
    torch.cuda.synchronize()
    # a full collection of the Python heap takes tens of ms: run it now, not inside the
    # timed steps, whose host side is only the launch sequence
    gc.collect()
    gc.disable()
    t0 = time.perf_counter()
    for i in range(steps):
        out = one(i, True)
    t_host = time.perf_counter() - t0  # host time to queue the steps
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    gc.enable()
    kev = KERNEL_EVENTS_CONCURRENT if st.scatter is not None else KERNEL_EVENTS
    kern = {k: float(np.mean([e.elapsed_ms(i, j) for e in ev])) for k, (i, j) in kev.items()}
    loss = out["loss"].item()
    # the step's launches outside the render call, each timed alone on the last step's inputs:
    # the sampler (with its V_R marks as the step makes them: the cell clear, the per-cell
    # marks and the bitmask written from them) and the regularisers over the V_R bitmask
    o, d, _ = gb[(steps - 1) % len(gb)]

    def sample():
        if st.vr_cells is not None:
            st.vr_cells.zero_()
        st.sample(o, d, None, None, st.bitmask, st.vr_cells)

    kern["upsample"] = _event_ms(sample)
    kern["reg_vr"] = _event_ms(st._regularisers)
    n_vr = int(st.vr_counts[0].item())
    return {"ms_per_step": dt / steps * 1e3, "rays_per_s": RAYS * steps / dt, "kernel_ms": kern, "loss": loss,
            "host_ms_per_step": t_host / steps * 1e3, "n_vr": n_vr, "n_vox": st.grid.numel()}


def run_full_render(form):
    """Config 4: every pixel of the 49 hemisphere views at 800x600 (f = 1200), 64 + 4x16
    samples per ray, on one resident scene (320^3 FP32 grid, 16 x 2^19 hash), through the
    serving kernels with all views admitted at once."""
    import torch

    from voxneus_amd import _native as N
    from voxneus_amd import scenes as S
    from voxneus_amd.serving import RenderRequest, Server, make_resident_scene

    dev = torch.device("cuda:0")
    sc = make_resident_scene(LATTICE, 19, 4, dev, table_dtype=torch.bfloat16 if form == "bf16" else torch.float32)
    mode = N.VXN_MLP_BF16_MFMA if form == "bf16" else N.VXN_MLP_FP32
    srv = Server([sc], mlp_mode=mode, ray_capacity=1 << 17, point_capacity=1 << 10, ray_arena=49 * 800 * 600,
                 point_arena=1 << 10, max_segs=64)
    cams = S.hemisphere_cameras(49)
    reqs = [RenderRequest(0, cams[v], 800, 600, 1200.0, 1200.0, 400.0, 300.0) for v in range(49)]
    srv.serve(reqs[:1], realtime=False)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = srv.serve(reqs, realtime=False)
    dt = time.perf_counter() - t0
    rays = sum(r.rows for r in res)
    out = {"views": 49, "image": "800x600", "rays": rays, "samples_per_ray": SAMPLES, "seconds": dt,
           "rays_per_s": rays / dt, "samples_per_s": rays * SAMPLES / dt, "grid": f"{LATTICE}^3 FP32",
           "hash": "16 x 2^19 " + ("bf16" if form == "bf16" else "fp32")}
    del srv, sc
    torch.cuda.empty_cache()
    return out


def run_lattice(reps=10):
    """SDF + normal at every point of a 320^3 lattice over the box of a 320^3 grid."""
    import torch

    from voxneus_amd import _native as N
    from voxneus_amd import functional as Fn
    from voxneus_amd import scenes

    dev = torch.device("cuda:0")
    shape = scenes.MultiPrimitive.seeded(3)
    grid = torch.from_numpy(scenes.sdf_grid_from(shape.sdf, LATTICE)).to(dev)
    box = ((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    Fn.lattice_query(grid, box, LATTICE, LATTICE, LATTICE, box)  # warm-up
    evs = []
    for _ in range(reps):
        e = N.KernelEvents(2)
        s = torch.cuda.current_stream().cuda_stream
        e.record(0, s)
        Fn.lattice_query(grid, box, LATTICE, LATTICE, LATTICE, box)
        e.record(1, s)
        evs.append(e)
    ms = float(np.mean([e.elapsed_ms(0, 1) for e in evs]))
    pts = LATTICE ** 3
    achieved = LATTICE_BYTES_PER_POINT * pts / (ms * 1e-3) / 1e9
    return {"kernel": "lattice_query_kernel", "grid": f"{LATTICE}^3 FP32", "points": pts, "kernel_ms": ms,
            "bytes_per_point": LATTICE_BYTES_PER_POINT, "achieved": achieved, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "compulsory_GBps": (grid.numel() * 4 + pts * 16) / (ms * 1e-3) / 1e9}


def _event_ms(fn, reps=5):
    """Mean HIP-event time (ms) of ``fn`` on the current stream over ``reps`` calls after
    one warm-up call."""
    import torch

    from voxneus_amd import _native as N
    fn()
    evs = []
    for _ in range(reps):
        e = N.KernelEvents(2)
        st = torch.cuda.current_stream().cuda_stream
        e.record(0, st)
        fn()
        e.record(1, st)
        evs.append(e)
    torch.cuda.synchronize()
    return float(np.mean([e.elapsed_ms(0, 1) for e in evs]))


def run_pipeline_stages(sc, batches, form):
    """The steps around the training step and after training, at the bench sizes: both
    optimizers after a fused step (256^3 grid, 16 x 2^19 hash), the occupancy update (128^3:
    a full sweep and the trainer's quarter of random cells), progressive super-resolution
    160^3 -> 320^3, and marching cubes on a 320^3 grid."""
    import torch

    from voxneus_amd import _native as N
    from voxneus_amd import functional as Fn
    from voxneus_amd import scenes
    from voxneus_amd.hash_grid import level_resolutions
    from voxneus_amd.mesh import marching_cubes
    from voxneus_amd.training import FusedOptimizer, FusedTrainStep, StepConfig

    dev = torch.device("cuda:0")
    table = torch.from_numpy(sc.embeddings).to(dev)
    mlp = mlp_flat(sc).to(dev)
    if form == "bf16":
        table = table.to(torch.bfloat16)
    st = FusedTrainStep(torch.from_numpy(sc.grid).to(dev), table, mlp, torch.tensor(0.3, device=dev),
                        level_resolutions(16, 16, sc.hash_max_resolution),
                        cfg=StepConfig(perturb=True, mlp_mode=N.VXN_MLP_BF16_MFMA if form == "bf16" else N.VXN_MLP_FP32))
    o, d, tgt = (torch.from_numpy(a).to(dev) for a in batches[0])
    st.step(o, d, tgt)
    opt = FusedOptimizer(st, total_steps=40000)
    it = [0]

    def opt_step():
        opt.step(it[0])
        it[0] += 1

    out = {}
    ms = _event_ms(opt_step)
    n_grid, n_tab = st.grid.numel(), st.table.numel()
    upd = int(opt.n_updated.item())
    # masked grid Adam: gradient read for every voxel, 28 B per updated voxel; the network
    # Adam: 28 B per element (+2 B bf16 copy)
    nbytes = 4 * n_grid + 28 * upd + (28 + (2 if form == "bf16" else 0)) * n_tab + 28 * (mlp.numel() + 1)
    out["optimizers"] = {"ms": ms, "grid_voxels": n_grid, "grid_updated": upd, "table_entries": n_tab,
                         "algorithmic_GBps": nbytes / (ms * 1e-3) / 1e9}
    out["occupancy_sweep_128"] = {"ms": _event_ms(lambda: st.update_occupancy(full_sweep=True))}
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    out["occupancy_quarter_128"] = {"ms": _event_ms(lambda: st.update_occupancy(0.25, generator=gen))}
    g160 = torch.from_numpy(scenes.sdf_grid_from(scenes.MultiPrimitive.seeded(3).sdf, 160)).to(dev)
    ms = _event_ms(lambda: Fn.upsample_grid(g160, (LATTICE,) * 3))
    out["upsample_160_to_320"] = {"ms": ms, "algorithmic_GBps": (g160.numel() * 4 + LATTICE ** 3 * 4) / (ms * 1e-3) / 1e9}
    g320 = Fn.upsample_grid(g160, (LATTICE,) * 3)
    box6 = (-1.0, -1.0, -1.0, 1.0, 1.0, 1.0)
    v, f = marching_cubes(g320, box6)
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        v, f = marching_cubes(g320, box6)
    torch.cuda.synchronize()
    out["marching_cubes_320"] = {"ms_wall_incl_count_read": (time.perf_counter() - t0) / reps * 1e3,
                                 "vertices": int(v.shape[0]), "faces": int(f.shape[0])}
    del st, opt, g160, g320, v, f
    torch.cuda.empty_cache()
    return out


SERVE_RENDER_SIZES = ((64, 64), (200, 150), (400, 300), (800, 600))


def serving_requests(n, n_scenes, lam, seed=11, render_frac=2.0 / 3.0):
    """Poisson arrivals (rate ``lam`` per second) of renders of the four image sizes from
    random hemisphere poses (f = 1.5 W) and SDF/normal queries of 10^4..10^6 uniform points
    (log-uniform), each on a random resident scene."""
    import torch

    from voxneus_amd import scenes as S
    from voxneus_amd.serving import QueryRequest, RenderRequest

    rng = np.random.default_rng(seed)
    cams = S.hemisphere_cameras(49)
    t = 0.0
    reqs = []
    for i in range(n):
        t += rng.exponential(1.0 / lam)
        sc = int(rng.integers(n_scenes))
        if rng.random() < render_frac:
            W, H = SERVE_RENDER_SIZES[int(rng.integers(len(SERVE_RENDER_SIZES)))]
            reqs.append(RenderRequest(sc, cams[int(rng.integers(49))], W, H, 1.5 * W, 1.5 * W, W / 2, H / 2, t))
        else:
            npts = int(round(10 ** rng.uniform(4.0, 6.0)))
            g = torch.Generator(device="cuda:0")
            g.manual_seed(seed * 100003 + i)
            pts = torch.rand((npts, 3), generator=g, device="cuda:0") * 2.0 - 1.0
            reqs.append(QueryRequest(sc, pts, t))
    return reqs


def run_serving(form, n_requests=160, utilisation=0.7):
    """Config-5 serving loop: 24 resident scenes, Poisson arrivals at a fixed offered load
    (``utilisation`` of the measured service rate), p50/p99 latency per request kind."""
    import torch

    from voxneus_amd import _native as N
    from voxneus_amd.serving import Server, latency_stats, make_resident_scene

    dev = torch.device("cuda:0")
    grids, logs = (96, 128, 192, 256), (17, 18, 19)
    tdt = torch.bfloat16 if form == "bf16" else torch.float32
    scenes = [make_resident_scene(grids[i % 4], logs[i % 3], 1000 + i, dev, table_dtype=tdt) for i in range(24)]
    mode = N.VXN_MLP_BF16_MFMA if form == "bf16" else N.VXN_MLP_FP32
    srv = Server(scenes, mlp_mode=mode, ray_capacity=1 << 16, point_capacity=1 << 20, ray_arena=1 << 23,
                 point_arena=1 << 24)
    # service rate: the same request mix, all admitted at once
    cal = serving_requests(24, len(scenes), 1.0, seed=7)
    srv.serve(cal[:4], realtime=False)  # warm-up
    t0 = time.perf_counter()
    srv.serve(cal, realtime=False)
    busy = time.perf_counter() - t0
    lam = utilisation * len(cal) / busy
    reqs = serving_requests(n_requests, len(scenes), lam)
    t0 = time.perf_counter()
    res = srv.serve(reqs, realtime=True)
    wall = time.perf_counter() - t0
    out = {"scenes": len(scenes), "resident_MB": sum(sc.nbytes for sc in scenes) / 1e6, "requests": n_requests,
           "offered_req_per_s": lam, "utilisation_target": utilisation, "wall_s": wall,
           "render": latency_stats(res, "render"), "query": latency_stats(res, "query")}
    del srv, scenes
    torch.cuda.empty_cache()
    return out


def profiled_traffic(kernel, form):
    """HBM bytes per dispatch of ``kernel`` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (already scaled by the copy calibration), or None."""
    import re

    path = os.path.join(REPO, "profiles", f"fused_step_grid256_rays8192_{form}.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_dispatch %s: (\d+)" % re.escape(kernel), text, re.M)
    return int(m.group(1)) if m else None


def rgb_error_vs_fp64(sc, batch, form="fp32", n_rays=256):
    """Max-abs RGB of the HIP forward (FP32 or BF16 form) vs the FP64 oracle on the same z
    values and the same (FP32) parameters."""
    import torch

    from oracle import ref_cpu as R
    from voxneus_amd import functional as Fn

    dev = torch.device("cuda:0")
    o, d = (torch.from_numpy(a[:n_rays]) for a in batch[:2])
    res = R.hash_level_resolutions(16, 16, sc.hash_max_resolution)
    m32 = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    near, far = R.near_far_from_sphere(o, d)
    zs = m32.sample(o, d, near, far)
    m64 = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution,
                     dtype=torch.float64)
    with torch.no_grad():
        ref = m64.render(o.double(), d.double(), None, None, zs=tuple(z.double() for z in zs))["rgb"]
    from voxneus_amd import _native as N

    table = torch.from_numpy(sc.embeddings).to(dev)
    mode = N.VXN_MLP_FP32
    if form == "bf16":
        table, mode = table.to(torch.bfloat16), N.VXN_MLP_BF16_MFMA
    cfg = Fn.RenderConfig((-1, -1, -1, 1, 1, 1), res, mlp_mode=mode)
    s = torch.clamp(torch.exp(torch.tensor(0.3) * 10.0), 1e-4, 300.0)
    out = Fn.RenderFunction.apply(torch.from_numpy(sc.grid).to(dev), table, mlp_flat(sc).to(dev), s.to(dev),
                                  o.to(dev), d.to(dev), zs[2].to(dev), zs[1].to(dev), cfg, False, None)
    return float((out[0].detach().cpu().double() - ref).abs().max())


def cpu_baseline(sc, batch, n_rays):
    import torch

    from oracle import ref_cpu as R

    threads = len(os.sched_getaffinity(0))
    torch.set_num_threads(threads)
    m = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    o, d, t = (torch.from_numpy(a[:n_rays]) for a in batch)
    m.train_step(o, d, t, perturb=True)  # warm-up
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        m.train_step(o, d, t, perturb=True)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except Exception:
        pass
    return {"value": n_rays / med, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"{n_rays} of the 8192 rays of one step (same scene, 256^3 grid, 2^19 hash), median of 3 "
                      f"timed oracle fwd+bwd steps after 1 warm-up, {model}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--form", default="both", choices=["both", "fp32", "bf16"],
                    help="both (default): the BF16 form is the headline value, the FP32 form is reported beside it")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rays", type=int, default=128)
    ap.add_argument("--no-error-check", action="store_true", help="skip the FP64 RGB error check (profiling runs)")
    ap.add_argument("--no-lattice", action="store_true",
                    help="skip the 320^3 measurements (lattice query, 49-view 800x600 render)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop latency measurement")
    ap.add_argument("--serial-grid", action="store_true",
                    help="run the grid gradient on the main stream, after the render kernel (default: on its own "
                         "stream beside the hash-table gradient)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy (3 dispatches): the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (rays shard across GPUs by data "
              "parallelism with one gradient all-reduce per step, which is not implemented here).",
              file=sys.stderr)
        sys.exit(2)
    import torch

    if a.copy_calibration:
        src = torch.ones(64 << 20, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        log("copied 3 x 256 MiB")
        return
    sc, batches = make_workload()
    forms = ["bf16", "fp32"] if a.form == "both" else [a.form]
    ns = RAYS * SAMPLES
    res = {}
    for form in forms:
        log(f"GPU: {torch.cuda.get_device_name(0)}; form {form}; warmup {a.warmup}; steps {a.steps}")
        r = run_gpu(sc, batches, form, a.steps, a.warmup, a.serial_grid)
        log(f"gpu {form}: {r}")
        kernels = {}
        for k, ms in r["kernel_ms"].items():
            if k not in KERNEL_BYTES[form] and k not in ("reductions", "upsample", "reg_vr"):
                continue
            bps = KERNEL_BYTES[form].get(k, SIDE_BYTES_PER_SAMPLE.get(k))
            nbytes = None if bps is None else bps * ns
            if k == "reg_vr":  # the bitmask, then 7 grid words read and 7 gradient words RMW per V_R vertex
                nbytes = r["n_vox"] / 8 + r["n_vr"] * (7 * 4 + 7 * 8)
            ach = None if nbytes is None else nbytes / (ms * 1e-3) / 1e9
            traffic = profiled_traffic(PROFILE_NAMES.get(k, k), form)
            kernels[k] = {"ms": ms, "bytes_per_sample": bps, "algorithmic_bytes": nbytes, "achieved_GBps": ach,
                          "frac": None if ach is None else ach / HBM_PEAK_GBS, "traffic_bytes": traffic,
                          "traffic_frac": None if traffic is None else traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        err = None if a.no_error_check else rgb_error_vs_fp64(sc, batches[0], form)
        log(f"rgb max-abs err vs fp64 ({form}): {err}")
        res[form] = {"rays_per_s": r["rays_per_s"], "samples_per_s": r["rays_per_s"] * SAMPLES,
                     "ms_per_step": r["ms_per_step"], "host_ms_per_step": r["host_ms_per_step"],
                     "rgb_max_abs_err_vs_fp64": err, "kernels": kernels}
    main_form = forms[0]
    r = res[main_form]
    kms = {k: v["ms"] for k, v in r["kernels"].items()}
    dom = max(KERNEL_BYTES[main_form], key=lambda k: kms[k])  # of the render call's kernels
    nbytes = KERNEL_BYTES[main_form][dom] * ns
    achieved = r["kernels"][dom]["achieved_GBps"]
    lat = None if a.no_lattice else run_lattice()
    stages = None if a.no_lattice else run_pipeline_stages(sc, batches, main_form)
    if stages:
        log(f"pipeline stages: {stages}")
    if lat:
        log(f"lattice: {lat}")
    full = None
    if not a.no_lattice:
        full = {form: run_full_render(form) for form in forms}
        log(f"full-scale render: {full}")
    serving = None
    if not a.no_serving:
        serving = {}
        for form in forms:
            serving[form] = run_serving(form)
            log(f"serving {form}: {serving[form]}")
    cpu = None if a.no_cpu_baseline else cpu_baseline(sc, batches[0], a.cpu_rays)
    if cpu:
        log(f"cpu baseline: {cpu}")
    line = {
        "metric": "train rays/s fwd+bwd @256^3 (BF16 vs FP32 at stated RGB err); render p50/p99 ms",
        "value": r["rays_per_s"], "unit": "rays/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if main_form == "fp32" else "bf16", "data": "synthetic",
        "config": {"workload": "fused fwd+bwd train step: 256^3 FP32 SDF grid (seeded multi-primitive), 16x2 hash "
                               "2^19/level, 8192 rays x (64 + 4x16) samples, hemisphere cameras; BF16 form = bf16 "
                               "hash table + bf16 MLP on matrix cores with FP32 accumulation", "form": main_form,
                   "rays_per_step": RAYS, "samples_per_ray": SAMPLES, "grid": 256, "log2_hash": 19,
                   "inv_s_reading": "init_s_val=0.3 -> s = exp(10*0.3)"},
        "samples_per_s": r["samples_per_s"],
        "rgb_max_abs_err_vs_fp64": r["rgb_max_abs_err_vs_fp64"],
        "forms": res,
        "bf16_over_fp32": (res["bf16"]["rays_per_s"] / res["fp32"]["rays_per_s"]) if len(res) == 2 else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(dom, main_form),
                     "kernel": dom, "kernel_ms": kms[dom], "bytes_per_launch": nbytes},
        "grid_gather": lat,
        "pipeline_stages": stages,
        "full_scale_render": full,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
