"""Benchmark: fused VoxNeuS fwd+bwd training step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form fp32] [--no-cpu-baseline]

Workload (the precision/throughput configuration): FP32 SDF grid 256^3 of a seeded
multi-primitive scene, 16-level x 2-feature hash with 2^19 entries per level (per-level
scale 1.5), 51->64->3 radiance MLP, 8192 rays per step drawn from 49 hemisphere pinhole
cameras (400x300, f=600), 64 stratified + 4x16 upsampled samples per ray.  One step =
ray bounds, 4-step upsampler, fused render forward + Huber loss + backward, V_R build,
closed-form eikonal (w=0.1) + curvature (w=5e-4) gradients, k=5 s-amplification
(training/trainer.py:188-326).  Inputs are resident in HBM before timing starts.

Prints one JSON line.  ``roofline`` covers the dominant kernel (the fused render
training kernel), timed live with HIP events on the stream it runs on; ``cpu_baseline``
times the CPU oracle (an op-for-op restatement of the reference's PyTorch path) on a
bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
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
# Algorithmic bytes per sample of the fused render training kernel (no cache credit):
# grid forward 32 words (8 corners + 24 face neighbours) = 128 B; hash forward
# 16 levels x 8 corners x 2 features = 256 entries -> 1024 B (FP32) / 512 B (BF16);
# grid backward 32 words read-modify-write = 256 B; hash backward 256 FP32 RMW = 2048 B;
# per-ray streaming (o, d, target, z/dist in, outputs) ~= 9 B per sample.
BYTES_PER_SAMPLE = {"fp32": 128 + 1024 + 256 + 2048 + 9, "bf16": 128 + 512 + 256 + 2048 + 9}


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


def run_gpu(sc, batches, form, steps, warmup):
    import torch

    from oracle import ref_cpu as R  # only for the hash level table
    from voxneus_amd.hash_grid import level_resolutions
    from voxneus_amd.training import FusedTrainStep, StepConfig

    dev = torch.device("cuda:0")
    table = torch.from_numpy(sc.embeddings).to(dev)
    if form == "bf16":
        table = table.to(torch.bfloat16)
    st = FusedTrainStep(torch.from_numpy(sc.grid).to(dev), table, mlp_flat(sc).to(dev),
                        torch.tensor(0.3, device=dev), level_resolutions(16, 16, sc.hash_max_resolution),
                        cfg=StepConfig(perturb=True))
    gb = [tuple(torch.from_numpy(a).to(dev) for a in b) for b in batches]
    ev = []

    def one(i, timed):
        o, d, t = gb[i % len(gb)]
        if timed:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            st.render_events = (e0, e1)
            ev.append((e0, e1))
        else:
            st.render_events = None
        return st.step(o, d, t)

    for i in range(warmup):
        one(i, False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        out = one(i, True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    loss = out["loss"].item()
    return {"ms_per_step": dt / steps * 1e3, "rays_per_s": RAYS * steps / dt, "render_kernel_ms": kern_ms,
            "loss": loss}


def rgb_error_vs_fp64(sc, batch, n_rays=256):
    """Max-abs RGB of the HIP FP32 forward vs the FP64 oracle on the same z values."""
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
    cfg = Fn.RenderConfig((-1, -1, -1, 1, 1, 1), res)
    s = torch.clamp(torch.exp(torch.tensor(0.3) * 10.0), 1e-4, 300.0)
    out = Fn.RenderFunction.apply(torch.from_numpy(sc.grid).to(dev), torch.from_numpy(sc.embeddings).to(dev),
                                  mlp_flat(sc).to(dev), s.to(dev), o.to(dev), d.to(dev), zs[2].to(dev),
                                  zs[1].to(dev), cfg, False, None)
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
    ap.add_argument("--form", default="fp32", choices=["fp32", "bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rays", type=int, default=128)
    ap.add_argument("--no-error-check", action="store_true", help="skip the FP64 RGB error check (profiling runs)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (rays shard across GPUs by data "
              "parallelism with one gradient all-reduce per step, which is not implemented here).",
              file=sys.stderr)
        sys.exit(2)
    import torch

    sc, batches = make_workload()
    log(f"GPU: {torch.cuda.get_device_name(0)}; form {a.form}; warmup {a.warmup}; steps {a.steps}")
    r = run_gpu(sc, batches, a.form, a.steps, a.warmup)
    log(f"gpu: {r}")
    nbytes = BYTES_PER_SAMPLE[a.form] * RAYS * SAMPLES
    achieved = nbytes / (r["render_kernel_ms"] * 1e-3) / 1e9
    err = None if a.no_error_check else rgb_error_vs_fp64(sc, batches[0])
    log(f"rgb max-abs err vs fp64: {err}")
    cpu = None if a.no_cpu_baseline else cpu_baseline(sc, batches[0], a.cpu_rays)
    if cpu:
        log(f"cpu baseline: {cpu}")
    line = {
        "metric": "train rays/s fwd+bwd @256^3 (BF16 vs FP32 at stated RGB err); render p50/p99 ms",
        "value": r["rays_per_s"], "unit": "rays/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if a.form == "fp32" else "bf16", "data": "synthetic",
        "config": {"workload": "fused fwd+bwd train step: 256^3 FP32 SDF grid (seeded multi-primitive), 16x2 hash "
                               "2^19/level, 8192 rays x (64 + 4x16) samples, hemisphere cameras", "form": a.form,
                   "rays_per_step": RAYS, "samples_per_ray": SAMPLES, "grid": 256, "log2_hash": 19},
        "samples_per_s": r["rays_per_s"] * SAMPLES,
        "rgb_max_abs_err_vs_fp64": err,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "render_kernel<2> (fused render train)", "kernel_ms": r["render_kernel_ms"],
                     "bytes_per_launch": nbytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
