"""Write the golden vectors that pin the CPU oracle to the reference.

Run once, in the build container only (the reference checkout is not present on GPU
hosts):

    python tests/golden/make_golden.py /root/reference/Code [fixture-name ...]

(with fixture names, only those fixtures are written)

It imports the reference's pure-PyTorch path (``force_pytorch=True``; no compiled backend
exists on this host) on the CPU, feeds it the seeded synthetic scenes from
``voxneus_amd.scenes``, and stores only inputs' SHA-256 digests and the reference's
outputs as compressed ``.npz`` fixtures next to this script.  The training-step fixtures
are produced by the reference's own ``VoxNeuSTrainer.train_step`` (its optimizer step is
replaced by a no-op so the gradients it produced can be read back).

Fixtures:
  * ``unit_grid.npz``    - query_sdf / query_normals (sdf, grad, normal, corners, weights)
                           and the grid gradient of a random linear loss, 64^3 sphere grid;
  * ``unit_grid_analytical.npz`` - the same with gradient_mode="analytical" (Eq. 4:
                           compute_analytical_gradient) and its grid gradient;
  * ``unit_hash.npz``    - hash encoding (16 levels, T=2^15, per-level scale 1.5) with and
                           without the lambda mask, plus embedding gradients;
  * ``unit_mlp.npz``     - SH embedding + radiance MLP outputs and parameter gradients;
  * ``unit_alpha.npz``   - NeuS alpha, compositing and their input gradients;
  * ``unit_occupancy.npz`` - three OccupancyGrid.update_from_sdf updates (seeded draws);
  * ``unit_sampler.npz`` - 4-step upsampler (deterministic and injected-uniform), with a
                           seeded occupancy mask: per-step weights and new samples;
  * ``step_c1.npz`` / ``step_c2.npz`` - full training-step composition (render, Huber
                           loss, backward, V_R, eikonal, curvature, s-amplification).
"""

from __future__ import annotations

import hashlib
import os
import sys
import tempfile

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from voxneus_amd import scenes  # noqa: E402


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def det(t):
    return t.detach().cpu().numpy()


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"wrote {path} ({os.path.getsize(path) / 1024:.1f} KiB)")


def sparse(a):
    a = np.asarray(a).reshape(-1)
    nz = np.nonzero(a)[0].astype(np.int32)
    return nz, a[nz].astype(a.dtype)


def main(ref_code: str, only=()):
    sys.path.insert(0, ref_code)
    torch.set_num_threads(1)
    from voxneus.models.sdf_grid import DenseSDFGrid
    from voxneus.models.radiance_hash import PyTorchHashGrid
    from voxneus.models.embedder import SphericalHarmonicsEmbedder
    from voxneus.models.radiance_mlp import RadianceMLP
    from voxneus.models.voxneus import VoxNeuS
    from voxneus.rendering import sampler as ref_sampler
    from voxneus.rendering.neus_alpha import compute_neus_alpha
    from voxneus.rendering.volume_render import volume_render_from_samples
    from voxneus.rendering.occupancy_grid import OccupancyGrid
    from voxneus.rendering.rays import near_far_from_sphere
    import voxneus.training.trainer as ref_trainer

    # ------------------------------------------------------------------ unit_grid
    sc = scenes.config1_scene(seed=0)
    g = DenseSDFGrid(resolution=64)
    with torch.no_grad():
        g.grid.copy_(torch.from_numpy(sc.grid)[None, None])
    rng = np.random.default_rng(11)
    pts = rng.uniform(-1.1, 1.1, size=(4096, 3)).astype(np.float32)
    pts[:64] = rng.uniform(-1, 1, size=(64, 3)).astype(np.float32).round(1)  # lattice-ish / exact faces
    pts[64:72] = np.array([[-1, -1, -1], [1, 1, 1], [1, -1, 1], [-1, 1, -1], [0, 0, 0],
                           [1.0, 0.3, -0.2], [-1.0, 0.5, 1.0], [0.99999, -0.99999, 0.5]], np.float32)
    xyz = torch.from_numpy(pts)
    sdf = g.query_sdf(xyz)
    grad, normal, corners, weights = g.query_normals(xyz)
    r_g = rng.normal(size=(4096, 3)).astype(np.float32)
    r_s = rng.normal(size=(4096, 1)).astype(np.float32)
    loss = (grad * torch.from_numpy(r_g)).sum() + (sdf * torch.from_numpy(r_s)).sum()
    loss.backward()
    gi, gv = sparse(det(g.grid.grad))
    if not only or "unit_grid.npz" in only:
        save("unit_grid.npz", inputs_sha=sha(sc.grid, pts), pts=pts, r_g=r_g, r_s=r_s, sdf=det(sdf),
             grad=det(grad), normal=det(normal), corners=det(corners).astype(np.int16), weights=det(weights),
             grid_grad_idx=gi, grid_grad_val=gv)
    # the Eq. 4 analytical gradient (gradient_mode="analytical", ablation E1)
    g.grid.grad = None
    grad_a, normal_a, _, _ = g.query_normals(xyz, gradient_mode="analytical")
    (grad_a * torch.from_numpy(r_g)).sum().backward()
    ai, av = sparse(det(g.grid.grad))
    if not only or "unit_grid_analytical.npz" in only:
        save("unit_grid_analytical.npz", inputs_sha=sha(sc.grid, pts), pts=pts, r_g=r_g, grad=det(grad_a),
             normal=det(normal_a), grid_grad_idx=ai, grid_grad_val=av)
    # ------------------------------------------------------------------ unit_occupancy
    # OccupancyGrid.update_from_sdf over the config-1 grid: three updates with seeded
    # torch.randint draws (replayed from the same seeds to store them), s = exp(3), exp(3),
    # 60; the density grid and mask after each update
    if not only or "unit_occupancy.npz" in only:
        occ = OccupancyGrid(resolution=32)
        s_vals = (float(torch.exp(torch.tensor(3.0))), float(torch.exp(torch.tensor(3.0))), 60.0)
        rec = {}
        for k, s_v in enumerate(s_vals):
            torch.manual_seed(100 + k)
            occ.update_from_sdf(g.query_sdf, torch.tensor(s_v), subsample_ratio=0.25)
            torch.manual_seed(100 + k)
            n = max(int(32 ** 3 * 0.25), 1024)
            draws = [torch.randint(0, 32, (n,)) for _ in range(3)]
            rec[f"idx{k}"] = torch.stack(draws).numpy().astype(np.int16)
            rec[f"density{k}"] = det(occ.density_grid).copy()
            rec[f"mask{k}"] = np.packbits(det(occ.occupancy_mask).reshape(-1))
        save("unit_occupancy.npz", inputs_sha=sha(sc.grid), s_vals=np.array(s_vals, np.float32), **rec)
    if only and not any(n not in ("unit_grid.npz", "unit_grid_analytical.npz", "unit_occupancy.npz")
                        for n in only):
        return

    # ------------------------------------------------------------------ unit_hash
    maxres = scenes.hash_max_resolution(16, 1.5, 16)
    hg = PyTorchHashGrid(n_levels=16, n_features_per_level=2, log2_hashmap_size=15, base_resolution=16,
                         max_resolution=maxres)
    with torch.no_grad():
        hg.embeddings.copy_(torch.from_numpy(sc.embeddings))
    x = rng.uniform(0.0, 1.0, size=(4096, 3)).astype(np.float32)
    x[:8] = np.array([[0, 0, 0], [1, 1, 1], [1, 0, 1], [0.5, 0.5, 0.5], [0.25, 0.75, 1.0],
                      [1e-7, 1 - 1e-7, 0.5], [0.999999, 0.0, 0.3], [0.123, 0.456, 0.789]], np.float32)
    xt = torch.from_numpy(x)
    feats = hg(xt)
    feats_l = hg(xt, active_levels=5.5)
    r_h = rng.normal(size=feats.shape).astype(np.float32)
    (feats * torch.from_numpy(r_h)).sum().backward()
    eg = det(hg.embeddings.grad)
    hi, hv = sparse(eg)
    keep = np.random.default_rng(98).choice(hi.shape[0], size=min(6000, hi.shape[0]), replace=False)
    save("unit_hash.npz", inputs_sha=sha(sc.embeddings, x), x=x, b=np.float64(hg.b),
         level_res=np.array([int(np.floor(hg.base_resolution * (hg.b ** l))) for l in range(16)], np.int32),
         feats=det(feats), feats_lambda=det(feats_l), r=r_h, emb_grad_idx=hi[keep], emb_grad_val=hv[keep],
         emb_grad_nnz=np.int64(hi.shape[0]), emb_grad_sum=np.float64(eg.astype(np.float64).sum()),
         emb_grad_abs=np.float64(np.abs(eg.astype(np.float64)).sum()))

    # ------------------------------------------------------------------ unit_mlp
    mlp = RadianceMLP(hash_feature_dim=32, view_embed_dim=16, normal_dim=3, hidden_dim=64)
    with torch.no_grad():
        mlp.net[0].weight.copy_(torch.from_numpy(sc.mlp["W0"]))
        mlp.net[0].bias.copy_(torch.from_numpy(sc.mlp["b0"]))
        mlp.net[2].weight.copy_(torch.from_numpy(sc.mlp["W1"]))
        mlp.net[2].bias.copy_(torch.from_numpy(sc.mlp["b1"]))
    sh = SphericalHarmonicsEmbedder(4)
    hf = rng.normal(0, 1e-2, size=(512, 32)).astype(np.float32)
    dirs = scenes.normalize(rng.normal(size=(512, 3))).astype(np.float32)
    nrm = scenes.normalize(rng.normal(size=(512, 3))).astype(np.float32)
    hft = torch.from_numpy(hf).requires_grad_(True)
    nrt = torch.from_numpy(nrm).requires_grad_(True)
    she = sh(torch.from_numpy(dirs))
    rgb = mlp(hft, she, nrt)
    r_m = rng.normal(size=(512, 3)).astype(np.float32)
    (rgb * torch.from_numpy(r_m)).sum().backward()
    save("unit_mlp.npz", inputs_sha=sha(hf, dirs, nrm), hf=hf, dirs=dirs, nrm=nrm, r=r_m, sh=det(she), rgb=det(rgb),
         dW0=det(mlp.net[0].weight.grad), db0=det(mlp.net[0].bias.grad), dW1=det(mlp.net[2].weight.grad),
         db1=det(mlp.net[2].bias.grad), dhf=det(hft.grad), dnrm=det(nrt.grad))

    # ------------------------------------------------------------------ unit_alpha
    R, S = 64, 128
    sdf_a = rng.normal(0, 0.05, size=(R, S)).astype(np.float32)
    nrm_a = scenes.normalize(rng.normal(size=(R, S, 3))).astype(np.float32)
    d_a = scenes.normalize(rng.normal(size=(R, 3))).astype(np.float32)
    dists_a = rng.uniform(0.001, 0.05, size=(R, S)).astype(np.float32)
    z_a = np.cumsum(dists_a, -1).astype(np.float32) + 1.0
    col_a = rng.uniform(0, 1, size=(R, S, 3)).astype(np.float32)
    pts_a = rng.uniform(-0.8, 0.8, size=(R, S, 3)).astype(np.float32)
    pts_a[:4] *= 1.6  # some samples outside the unit sphere
    ts = [torch.from_numpy(a).requires_grad_(True) for a in (sdf_a, nrm_a, col_a)]
    s_t = torch.tensor(20.085536923187668, dtype=torch.float32, requires_grad=True)
    out = volume_render_from_samples(colors=ts[2], sdf=ts[0], normals=ts[1], ray_dirs=torch.from_numpy(d_a),
                                     z_vals=torch.from_numpy(z_a), dists=torch.from_numpy(dists_a), s=s_t,
                                     pts=torch.from_numpy(pts_a), sphere_radius=1.0)
    r_rgb = rng.normal(size=(R, 3)).astype(np.float32)
    r_d = rng.normal(size=(R, 1)).astype(np.float32)
    r_o = rng.normal(size=(R, 1)).astype(np.float32)
    r_n = rng.normal(size=(R, 3)).astype(np.float32)
    L = ((out["rgb"] * torch.from_numpy(r_rgb)).sum() + (out["depth"] * torch.from_numpy(r_d)).sum()
         + (out["opacity"] * torch.from_numpy(r_o)).sum() + (out["normal"] * torch.from_numpy(r_n)).sum())
    L.backward()
    a_single = compute_neus_alpha(torch.tensor([[0.0], [1.0]]), torch.tensor([[0.0, 0.0, 1.0]] * 2),
                                  torch.tensor([[0.0, 0.0, -1.0]] * 2), torch.tensor([[0.05], [0.05]]),
                                  torch.tensor(50.0))
    save("unit_alpha.npz", inputs_sha=sha(sdf_a, nrm_a, d_a, dists_a, z_a, col_a, pts_a), sdf=sdf_a, nrm=nrm_a,
         d=d_a, dists=dists_a, z=z_a, col=col_a, pts=pts_a, s=np.float32(s_t.item()), r_rgb=r_rgb, r_d=r_d,
         r_o=r_o, r_n=r_n, rgb=det(out["rgb"]), depth=det(out["depth"]), normal=det(out["normal"]),
         opacity=det(out["opacity"]), weights=det(out["weights"]), alpha=det(out["alpha"]),
         transmittance=det(out["transmittance"]), d_sdf=det(ts[0].grad), d_nrm=det(ts[1].grad),
         d_col=det(ts[2].grad), d_s=np.float32(s_t.grad.item()), alpha_probe=det(a_single))

    # ------------------------------------------------------------------ unit_sampler
    occ = OccupancyGrid(resolution=32)
    occ_mask = np.random.default_rng(5).uniform(size=(32, 32, 32)) < 0.7
    with torch.no_grad():
        occ.occupancy_mask.copy_(torch.from_numpy(occ_mask))
    o, d = scenes.sphere_rays(256, seed=21)
    ot, dt = torch.from_numpy(o), torch.from_numpy(d)
    near, far = near_far_from_sphere(ot, dt, radius=1.0)
    sampler = ref_sampler.RaySampler(64, 64, perturb=False)
    fn = lambda p: g.query_sdf(p).squeeze(-1)
    rec = {}
    orig_pdf = ref_sampler.sample_pdf

    def rec_pdf(bins, weights, n_samples, det=False, eps=1e-5):
        rec.setdefault("weights", []).append(det_np(weights))
        out = orig_pdf(bins, weights, n_samples, det=det, eps=eps)
        rec.setdefault("new_z", []).append(det_np(out))
        return out

    det_np = lambda t: t.detach().cpu().numpy().copy()
    ref_sampler.sample_pdf = rec_pdf
    try:
        with torch.no_grad():
            z1, dist1, mid1 = sampler.sample_multistep_sdf(ot, dt, near, far, fn, occupancy_grid=occ)
        det_rec = rec
        rec = {}
        # stochastic branch with the reference's own torch.rand / rand_like, captured
        draws = []
        orig_rand, orig_rand_like = torch.rand, torch.rand_like

        def rand(*a, **k):
            r = orig_rand(*a, **k)
            draws.append(r.numpy().copy())
            return r

        def rand_like(*a, **k):
            r = orig_rand_like(*a, **k)
            draws.append(r.numpy().copy())
            return r

        torch.manual_seed(123)
        torch.rand, torch.rand_like = rand, rand_like
        try:
            sp = ref_sampler.RaySampler(64, 64, perturb=True)
            with torch.no_grad():
                z2, dist2, mid2 = sp.sample_multistep_sdf(ot, dt, near, far, fn, occupancy_grid=occ)
        finally:
            torch.rand, torch.rand_like = orig_rand, orig_rand_like
    finally:
        ref_sampler.sample_pdf = orig_pdf
    save("unit_sampler.npz", inputs_sha=sha(sc.grid, o, d, occ_mask), rays_o=o, rays_d=d,
         occ_mask=np.packbits(occ_mask.reshape(-1)), occ_res=np.int32(32), near=det(near), far=det(far),
         z_det=det(z1), dists_det=det(dist1), mid_det=det(mid1),
         w_det=np.stack([np.pad(w, ((0, 0), (0, 127 - w.shape[1]))) for w in det_rec["weights"]], 0),
         newz_det=np.stack(det_rec["new_z"], 0),
         z_rnd=det(z2), dists_rnd=det(dist2), mid_rnd=det(mid2), t_rand=draws[0],
         u_rnd=np.stack(draws[1:], 0), newz_rnd=np.stack(rec["new_z"], 0))

    # ------------------------------------------------------------------ training steps
    def step_fixture(name, scene, o, d, target, R):
        model = VoxNeuS(grid_res=scene.grid_res, hash_log2_size=int(np.log2(scene.embeddings.shape[1])),
                        force_pytorch=True)
        enc = PyTorchHashGrid(n_levels=16, n_features_per_level=2,
                              log2_hashmap_size=int(np.log2(scene.embeddings.shape[1])), base_resolution=16,
                              max_resolution=scene.hash_max_resolution)
        model.radiance_hash.encoder = enc
        with torch.no_grad():
            model.sdf_grid.grid.copy_(torch.from_numpy(scene.grid)[None, None])
            enc.embeddings.copy_(torch.from_numpy(scene.embeddings))
            model.radiance_mlp.net[0].weight.copy_(torch.from_numpy(scene.mlp["W0"]))
            model.radiance_mlp.net[0].bias.copy_(torch.from_numpy(scene.mlp["b0"]))
            model.radiance_mlp.net[2].weight.copy_(torch.from_numpy(scene.mlp["W1"]))
            model.radiance_mlp.net[2].bias.copy_(torch.from_numpy(scene.mlp["b1"]))

        class Rays:
            def sample_rays(self, n):
                return {"rays_o": torch.from_numpy(o), "rays_d": torch.from_numpy(d),
                        "rgb": torch.from_numpy(target)}

        cfg = {"training": {"batch_size": R, "use_mask": False, "loss_type": "Huber", "progressive_hash": False,
                            "occupancy_update_interval": 10 ** 9, "sr_milestones": [10 ** 9, 2 * 10 ** 9],
                            "val_interval": 10 ** 9, "save_interval": 10 ** 9},
               "sampler": {"n_samples": 64, "n_importance": 64, "perturb": False, "up_sample_steps": 4},
               "scheduler": {"w_eik_init": 0.1, "w_curv_init": 5e-4, "s_amplification_k": 5.0,
                             "s_warmup_steps": 0},
               "logging": {"dashboard_interval": 10 ** 9, "print_interval": 10 ** 9}}
        cap = {}
        with tempfile.TemporaryDirectory() as td:
            tr = ref_trainer.VoxNeuSTrainer(model, Rays(), cfg, output_dir=td, device=torch.device("cpu"))
            for opt in tr.optimizers.values():
                opt.step = lambda *a, **k: None
            orig_render = model.render

            def render(*a, **k):
                out = orig_render(*a, **k)
                cap["render"] = {kk: det_np(v) for kk, v in out.items() if torch.is_tensor(v)}
                return out

            model.render = render
            orig_eik = ref_trainer.compute_eikonal_loss_and_gradient
            orig_amp = ref_trainer.amplify_s_gradient
            orig_samp = tr.sampler.sample_multistep_sdf

            def samp(*a, **k):
                out = orig_samp(*a, **k)
                cap["z"] = [det_np(t) for t in out]
                return out

            def eik(*a, **k):
                cap["vr"] = det_np(k["active_vertices"])
                return orig_eik(*a, **k)

            def amp(p, k=5.0):
                cap["dv_pre"] = float(p.grad.item())
                r = orig_amp(p, k)
                cap["dv_post"] = float(p.grad.item())
                return r

            tr.sampler.sample_multistep_sdf = samp
            ref_trainer.compute_eikonal_loss_and_gradient = eik
            ref_trainer.amplify_s_gradient = amp
            try:
                metrics = tr.train_step(1)
            finally:
                ref_trainer.compute_eikonal_loss_and_gradient = orig_eik
                ref_trainer.amplify_s_gradient = orig_amp
        rr = cap["render"]
        gg = det(model.sdf_grid.grid.grad)[0, 0]
        gi, gv = sparse(gg)
        eg = det(enc.embeddings.grad)
        ei, ev = sparse(eg)
        keep = np.random.default_rng(99).choice(ei.shape[0], size=min(4000, ei.shape[0]), replace=False)
        vr = cap["vr"]
        X = scene.grid_res
        vr_lin = (vr[:, 0] * X * X + vr[:, 1] * X + vr[:, 2]).astype(np.int32)
        save(name, inputs_sha=sha(scene.grid, scene.embeddings, scene.mlp["W0"], scene.mlp["W1"], o, d, target),
             rays_o=o, rays_d=d, target=target, z_vals=cap["z"][0], dists=cap["z"][1], mid=cap["z"][2],
             rgb=rr["rgb"], depth=rr["depth"], normal=rr["normal"], opacity=rr["opacity"], weights=rr["weights"],
             alpha=rr["alpha"], transmittance=rr["transmittance"], sdf_samples=rr["sdf_samples"],
             grad_samples=rr["grad_samples"], normal_samples=rr["normal_samples"],
             corner_c0=rr["corner_indices"][..., 0, :].astype(np.int16),
             loss_rgb=np.float64(metrics["loss_rgb"]), loss_eik=np.float64(metrics["loss_eik"]),
             loss_curv=np.float64(metrics["loss_curv"]), s=np.float64(metrics["s"]),
             grid_grad_idx=gi, grid_grad_val=gv, emb_grad_nnz=np.int64(ei.shape[0]),
             emb_grad_idx=ei[keep], emb_grad_val=ev[keep], emb_grad_sum=np.float64(eg.astype(np.float64).sum()),
             emb_grad_abs=np.float64(np.abs(eg.astype(np.float64)).sum()),
             dW0=det(model.radiance_mlp.net[0].weight.grad), db0=det(model.radiance_mlp.net[0].bias.grad),
             dW1=det(model.radiance_mlp.net[2].weight.grad), db1=det(model.radiance_mlp.net[2].bias.grad),
             dv_pre=np.float64(cap["dv_pre"]), dv_post=np.float64(cap["dv_post"]),
             amplified=np.bool_(metrics["s_amplified"]), vr_lin=vr_lin)

    o1, d1, t1 = scenes.config_rays("sphere", 512, seed=7)
    step_fixture("step_c1.npz", sc, o1[:64], d1[:64], t1[:64], 64)
    sc2 = scenes.make_scene("multi", 128, 17, seed=2)
    o2, d2 = scenes.hemisphere_ray_batch(32, seed=8)
    t2 = scenes.procedural_rgb(sc2.shape, o2, d2)
    step_fixture("step_c2.npz", sc2, o2, d2, t2, 32)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/Code", tuple(sys.argv[2:]))
