"""The training step's colour loss variants on the GPU (trainer.py:213-245): Huber or MSE,
unmasked or with a foreground mask (weighted colour loss plus mask_weight * BCE of the
clamped opacity), against the oracle's autograd on the step's own z values: loss, its two
parts, and the MLP, grid and variance gradients within 1e-5 (regularisers off, so the grid
gradient is the render's alone).

The rays are the second golden batch (the multi-primitive scene: some rays hit, some
miss, none reaches an opacity within 1e-4 of 1).  The masks follow the silhouette (mask 1
where the opacity exceeds 0.5) with a few mid-range rays flipped: the BCE gradient
(x - m) / (x (1 - x)) of a ray with opacity x near 1 but mask 0 would amplify the
last-bit difference of the two opacity sums (summed in different orders) by 1 / (1 - x)
and test the conditioning of the loss rather than the kernel."""

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R
from tests.conftest import load_golden
from voxneus_amd import scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _mlp_flat(sc):
    return torch.cat([torch.from_numpy(sc.mlp[k]).reshape(-1) for k in ("W0", "b0", "W1", "b1")])


@pytest.mark.parametrize("loss_type,mask", [("Huber", "silhouette"), ("MSE", None), ("MSE", "silhouette"),
                                            ("Huber", "zero"), ("Huber", "ones")])
def test_loss_variants_match_oracle(loss_type, mask):
    from voxneus_amd.training import FusedTrainStep, StepConfig
    g = load_golden("step_c2.npz")
    sc = scenes.make_scene("multi", 128, 17, seed=2)
    cfg = StepConfig(w_eik=0.0, w_curv=0.0, amplify=False, loss_type=loss_type, use_mask=mask is not None,
                     mask_weight=0.25)
    st = FusedTrainStep(torch.from_numpy(sc.grid).to(DEV), torch.from_numpy(sc.embeddings).to(DEV),
                        _mlp_flat(sc).to(DEV), torch.tensor(0.3, device=DEV),
                        R.hash_level_resolutions(16, 16, sc.hash_max_resolution), cfg=cfg)
    o, d, tgt = (torch.from_numpy(g[k]) for k in ("rays_o", "rays_d", "target"))
    n = o.shape[0]
    tm = None
    if mask == "silhouette":
        m0 = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
        with torch.no_grad():
            op = m0.render(o, d, None, None, zs=tuple(torch.from_numpy(g[k]) for k in ("z_vals", "dists", "mid")))[
                "opacity"]
        tm = (op > 0.5).float()
        mid = ((op > 0.05) & (op < 0.95)).reshape(-1).nonzero().reshape(-1)[:4]
        tm[mid] = 1.0 - tm[mid]
        assert 0 < float(tm.sum()) < n
    elif mask == "zero":
        tm = torch.zeros(n, 1)
    elif mask == "ones":
        tm = torch.ones(n, 1)
    out = st.step(o.to(DEV), d.to(DEV), tgt.to(DEV), target_mask=None if tm is None else tm.to(DEV))
    m = R.RefModel(sc.grid, sc.embeddings, **sc.mlp, hash_max_resolution=sc.hash_max_resolution)
    zs = tuple(x.cpu() for x in (out["z_vals"], out["dists"], out["mid_z_vals"]))
    ref = m.render(o, d, None, None, zs=zs)
    total, lrgb, lmask = R.color_loss(ref["rgb"], tgt, loss_type, tm is not None, tm, ref["opacity"], 0.25)
    total.backward()
    lt, lc, lm = float(total.detach()), float(lrgb.detach()), 0.0 if lmask is None else float(lmask.detach())
    assert abs(float(out["loss"]) - lt) <= 1e-5 * abs(lt)
    parts = out["loss_parts"].cpu().numpy()
    assert abs(parts[0] - lc) <= 1e-5 * abs(lc) and abs(parts[1] - lm) <= 1e-5 * max(abs(lm), 1e-6)
    ref_mlp = torch.cat([m.W0.grad.reshape(-1), m.b0.grad, m.W1.grad.reshape(-1), m.b1.grad]).numpy()
    assert rel_err(st.grad_mlp.cpu().numpy(), ref_mlp) < 1e-5
    assert rel_err(st.grad_grid.cpu().numpy(), m.grid.grad.numpy()[0, 0]) < 1e-5
    assert abs(float(st.grad_v.reshape(())) - float(m.variance.grad)) <= 1e-5 * max(1e-3, abs(float(m.variance.grad)))
