"""Pins the oracle's emulation of the BF16 matrix-core MLP (``_RadianceMlpBF16``), the
checker of the BF16 form's forward and backward on the GPU.

1. With its bf16 rounding points switched off and float64 operands, the emulation's
   forward and hand-written backward must equal float64 autograd of the reference MLP
   (``radiance_mlp``, radiance_mlp.py:81-99) to 1e-12: the contractions, the ReLU mask,
   the sigmoid derivative and the column blocks of dW0 are those of the reference.
2. With the rounding points on, its float32 results stay within a bf16-scale distance of
   the float64 reference, measured as ||a - b|| / ||b|| (bf16 rounding flips a few ReLU
   masks, which moves single entries by a whole term; measured 1e-4 .. 4.4e-2 on these
   synthetic weights, whose hidden pre-activations crowd around zero).
3. Mutants of the backward that differ from it by one edit (a dropped ReLU mask, a missing
   sigmoid derivative, the normal columns of dW0 taken from the wrong inputs, the bias
   gradient of the output layer without the sigmoid derivative) fail check 1.
"""

import inspect
import textwrap

import numpy as np
import pytest
import torch

from oracle import ref_cpu as R

MUTANTS = {
    "relu mask dropped": ("* mask.reshape(-1, 64)", ""),
    "sigmoid derivative dropped": ("(g * (1.0 - rgb) * rgb)", "(g * rgb)"),
    "dW0 normal columns from hash inputs": ("dh.T @ xf[:, 32:35]", "dh.T @ xf[:, 29:32]"),
    "db1 without the sigmoid derivative": ("dout.sum(0), None)", "g.reshape(-1, 3).sum(0), None)"),
}


def _inputs(V, dtype, n=700, seed=0):
    g = torch.Generator().manual_seed(seed)
    u = lambda *sh, s=1.0: ((torch.rand(*sh, generator=g, dtype=torch.float64) * 2 - 1) * s).to(dtype)
    feats = u(n, 32, s=1e-2)
    sh = u(n, V)
    nrm = torch.nn.functional.normalize(u(n, 3), dim=-1)
    W0, b0 = u(64, 35 + V, s=0.3), u(64, s=0.1)
    W1, b1 = u(3, 64, s=0.3), u(3, s=0.1)
    r = u(n, 3)
    return [feats, sh, nrm, W0, b0, W1, b1], r


def _grads(fn, xs, r):
    xs = [x.detach().clone() for x in xs]
    for k in (0, 2, 3, 4, 5, 6):
        xs[k].requires_grad_(True)
    out = fn(*xs)
    (out * r).sum().backward()
    return [out.detach()] + [xs[k].grad for k in (0, 2, 3, 4, 5, 6)]


def _rel(a, b):
    return float((a.double() - b.double()).abs().max() / b.double().abs().max())


def _rel_norm(a, b):
    return float((a.double() - b.double()).norm() / b.double().norm())


def _emulation_error(cls, V, rounding=False, dtype=torch.float64, metric=_rel):
    xs, r = _inputs(V, torch.float64)
    ref = _grads(R.radiance_mlp, xs, r)
    got = _grads(lambda *a: cls.apply(*a, rounding), [x.to(dtype) for x in xs], r.to(dtype))
    return [metric(a, b) for a, b in zip(got, ref)]


@pytest.mark.parametrize("V", [16, 27])
def test_emulation_without_rounding_equals_fp64_autograd(V):
    errs = _emulation_error(R._RadianceMlpBF16, V)
    print(f"V={V}: rgb, d feats, d normal, dW0, db0, dW1, db1 rel err " + " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) < 1e-12


@pytest.mark.parametrize("V", [16, 27])
def test_emulation_with_rounding_within_bf16_of_fp64(V):
    errs = _emulation_error(R._RadianceMlpBF16, V, rounding=True, dtype=torch.float32, metric=_rel_norm)
    print(f"V={V} bf16: rgb, d feats, d normal, dW0, db0, dW1, db1 norm rel err " + " ".join(f"{e:.1e}" for e in errs))
    # 8 significant bits: a rounding is within 2^-9 relative; ReLU flips add more
    assert max(errs) < 6e-2
    # and the rounding is really there: float32 alone is ~1e-7
    assert max(errs) > 1e-4


def _mutant(name):
    old, new = MUTANTS[name]
    src = textwrap.dedent(inspect.getsource(R._RadianceMlpBF16.backward))
    src = "\n".join(l for l in src.splitlines() if not l.strip().startswith("@"))
    assert src.count(old) == 1, name
    ns = {"torch": torch, "_bf16": R._bf16}
    exec(src.replace(old, new), ns)
    return type("Mutant", (torch.autograd.Function,),
                {"forward": staticmethod(R._RadianceMlpBF16.forward), "backward": staticmethod(ns["backward"])})


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_backward_mistakes_fail_the_fp64_check(name):
    errs = _emulation_error(_mutant(name), 16)
    print(f"{name}: max rel err {max(errs):.2e}")
    assert max(errs) > 1e-3


def test_render_with_bf16_emulation_close_to_fp64():
    """The whole BF16-form render (bf16 table values, emulated MLP) against the FP64 render
    on the golden config-1 rays: RGB within 5e-3, the grid / hash / MLP gradients of a
    fixed upstream gradient within 2e-2 / 5e-2 / 1e-2 in ||a - b|| / ||b||."""
    from tests.conftest import load_golden
    from voxneus_amd import scenes

    g = load_golden("step_c1.npz")
    sc = scenes.config1_scene(seed=0)
    o, d = torch.from_numpy(g["rays_o"]), torch.from_numpy(g["rays_d"])
    zs = tuple(torch.from_numpy(g[k]) for k in ("z_vals", "dists", "mid"))
    rr = torch.from_numpy(np.random.default_rng(2).normal(size=(64, 3)))
    res = {}
    for name, dt, form, emb in (("bf16", torch.float32, "bf16",
                                 torch.from_numpy(sc.embeddings).to(torch.bfloat16).float().numpy()),
                                ("fp64", torch.float64, "fp32", sc.embeddings)):
        m = R.RefModel(sc.grid, emb, **sc.mlp, hash_max_resolution=sc.hash_max_resolution, dtype=dt, mlp_form=form)
        out = m.render(o.to(dt), d.to(dt), None, None, zs=tuple(z.to(dt) for z in zs))
        (out["rgb"] * rr.to(dt)).sum().backward()
        res[name] = (out["rgb"].detach(), m.grid.grad, m.embeddings.grad,
                     torch.cat([m.W0.grad.reshape(-1), m.b0.grad, m.W1.grad.reshape(-1), m.b1.grad]))
    e_rgb = float((res["bf16"][0].double() - res["fp64"][0]).abs().max())
    e = [_rel_norm(a, b) for a, b in zip(res["bf16"][1:], res["fp64"][1:])]
    em = [_rel(a, b) for a, b in zip(res["bf16"][1:], res["fp64"][1:])]
    print(f"bf16 emulation vs fp64: rgb {e_rgb:.2e} norm grid {e[0]:.2e} hash {e[1]:.2e} mlp {e[2]:.2e}; "
          f"max entry grid {em[0]:.2e} hash {em[1]:.2e} mlp {em[2]:.2e}")
    assert e_rgb <= 5e-3 and e[0] < 2e-2 and e[1] < 5e-2 and e[2] < 1e-2
