"""Scenes stored in the reference trainer's checkpoint layout
(examples/simple_trainer_textured_gaussians.py:1766-1778: ``{"step", "splats": {means,
scales (log), quats, opacities (logit), sh0, shN, textures}}``), found with
``checkpoint.latest_checkpoint``, loaded onto the GPU with ``load_checkpoint`` and turned
into render inputs with ``render_inputs`` (the trainer's texture options, :445-485),
then rendered through the public API in float64 against the oracle's render of the same
stored parameters: images to 1e-9, gradients with respect to the stored (raw)
parameters to 1e-8.  Then the refinement stage (textures only, fused Adam) starting
from the loaded checkpoint: the loss falls, the refined scene saved as a later step is
the one ``latest_checkpoint`` picks, and it renders bit-identically after the reload."""
import os

import pytest
import torch

import textured_gaussians_amd as tg
from oracle.render import render as oracle_render
from textured_gaussians_amd import checkpoint as ck
from textured_gaussians_amd import scenes
from textured_gaussians_amd.neural import refine_textures

from gpu_util import DEV, cpu64, rel_err

pytestmark = pytest.mark.gpu

F64 = torch.float64
CAMS = [(3.5, 0.3, 0.4), (-2.8, 2.0, 0.9)]


def _stored(n=700, res=8, deg=2, seed=0):
    """Raw parameters as the trainer stores them."""
    g = torch.Generator().manual_seed(seed)
    q = torch.randn(n, 4, generator=g)
    return {
        "means": torch.randn(n, 3, generator=g) * 0.7,
        "scales": torch.randn(n, 3, generator=g) * 0.4 - 3.0,
        "quats": q,
        "opacities": torch.randn(n, generator=g) + 0.5,
        "sh0": torch.randn(n, 1, 3, generator=g) * 0.6,
        "shN": torch.randn(n, (deg + 1) ** 2 - 1, 3, generator=g) * 0.15,
        "textures": torch.rand(n, res, res, 4, generator=g) * 1.4 - 0.2,
    }


def _cameras(w=64, h=48):
    vm = torch.stack([scenes.look_at_viewmat(c, dtype=F64) for c in CAMS])
    K = scenes.intrinsics(w, h, 50.0, dtype=F64)[None].repeat(len(CAMS), 1, 1)
    return vm, K, w, h


def _render(ri, vm, K, w, h, bg):
    return tg.rasterization_textured_gaussians(ri["means"], ri["quats"], ri["scales"], ri["opacities"], ri["colors"],
                                               ri["textures"], vm, K, w, h, sh_degree=ri["sh_degree"], backgrounds=bg,
                                               grad_mode="exact")


@pytest.mark.parametrize("opts", [{}, {"rgb_clamp": "clamp", "alpha_clamp": "sigmoid"},
                                  {"textured_alpha": False, "rgb_clamp": "sigmoid"},
                                  {"rgb_clamp": "normalize", "alpha_clamp": "clamp"}])
def test_loaded_checkpoint_renders_as_the_oracle(tmp_path, opts):
    sp = _stored()
    ck.save_checkpoint(os.path.join(tmp_path, "ckpt_999.pt"), 999, _stored(seed=1))
    ck.save_checkpoint(os.path.join(tmp_path, "ckpt_6999.pt"), 6999, sp, gaussian_factor=1.0)
    path = ck.latest_checkpoint(str(tmp_path))
    assert os.path.basename(path) == "ckpt_6999.pt"
    data = ck.load_checkpoint(path, device=DEV)
    assert data["step"] == 6999 and data["splats"]["means"].is_cuda
    vm, K, w, h = _cameras()
    bg = torch.tensor([[0.2, 0.5, 0.9], [1.0, 1.0, 1.0]], dtype=F64)
    raw = {k: v.to(F64).clone().requires_grad_(True) for k, v in data["splats"].items()}
    ri = ck.render_inputs(raw, **opts)
    assert ri["sh_degree"] == 2
    out = _render(ri, vm.to(DEV), K.to(DEV), w, h, bg.to(DEV))
    assert float((out[-1]["radii"] > 0).float().mean()) > 0.5
    g = torch.Generator().manual_seed(4)
    wc = torch.randn(out[0].shape, generator=g, dtype=F64)
    wa = torch.randn(out[1].shape, generator=g, dtype=F64)
    ((out[0] * wc.to(DEV)).sum() + (out[1] * wa.to(DEV)).sum()).backward()
    # the oracle: the same stored parameters through the same stored-to-render maps, on the CPU
    rawc = {k: v.to(F64).clone().requires_grad_(True) for k, v in sp.items()}
    ric = ck.render_inputs(rawc, **opts)
    ref = oracle_render(ric["means"], ric["quats"], ric["scales"], ric["opacities"], ric["colors"], vm, K, w, h,
                        textures=ric["textures"], sh_degree=ric["sh_degree"], backgrounds=bg, uv_grad="exact")
    assert (cpu64(out[0]) - ref["render_colors"]).abs().max().item() < 1e-9
    assert (cpu64(out[1]) - ref["render_alphas"]).abs().max().item() < 1e-9
    ((ref["render_colors"] * wc).sum() + (ref["render_alphas"] * wa).sum()).backward()
    for k in sp:
        assert rel_err(raw[k].grad, rawc[k].grad) < 1e-8, (k, rel_err(raw[k].grad, rawc[k].grad))


def test_refinement_from_a_checkpoint_round_trips(tmp_path):
    """float32, as the trainer refines (the fused Adam takes float32 parameters)."""
    sp = _stored(n=500, res=8, deg=1, seed=2)
    ck.save_checkpoint(os.path.join(tmp_path, "ckpt_2999.pt"), 2999, sp)
    data = ck.load_checkpoint(ck.latest_checkpoint(str(tmp_path)), device=DEV)
    vm, K, w, h = _cameras()
    vm, K = vm.float().to(DEV), K.float().to(DEV)
    ri = ck.render_inputs(data["splats"])
    # targets: the same scene with other textures (what refinement has to recover)
    tgt_tex = torch.rand(ri["textures"].shape, generator=torch.Generator().manual_seed(7)).to(DEV)
    with torch.no_grad():
        targets = _render(dict(ri, textures=tgt_tex), vm, K, w, h, None)[0]

    def render(tex):
        return _render(dict(ri, textures=tex), vm, K, w, h, None)[0]
    with torch.no_grad():
        before = float(((render(ri["textures"]) - targets) ** 2).mean())
    refined, seen = refine_textures(render, ri["textures"], targets, steps=40, lr=0.05)
    with torch.no_grad():
        after = float(((render(refined) - targets) ** 2).mean())
    assert after < 0.5 * before, (before, after)
    assert float(seen.sum()) > 0
    # store the refined textures as a later step (rgb and alpha as rendered: no clamp modes)
    out_sp = {k: v.detach().cpu() for k, v in data["splats"].items()}
    out_sp["textures"] = refined.cpu()
    ck.save_checkpoint(os.path.join(tmp_path, "ckpt_3039.pt"), 3039, out_sp)
    path = ck.latest_checkpoint(str(tmp_path))
    assert os.path.basename(path) == "ckpt_3039.pt"
    again = ck.render_inputs(ck.load_checkpoint(path, device=DEV)["splats"])
    with torch.no_grad():
        a = render(again["textures"])
        b = render(refined)
    assert torch.equal(a, b)
    # an earlier limit still finds the original
    assert os.path.basename(ck.latest_checkpoint(str(tmp_path), limit=3000)) == "ckpt_2999.pt"
