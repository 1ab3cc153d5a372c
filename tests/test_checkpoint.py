"""The reference trainer's checkpoint layout (examples/simple_trainer_textured_gaussians.py:1766-1778)
read back into render inputs (:793-813, :445-485).  The checkpoints here are written in that
layout by the test itself (the reference ships none)."""
import math
import os

import pytest
import torch

from textured_gaussians_amd import checkpoint as ck


def _splats(n=50, L=8, deg=3, seed=0):
    g = torch.Generator().manual_seed(seed)
    return {
        "means": torch.randn(n, 3, generator=g),
        "scales": torch.randn(n, 3, generator=g) - 3,
        "quats": torch.randn(n, 4, generator=g),
        "opacities": torch.randn(n, generator=g),
        "sh0": torch.randn(n, 1, 3, generator=g),
        "shN": torch.randn(n, (deg + 1) ** 2 - 1, 3, generator=g),
        "textures": torch.randn(n, L, L, 4, generator=g),
    }


def test_round_trip_and_render_inputs(tmp_path):
    sp = _splats()
    path = os.path.join(tmp_path, "ckpt_6999.pt")
    ck.save_checkpoint(path, 6999, sp, gaussian_factor=1.0)
    data = ck.load_checkpoint(path)
    assert data["step"] == 6999 and data["gaussian_factor"] == 1.0
    ri = ck.render_inputs(data["splats"])
    assert torch.equal(ri["scales"], torch.exp(sp["scales"]))
    assert torch.equal(ri["opacities"], torch.sigmoid(sp["opacities"]))
    assert ri["colors"].shape == (50, 16, 3) and ri["sh_degree"] == 3
    assert torch.equal(ri["colors"][:, 0], sp["sh0"][:, 0]) and torch.equal(ri["colors"][:, 1:], sp["shN"])
    assert torch.equal(ri["textures"], sp["textures"])


def test_texture_options_follow_the_trainer():
    sp = _splats(seed=1)
    t = sp["textures"]
    out = ck.textures_from_splats(sp, textured_rgb=False, textured_alpha=False)
    assert torch.all(out[..., :3] == 0) and torch.all(out[..., 3] == 1)
    out = ck.textures_from_splats(sp, rgb_clamp="clamp", alpha_clamp="clamp")
    assert float(out[..., :3].max()) <= 1.0 and float(out[..., 3].max()) <= 0.9999
    out = ck.textures_from_splats(sp, rgb_clamp="sigmoid", alpha_clamp="sigmoid")
    assert torch.allclose(out[..., :3], t[..., :3].sigmoid()) and torch.allclose(out[..., 3], t[..., 3].sigmoid() * 0.9999)
    out = ck.textures_from_splats(sp, rgb_clamp="normalize", alpha_clamp="normalize")
    ref = (t[..., 3:4] / (t[..., 3:4].amax(dim=[1, 2], keepdim=True) + 1e-6)).clamp(0, 1)
    assert torch.allclose(out[..., 3:4], ref)
    with pytest.raises(ValueError):
        ck.textures_from_splats(sp, rgb_clamp="bogus")


def test_latest_checkpoint_and_errors(tmp_path):
    for s in (99, 1999, 29999):
        ck.save_checkpoint(os.path.join(tmp_path, f"ckpt_{s}.pt"), s, _splats(n=3))
    open(os.path.join(tmp_path, "ckpt_latest.pt"), "w").close()  # not a ckpt_{step}.pt name
    assert ck.latest_checkpoint(str(tmp_path)).endswith("ckpt_29999.pt")
    assert ck.latest_checkpoint(str(tmp_path), limit=2000).endswith("ckpt_1999.pt")
    assert ck.latest_checkpoint(os.path.join(tmp_path, "missing")) is None
    torch.save({"step": 1}, os.path.join(tmp_path, "other.pt"))
    with pytest.raises(KeyError):
        ck.load_checkpoint(os.path.join(tmp_path, "other.pt"))


def test_degree_from_coefficient_count():
    for deg in range(4):
        assert ck.render_inputs(_splats(deg=deg))["sh_degree"] == deg
    sp = _splats()
    del sp["shN"]
    assert ck.render_inputs(sp)["sh_degree"] == 0
