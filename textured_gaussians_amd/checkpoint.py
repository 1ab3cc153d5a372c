"""Reader for the reference trainer's checkpoints, so trained scenes (not only the
synthetic ones) can drive the render path.

Format (examples/simple_trainer_textured_gaussians.py:1766-1778): ``ckpt_{step}.pt``
holds ``{"step": int, "splats": ParameterDict.state_dict(), ...}`` with optional
``texture_model``, ``base_color_factor``, ``sigmoid_factor`` and ``gaussian_factor``
entries.  The splats are stored in their optimisation parameterisation: log
scales, opacity logits, SH as ``sh0`` [N,1,3] + ``shN`` [N,K-1,3] (or ``colors``
with an appearance module), textures [N,L,L,4] raw.  The trainer turns them into
render inputs in ``rasterize_splats`` (:793-813) and ``get_textures`` (:445-485);
``render_inputs`` below does the same.

Only tensors, ints and floats are read: ``torch.load(weights_only=True)``.
"""
import math
import os
import re
from typing import Dict, Optional

import torch
from torch import Tensor

_CKPT = re.compile(r"^ckpt_(\d+)\.pt$")


def latest_checkpoint(ckpt_dir: str, limit: Optional[int] = None) -> Optional[str]:
    """Path of ``ckpt_{step}.pt`` with the largest step (<= limit when given), as the
    reference's loader picks it (examples/scene_args_loader.py:406-428), or None."""
    best, best_step = None, -1
    if not os.path.isdir(ckpt_dir):
        return None
    for name in os.listdir(ckpt_dir):
        m = _CKPT.match(name)
        if not m:
            continue
        step = int(m.group(1))
        if (limit is None or step <= limit) and step > best_step:
            best, best_step = os.path.join(ckpt_dir, name), step
    return best


def load_checkpoint(path: str, device=None) -> Dict:
    """The checkpoint dict with its splats on ``device`` (tensors only; nothing is
    unpickled beyond tensors and plain containers)."""
    ck = torch.load(path, map_location="cpu", weights_only=True)
    if "splats" not in ck:
        raise KeyError(f"{path}: not a trainer checkpoint (no 'splats')")
    if device is not None:
        ck["splats"] = {k: v.to(device) for k, v in ck["splats"].items()}
    return ck


def save_checkpoint(path: str, step: int, splats: Dict[str, Tensor], **extra) -> None:
    """Writes the reference trainer's checkpoint layout (for round trips and for
    handing scenes fitted here back to the reference's tools)."""
    data = {"step": int(step), "splats": {k: v.detach().cpu() for k, v in splats.items()}}
    data.update(extra)
    torch.save(data, path)


def _clamp(t: Tensor, mode: str, alpha: bool) -> Tensor:
    # examples/simple_trainer_textured_gaussians.py:452-483
    if mode == "none":
        return t
    if mode == "normalize":
        return (t / (t.amax(dim=[1, 2], keepdim=True) + 1e-6)).clamp(0.0, 1.0)
    if mode == "clamp":
        return t.clamp(0.0, 0.9999) if alpha else t.clamp(0.0, 1.0)
    if mode == "sigmoid":
        return t.sigmoid() * 0.9999 if alpha else t.sigmoid()
    raise ValueError(f"unknown clamp type {mode}")


def textures_from_splats(splats: Dict[str, Tensor], textured_rgb: bool = True, textured_alpha: bool = True,
                         rgb_clamp: str = "none", alpha_clamp: str = "none") -> Tensor:
    """[N,L,L,4] render textures from the stored ones, with the trainer's options
    (``textured_rgb``, ``textured_alpha`` and their clamp modes; :445-485)."""
    t = splats["textures"]
    rgb = _clamp(t[..., :3], rgb_clamp, False) if textured_rgb else torch.zeros_like(t[..., :3])
    a = _clamp(t[..., 3:4], alpha_clamp, True) if textured_alpha else torch.ones_like(t[..., 3:4])
    return torch.cat([rgb, a], dim=-1).contiguous()


def render_inputs(splats: Dict[str, Tensor], **texture_options) -> Dict:
    """Arguments for ``rasterization_textured_gaussians`` from stored splats:
    means, quats, exp(scales), sigmoid(opacities), colors (SH coefficients
    [N,K,3] from sh0/shN, or the stored colors) with the matching sh_degree, and
    the textures (:793-813)."""
    out = dict(means=splats["means"], quats=splats["quats"], scales=torch.exp(splats["scales"]),
               opacities=torch.sigmoid(splats["opacities"]))
    if "sh0" in splats:
        colors = torch.cat([splats["sh0"], splats["shN"]], 1) if "shN" in splats else splats["sh0"]
        out["colors"] = colors.contiguous()
        out["sh_degree"] = int(round(math.sqrt(colors.shape[1]))) - 1
    else:
        out["colors"] = splats["colors"]
        out["sh_degree"] = None
    if "textures" in splats:
        out["textures"] = textures_from_splats(splats, **texture_options)
    return out
