"""One full training iteration of textured 2DGS at scale, built around the render path:
projection, tile sort, textured rasterizer forward/backward (transmittance
reconstructed by division in the backward, nothing stored per pair), the trainer's
loss (0.8 L1 + 0.2 D-SSIM) and a fused Adam step of every parameter.

Parameters follow the reference trainer's layout and learning rates
(examples/simple_trainer_textured_gaussians.py:124-157 means, scales (log), quats,
opacities (logit), sh0, shN; :176 textures 2.5e-3; Adam eps 1e-15, betas (0.9,
0.999), :200-207) and its render call (:793-813).  Differences, for memory:
- textures are bfloat16 with a float32 master and bfloat16 moments, stepped inside
  the backward from per-pair records (``BF16TextureAdam``); at 2M x 64x64 the float32
  textures, gradient and moments (524 GB) do not fit in HBM, this state (262 GB) does;
- every optimizer is the selective Adam of the reference's ``SelectiveAdam`` (no bias
  correction, rows not visible in the step untouched) where the trainer uses
  ``torch.optim.Adam``.
"""
from dataclasses import dataclass
from typing import Dict

import torch
from torch import Tensor

from .losses import l1_dssim_loss
from .optimizers import SelectiveAdam
from .rendering import rasterization_textured_gaussians
from .texture_adam import BF16TextureAdam

GEOMETRY_LR = dict(means=1.6e-4, scales=5e-3, quats=1e-3, opacities=5e-2, sh0=2.5e-3, shN=2.5e-3 / 20)
TEXTURE_LR = 2.5e-3


@dataclass
class TrainState:
    params: Dict[str, Tensor]
    geometry_adam: SelectiveAdam
    texture_adam: BF16TextureAdam


def init_state(means: Tensor, quats: Tensor, scales: Tensor, opacities: Tensor, sh: Tensor,
               texture_adam: BF16TextureAdam, scene_scale: float = 1.0) -> TrainState:
    """From render-space Gaussians (scales, opacities activated; SH [N, K, 3]); the
    textures are already in ``texture_adam``."""
    params = dict(means=means.clone(), scales=torch.log(scales), quats=quats.clone(),
                  opacities=torch.logit(opacities.clamp(1e-6, 1 - 1e-6)), sh0=sh[:, :1].clone(),
                  shN=sh[:, 1:].clone())
    params = {k: torch.nn.Parameter(v.contiguous()) for k, v in params.items()}
    groups = [{"params": [p], "lr": GEOMETRY_LR[k] * (scene_scale if k == "means" else 1.0), "name": k}
              for k, p in params.items()]
    opt = SelectiveAdam(groups, eps=1e-15, betas=(0.9, 0.999))
    return TrainState(params, opt, texture_adam)


def train_step(state: TrainState, viewmats: Tensor, Ks: Tensor, width: int, height: int, target: Tensor,
               backgrounds=None, sh_degree: int = 3, ssim_lambda: float = 0.2) -> Tensor:
    """One iteration; returns the loss (a device scalar, not synchronised)."""
    p = state.params
    for t in p.values():
        t.grad = None
    colors = torch.cat([p["sh0"], p["shN"]], 1)
    out = rasterization_textured_gaussians(
        p["means"], p["quats"], torch.exp(p["scales"]), torch.sigmoid(p["opacities"]), colors,
        state.texture_adam.textures, viewmats, Ks, width, height, sh_degree=sh_degree, backgrounds=backgrounds,
        texture_adam=state.texture_adam)
    loss = l1_dssim_loss(out[0], target, ssim_lambda)
    loss.backward()  # also steps the textures
    visibility = (out[-1]["radii"] > 0).any(0)
    state.geometry_adam.step(visibility)
    return loss.detach()


def scene_texture_adam(n: int, res: int, device, lr: float = TEXTURE_LR, seed: int = 1000,
                       chunk_rows: int = 16384) -> BF16TextureAdam:
    """Textures of the synthetic scenes (RGB N(0, 0.1), alpha U(0, 1)) drawn on the
    device chunk by chunk straight into the bfloat16 state (no float32 copy of the
    whole set)."""
    opt = BF16TextureAdam(n, res, res, lr, device=device)
    g = torch.Generator(device=device).manual_seed(seed)
    for a in range(0, n, chunk_rows):
        b = min(n, a + chunk_rows)
        blk = torch.empty(b - a, res, res, 4, device=device)
        blk[..., :3].normal_(0.0, 0.1, generator=g)
        blk[..., 3].uniform_(0.0, 1.0, generator=g)
        opt.load_master(blk, start=a, chunk_rows=chunk_rows)
    return opt


def state_bytes(state: TrainState) -> Dict[str, int]:
    t = state.texture_adam
    tex = sum(x.numel() * x.element_size() for x in (t.textures, t.lo, t.exp_avg, t.exp_avg_sq))
    geo = sum(p.numel() * p.element_size() * 4 for p in state.params.values())  # param, grad, m, v
    return {"texture_state": tex, "geometry_state": geo}


__all__ = ["TrainState", "init_state", "train_step", "scene_texture_adam", "state_bytes", "GEOMETRY_LR",
           "TEXTURE_LR"]
