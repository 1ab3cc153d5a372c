"""Synthetic scenes of the benchmark workloads (no datasets are needed).

Cameras are pinhole with fx = fy = (W/2) / tan(fov/2), principal point at the
image centre, looking at the origin with world +z up (OpenCV camera axes:
x right, y down, z forward).  Generation order: means, log-scales, quats,
opacity logits, colours/SH, texture RGB, texture A.
"""
import math
from dataclasses import dataclass
from typing import Optional

import torch


def look_at_viewmat(eye, target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), dtype=torch.float32) -> torch.Tensor:
    """World-to-camera 4x4 for a camera at ``eye`` looking at ``target``."""
    eye = torch.tensor(eye, dtype=torch.float64)
    f = torch.tensor(target, dtype=torch.float64) - eye
    f = f / f.norm()
    r = torch.linalg.cross(f, torch.tensor(up, dtype=torch.float64))
    r = r / r.norm()
    d = torch.linalg.cross(f, r)
    R = torch.stack([r, d, f], 0)
    V = torch.eye(4, dtype=torch.float64)
    V[:3, :3] = R
    V[:3, 3] = -R @ eye
    return V.to(dtype)


def intrinsics(width: int, height: int, fov_deg: float, dtype=torch.float32) -> torch.Tensor:
    fx = (width / 2.0) / math.tan(math.radians(fov_deg) / 2.0)
    return torch.tensor([[fx, 0.0, width / 2.0], [0.0, fx, height / 2.0], [0.0, 0.0, 1.0]], dtype=dtype)


@dataclass
class Scene:
    means: torch.Tensor
    quats: torch.Tensor
    scales: torch.Tensor
    opacities: torch.Tensor
    colors: torch.Tensor          # [N,3] post-activation or SH coefficients [N,K,3]
    textures: Optional[torch.Tensor]
    viewmats: torch.Tensor
    Ks: torch.Tensor
    width: int
    height: int
    sh_degree: Optional[int]
    backgrounds: torch.Tensor
    target: torch.Tensor

    def to(self, device=None, dtype=None):
        def cv(t):
            if t is None:
                return None
            t = t.to(device) if device is not None else t
            return t.to(dtype) if (dtype is not None and t.is_floating_point()) else t
        return Scene(cv(self.means), cv(self.quats), cv(self.scales), cv(self.opacities), cv(self.colors),
                     cv(self.textures), cv(self.viewmats), cv(self.Ks), self.width, self.height, self.sh_degree,
                     cv(self.backgrounds), cv(self.target))


def make_scene(n: int, width: int, height: int, fov_deg: float, cam_positions, tex_res: Optional[int],
               mean_dist: str = "uniform", mean_scale: float = 1.0, logscale_mu: float = -3.5,
               logscale_sigma: float = 0.5, sh_degree: Optional[int] = 0, sh_sigma: Optional[float] = None,
               seed: int = 0, target_seed: int = 1, device="cpu", texture_device=None,
               dtype=torch.float32) -> Scene:
    """Generic generator; ``texture_device`` lets the (large) textures be drawn
    directly on the GPU with a device generator of the same seed."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    if mean_dist == "uniform":
        means = (torch.rand(n, 3, generator=g) * 2 - 1) * mean_scale
    else:
        means = torch.randn(n, 3, generator=g) * mean_scale
    scales = torch.exp(torch.randn(n, 3, generator=g) * logscale_sigma + logscale_mu)
    quats = torch.randn(n, 4, generator=g)
    quats = quats / quats.norm(dim=-1, keepdim=True)
    opacities = torch.sigmoid(torch.randn(n, generator=g))
    if sh_degree is None:
        colors = torch.rand(n, 3, generator=g)
    elif sh_sigma is None:
        colors = torch.rand(n, (sh_degree + 1) ** 2, 3, generator=g)
    else:
        colors = torch.randn(n, (sh_degree + 1) ** 2, 3, generator=g) * sh_sigma
    textures = None
    if tex_res is not None:
        tdev = texture_device if texture_device is not None else "cpu"
        tg = torch.Generator(device=tdev).manual_seed(seed + 1000)
        textures = torch.empty(n, tex_res, tex_res, 4, device=tdev, dtype=dtype)
        textures[..., :3].normal_(0.0, 0.1, generator=tg)
        textures[..., 3].uniform_(0.0, 1.0, generator=tg)
    viewmats = torch.stack([look_at_viewmat(p) for p in cam_positions])
    K = intrinsics(width, height, fov_deg)
    Ks = K[None].repeat(len(cam_positions), 1, 1)
    C = len(cam_positions)
    backgrounds = torch.ones(C, 3)
    tg2 = torch.Generator(device="cpu").manual_seed(target_seed)
    target = torch.rand(C, height, width, 3, generator=tg2)
    sc = Scene(means, quats, scales, opacities, colors, None, viewmats, Ks, width, height, sh_degree, backgrounds,
               target).to(device, dtype)
    sc.textures = textures if textures is None else textures.to(device)
    return sc


def config1(**kw) -> Scene:
    """Plain-sized case: N=10,000, 16x16 textures, 256x256, fov 40, camera at (4,0,0), SH degree 0."""
    args = dict(n=10_000, width=256, height=256, fov_deg=40.0, cam_positions=[(4.0, 0.0, 0.0)], tex_res=16,
                mean_dist="uniform", mean_scale=1.0, logscale_mu=-3.5, logscale_sigma=0.5, sh_degree=0)
    args.update(kw)
    return make_scene(**args)


def config2(**kw) -> Scene:
    """Scene scale: N=1,000,000, 50x50 textures, 1297x840, fov 60, camera at radius 5, SH degree 3."""
    args = dict(n=1_000_000, width=1297, height=840, fov_deg=60.0, cam_positions=[(5.0, 0.0, 0.0)], tex_res=50,
                mean_dist="normal", mean_scale=2.0, logscale_mu=-4.5, logscale_sigma=0.7, sh_degree=3,
                sh_sigma=0.2)
    args.update(kw)
    return make_scene(**args)


def config3(**kw) -> Scene:
    """Batched cameras: N=500,000, 32x32 textures, 8 cameras 800x800 (fov 60) on the z=0 ring of radius 4."""
    cams = [(4.0 * math.cos(2 * math.pi * k / 8), 4.0 * math.sin(2 * math.pi * k / 8), 0.0) for k in range(8)]
    args = dict(n=500_000, width=800, height=800, fov_deg=60.0, cam_positions=cams, tex_res=32,
                mean_dist="normal", mean_scale=2.0, logscale_mu=-4.5, logscale_sigma=0.7, sh_degree=3,
                sh_sigma=0.2)
    args.update(kw)
    return make_scene(**args)
