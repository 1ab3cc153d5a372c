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
    # textures drawn on texture_device stay there (tens of GB at the large configs)
    sc.textures = textures if textures is None or texture_device is not None else textures.to(device)
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


def config5(**kw) -> Scene:
    """Full training iteration at scale: N=2,000,000, 1920x1080 (fov 60, camera at radius 5),
    SH degree 3; the 64x64 bfloat16 textures are drawn straight into the optimizer state
    (``train.scene_texture_adam``), so none here."""
    args = dict(n=2_000_000, width=1920, height=1080, fov_deg=60.0, cam_positions=[(5.0, 0.0, 0.0)], tex_res=None,
                mean_dist="normal", mean_scale=2.0, logscale_mu=-4.5, logscale_sigma=0.7, sh_degree=3,
                sh_sigma=0.2)
    args.update(kw)
    return make_scene(**args)


def smooth_textures(n: int, res: int = 16, seed: int = 7, max_cycles: float = 0.75) -> torch.Tensor:
    """Procedural ground-truth textures [n, res, res, 4] in (0, 1): per texture a base
    colour plus one low-frequency cosine wave per channel group (RGB shares a wave,
    alpha has its own), at most ``max_cycles`` cycles across the texture."""
    g = torch.Generator().manual_seed(seed)
    u = (torch.arange(res) + 0.5) / res
    uu, vv = torch.meshgrid(u, u, indexing="xy")  # [res(v), res(u)]
    def wave():
        f = (torch.rand(n, 2, generator=g) * 2 - 1) * max_cycles
        ph = torch.rand(n, generator=g) * 2 * math.pi
        return torch.cos(2 * math.pi * (f[:, 0, None, None] * uu + f[:, 1, None, None] * vv) + ph[:, None, None])
    base = torch.rand(n, 3, generator=g) * 0.6 + 0.2
    amp = torch.rand(n, 3, generator=g) * 0.2
    rgb = base[:, None, None, :] + amp[:, None, None, :] * wave()[..., None]
    alpha = 0.6 + 0.35 * wave()
    return torch.cat([rgb, alpha[..., None]], -1).clamp(0.01, 0.99)


def atom_textures(n: int, res: int = 16, seed: int = 7, n_atoms: int = 24, per_texture: int = 2,
                  cycles=(1.5, 3.0)) -> torch.Tensor:
    """Ground-truth textures [n, res, res, 4] in (0, 1) that share structure, as the
    textures of one scene do: a dictionary of ``n_atoms`` RGBA patterns (each a
    sum of two cosine waves of ``cycles`` cycles across the texture), and per
    texture a base colour/alpha plus ``per_texture`` atoms with random weights."""
    g = torch.Generator().manual_seed(seed)
    u = (torch.arange(res) + 0.5) / res
    uu, vv = torch.meshgrid(u, u, indexing="xy")
    atoms = torch.zeros(n_atoms, res, res, 4)
    for k in range(n_atoms):
        for _ in range(2):
            ang = torch.rand(1, generator=g) * 2 * math.pi
            f = cycles[0] + torch.rand(1, generator=g) * (cycles[1] - cycles[0])
            ph = torch.rand(1, generator=g) * 2 * math.pi
            wave = torch.cos(2 * math.pi * f * (torch.cos(ang) * uu + torch.sin(ang) * vv) + ph)
            atoms[k] += wave[..., None] * (torch.rand(4, generator=g) * 2 - 1) * 0.5
    idx = torch.stack([torch.randperm(n_atoms, generator=g)[:per_texture] for _ in range(n)])
    wts = torch.rand(n, per_texture, generator=g) * 0.3 + 0.1
    base = torch.cat([torch.rand(n, 3, generator=g) * 0.5 + 0.25, torch.rand(n, 1, generator=g) * 0.3 + 0.6], 1)
    tex = base[:, None, None, :] + (atoms[idx] * wts[:, :, None, None, None]).sum(1)
    return tex.clamp(0.01, 0.99)


def sphere_cameras(n_cams: int, radius: float = 4.0, max_elev_deg: float = 50.0, offset: float = 0.0):
    """``n_cams`` camera centres spread over a band of the sphere (golden-angle spiral in
    azimuth, elevations evenly within +-max_elev_deg), all looking at the origin."""
    cams = []
    for k in range(n_cams):
        el = math.radians(max_elev_deg) * (2 * (k + 0.5) / n_cams - 1)
        az = k * math.pi * (3 - math.sqrt(5)) + offset
        cams.append((radius * math.cos(el) * math.cos(az), radius * math.cos(el) * math.sin(az),
                     radius * math.sin(el)))
    return cams


def neural_scene(n: int = 10_000, size: int = 256, n_cams: int = 16, seed: int = 0) -> Scene:
    """Config-1-style Gaussians seen by ``n_cams`` cameras around them (fov 40, radius 4),
    with dim base colours (the textures carry the colour) and ground-truth textures
    that share structure (``atom_textures``): the scene the neural-texture
    decoder's training textures are refined on."""
    cams = sphere_cameras(n_cams)
    sc = make_scene(n, size, size, 40.0, cams, 16, sh_degree=None, seed=seed)
    sc.colors = torch.full_like(sc.colors, 0.05)
    sc.textures = atom_textures(n, 16, seed + 7).to(sc.means.device)
    return sc
