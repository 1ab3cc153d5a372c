"""Depth-to-normal helpers used by the depth render modes.

Same computation as the reference's ``textured_gaussians/utils.py:80-168``
(``depth_to_points`` / ``depth_to_normal``): back-project every pixel centre
with the intrinsics, take central differences of the points, and normalise
their cross product; the one-pixel border is zero.
"""
from typing import Literal

import torch
import torch.nn.functional as F
from torch import Tensor

Filtering = Literal["bilinear", "bilinear_bwd2", "bilinear_exact"]


def depth_to_points(depths: Tensor, camtoworlds: Tensor, Ks: Tensor, z_depth: bool = True) -> Tensor:
    """depths [..., H, W, 1], camtoworlds [..., 4, 4], Ks [..., 3, 3] -> points [..., H, W, 3]."""
    assert depths.shape[-1] == 1, f"Invalid depth shape: {depths.shape}"
    assert camtoworlds.shape[-2:] == (4, 4), f"Invalid viewmats shape: {camtoworlds.shape}"
    assert Ks.shape[-2:] == (3, 3), f"Invalid Ks shape: {Ks.shape}"
    device = depths.device
    height, width = depths.shape[-3:-1]
    x, y = torch.meshgrid(torch.arange(width, device=device), torch.arange(height, device=device), indexing="xy")
    fx, fy = Ks[..., 0, 0], Ks[..., 1, 1]
    cx, cy = Ks[..., 0, 2], Ks[..., 1, 2]
    cam_dirs = F.pad(torch.stack([(x - cx[..., None, None] + 0.5) / fx[..., None, None],
                                  (y - cy[..., None, None] + 0.5) / fy[..., None, None]], dim=-1),
                     (0, 1), value=1.0)
    R = camtoworlds[..., :3, :3]
    dirs = (R[..., None, None, :, :] @ cam_dirs[..., None]).squeeze(-1)
    origins = camtoworlds[..., :3, -1]
    if not z_depth:
        dirs = F.normalize(dirs, dim=-1)
    return origins[..., None, None, :] + depths * dirs


def depth_to_normal(depths: Tensor, camtoworlds: Tensor, Ks: Tensor, z_depth: bool = True) -> Tensor:
    """depths [..., H, W, 1] -> world-frame surface normals [..., H, W, 3]."""
    points = depth_to_points(depths, camtoworlds, Ks, z_depth=z_depth)
    dx = points[..., 2:, 1:-1, :] - points[..., :-2, 1:-1, :]
    dy = points[..., 1:-1, 2:, :] - points[..., 1:-1, :-2, :]
    normals = F.normalize(torch.cross(dx, dy, dim=-1), dim=-1)
    return F.pad(normals, (0, 0, 1, 1, 1, 1), value=0.0)
