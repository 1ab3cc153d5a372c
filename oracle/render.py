"""Full textured-2DGS render pipeline on the CPU (test infrastructure only).

Restates ``textured_gaussians/rendering.py:1342-1685`` (``rasterization_textured_gaussians``)
and ``:1009-1338`` (``rasterization_2dgs``) for the unpacked layout with the
oracle's stages: projection (autograd), tile intersection and sort (numpy,
bit-exact, no gradient), SH colours (autograd) and the rasterizer (autograd).
"""
import math

import numpy as np
import torch

from . import isect as O_isect
from .projection import project_2dgs
from .raster import rasterize_to_pixels
from .sh import spherical_harmonics


def isect_torch(means2d, radii, depths, tile_size, tile_width, tile_height):
    """Tile intersection of torch tensors via the numpy restatement."""
    tpg, ids, flat = O_isect.isect_tiles(means2d.detach().cpu().float().numpy(),
                                         radii.detach().cpu().numpy(),
                                         depths.detach().cpu().float().numpy(), tile_size, tile_width,
                                         tile_height)
    C = radii.shape[0]
    offs = O_isect.isect_offset_encode(ids, C, tile_width, tile_height)
    return (torch.from_numpy(tpg), torch.from_numpy(ids), torch.from_numpy(flat), torch.from_numpy(offs))


def render(means, quats, scales, opacities, colors, viewmats, Ks, width, height, textures=None,
           texture_range=(3.0, 3.0), near_plane=0.01, far_plane=1e10, radius_clip=0.0, sh_degree=None,
           tile_size=16, backgrounds=None, render_mode="RGB", g_weight=1.0, uv_grad="exact",
           gs_contrib_threshold=0.0, isect=None):
    """Returns a dict with the rasterizer outputs plus ``meta`` (projection and
    intersection results).  ``isect`` may supply (flatten_ids, isect_offsets)
    computed elsewhere (e.g. from the GPU's projection) to share the lists."""
    C = viewmats.shape[0]
    radii, means2d, depths, rt, normals = project_2dgs(means, quats, scales, viewmats, Ks, width, height,
                                                       near_plane, far_plane, radius_clip)
    tw = math.ceil(width / tile_size)
    th = math.ceil(height / tile_size)
    if isect is None:
        tpg, isect_ids, flatten_ids, offsets = isect_torch(means2d, radii, depths, tile_size, tw, th)
    else:
        flatten_ids, offsets = isect
        tpg = isect_ids = None
    if not (colors.dim() == 3 and sh_degree is None):
        colors = colors.expand(C, *([-1] * colors.dim()))
    if sh_degree is not None:
        camtoworlds = torch.inverse(viewmats)
        dirs = means[None, :, :] - camtoworlds[:, None, :3, 3]
        colors = spherical_harmonics(sh_degree, dirs, colors, masks=radii > 0)
        colors = torch.clamp_min(colors + 0.5, 0.0)
    if render_mode in ("RGB+D", "RGB+ED"):
        colors = torch.cat((colors, depths[..., None]), dim=-1)
    elif render_mode in ("D", "ED"):
        colors = depths[..., None]
    opac = opacities.repeat(C, 1)
    out = rasterize_to_pixels(means2d, rt, colors, opac, normals, width, height, tile_size, offsets,
                              flatten_ids, textures=textures, texture_range=texture_range,
                              backgrounds=backgrounds, g_weight=g_weight,
                              gs_contrib_threshold=gs_contrib_threshold, uv_grad=uv_grad)
    rc = out["render_colors"]
    if render_mode in ("ED", "RGB+ED"):
        rc = torch.cat([rc[..., :-1], rc[..., -1:] / out["render_alphas"].clamp(min=1e-10)], dim=-1)
    R = torch.linalg.inv(viewmats)[..., :3, :3]
    out["render_colors"] = rc
    out["render_normals"] = (R[:, None, None] @ out["render_normals"][..., None]).squeeze(-1)
    out["meta"] = dict(radii=radii, means2d=means2d, depths=depths, ray_transforms=rt, normals=normals,
                       tiles_per_gauss=tpg, isect_ids=isect_ids, flatten_ids=flatten_ids, isect_offsets=offsets,
                       colors=colors, opacities=opac)
    return out
