"""Rendering entry points: ``rasterization_textured_gaussians`` and the plain
``rasterization_2dgs`` comparator.

Same signatures, defaults, assertions and 9-tuple returns as the reference's
``textured_gaussians/rendering.py:1342-1685`` and ``:1009-1338``; the stages
are the MI355X ops of ``textured_gaussians_amd.cuda``.  Additions:
``proj_grad="exact"`` (true projection VJP) and ``filtering="bilinear_exact"``
(true texture-coordinate gradient); ``grad_mode="exact"`` sets both.
Differences: the ``densify`` dummy lives on the inputs' device (the reference
hard-codes ``device="cuda"``, ``rendering.py:1541-1543``); with ``packed=True``
each entry samples the texture of its own Gaussian (the reference uses the
entry index, Q3).
"""
import math
from typing import Dict, Literal, Optional, Tuple

import torch
from torch import Tensor

from .cuda._wrapper import (
    fully_fused_projection_2dgs,
    isect_offset_encode,
    isect_tiles,
    rasterize_to_pixels_2dgs,
    rasterize_to_pixels_textured_gaussians,
    spherical_harmonics,
    spherical_harmonics_views,
)
from .cuda._ops import wait_event
from .utils import depth_to_normal


def _common_checks(means, quats, scales, opacities, colors, viewmats, Ks, render_mode, distloss, sh_degree):
    N = means.shape[0]
    C = viewmats.shape[0]
    assert means.shape == (N, 3), means.shape
    assert quats.shape == (N, 4), quats.shape
    assert scales.shape == (N, 3), scales.shape
    assert opacities.shape == (N,), opacities.shape
    assert viewmats.shape == (C, 4, 4), viewmats.shape
    assert Ks.shape == (C, 3, 3), Ks.shape
    assert render_mode in ["RGB", "D", "ED", "RGB+D", "RGB+ED"], render_mode
    if distloss:
        assert render_mode in ["D", "ED", "RGB+D", "RGB+ED"], (
            f"distloss requires depth rendering, render_mode should be D, ED, RGB+D, RGB+ED, but got {render_mode}")
    if sh_degree is None:
        assert (colors.dim() == 2 and colors.shape[0] == N) or (
            colors.dim() == 3 and colors.shape[:2] == (C, N)), colors.shape
    else:
        assert colors.dim() == 3 and colors.shape[0] == N and colors.shape[2] == 3, colors.shape
        assert (sh_degree + 1) ** 2 <= colors.shape[1], colors.shape
    return N, C


def _inverse_deferred(viewmats: Tensor):
    """``torch.inverse(viewmats)`` by the same routine, without the host wait of its
    singularity check: the check's status comes back by an asynchronous copy and
    ``_check_inverse`` reads it after the intersection count read, where the host
    waits for the device anyway (so the check costs no second wait).  The
    reference inverts twice per render ([RD]:1573, and for the normals), each
    time with that wait."""
    if not viewmats.is_cuda:
        return torch.inverse(viewmats), None
    inv, info = torch.linalg.inv_ex(viewmats)
    info_host = torch.empty(info.shape, dtype=info.dtype, pin_memory=True)
    info_host.copy_(info, non_blocking=True)
    done = torch.cuda.Event()
    done.record()
    return inv, (info_host, done)


def _check_inverse(pending) -> None:
    if pending is None:
        return
    info_host, done = pending
    wait_event(done)
    if bool((info_host != 0).any()):
        raise torch.linalg.LinAlgError(
            "linalg.inv: a view matrix is singular, the inversion could not be completed")


def _prepare(means, quats, scales, opacities, colors, viewmats, Ks, width, height, near_plane, far_plane,
             radius_clip, eps2d, sh_degree, packed, tile_size, sparse_grad, norm_rot_grad, render_mode,
             proj_grad):
    """Projection, tile intersection, per-view colours ([RD]:1507-1594).  Also
    returns the camera-to-world matrices (the inverse view matrices)."""
    N, C = means.shape[0], viewmats.shape[0]
    camtoworlds, pending = _inverse_deferred(viewmats)
    proj = fully_fused_projection_2dgs(
        means, quats, scales, viewmats, Ks, width, height, eps2d, near_plane, far_plane, radius_clip, packed,
        sparse_grad, norm_rot_grad, proj_grad=proj_grad)
    if packed:
        camera_ids, gaussian_ids, radii, means2d, depths, ray_transforms, normals = proj
        # per-entry gathers by index_select (the reference indexes, opacities[gaussian_ids]):
        # the same values, and a backward that scatter-adds with atomics instead of
        # sorting the indices first (torch's advanced-indexing backward)
        opacities = opacities.index_select(0, gaussian_ids)
    else:
        radii, means2d, depths, ray_transforms, normals = proj
        # one camera: a view (repeat's backward would be a reduction over the camera axis)
        opacities = opacities[None] if C == 1 else opacities.repeat(C, 1)
        camera_ids, gaussian_ids = None, None
    densify = torch.zeros_like(means2d, dtype=means.dtype, requires_grad=True, device=means.device)
    # per-view colours before the tile intersection: they need only the projection, and the
    # intersection count read makes the host wait for the device, so work queued before it
    # overlaps that wait (the reference computes them after the sort, [RD]:1561-1594)
    if sh_degree is not None and not packed:
        # grid layout: one coefficient set seen from C views (the reference expands the
        # coefficients to [C, N, K, 3] first; the views kernel loops over them instead)
        if camtoworlds.shape[0] == 1:
            # one camera: the gradient to the means is the directions' gradient itself (a
            # broadcast over a camera axis would cost a reduction kernel in the backward)
            dirs = (means - camtoworlds[0, :3, 3])[None]
        else:
            dirs = means[None, :, :] - camtoworlds[:, None, :3, 3]
        colors = spherical_harmonics_views(sh_degree, dirs, colors, masks=radii > 0)
        colors = torch.clamp_min(colors + 0.5, 0.0)
    else:
        if not (colors.dim() == 3 and sh_degree is None):
            colors = colors.index_select(0, gaussian_ids) if packed else colors.expand(C, *([-1] * colors.dim()))
        elif packed:
            colors = colors.reshape(C * N, -1).index_select(0, camera_ids * N + gaussian_ids)
        if sh_degree is not None:
            dirs = means.index_select(0, gaussian_ids) - camtoworlds[camera_ids, :3, 3]
            colors = spherical_harmonics(sh_degree, dirs, colors, masks=radii > 0)
            colors = torch.clamp_min(colors + 0.5, 0.0)
    if render_mode in ["RGB+D", "RGB+ED"]:
        colors = torch.cat((colors, depths[..., None]), dim=-1)
    elif render_mode in ["D", "ED"]:
        colors = depths[..., None]
    tile_width = math.ceil(width / float(tile_size))
    tile_height = math.ceil(height / float(tile_size))
    tiles_per_gauss, isect_ids, flatten_ids = isect_tiles(
        means2d, radii, depths, tile_size, tile_width, tile_height, packed=packed, n_cameras=C,
        camera_ids=camera_ids, gaussian_ids=gaussian_ids)
    isect_offsets = isect_offset_encode(isect_ids, C, tile_width, tile_height)
    _check_inverse(pending)
    meta = {
        "camera_ids": camera_ids, "gaussian_ids": gaussian_ids, "radii": radii, "means2d": means2d,
        "depths": depths, "ray_transforms": ray_transforms, "opacities": opacities, "normals": normals,
        "tile_width": tile_width, "tile_height": tile_height, "tiles_per_gauss": tiles_per_gauss,
        "isect_ids": isect_ids, "flatten_ids": flatten_ids, "isect_offsets": isect_offsets, "width": width,
        "height": height, "tile_size": tile_size, "n_cameras": C, "gradient_2dgs": densify,
    }
    return colors, opacities, densify, meta, camtoworlds


def _finish(render_colors, render_alphas, render_normals, render_median, render_mode, depth_mode, camtoworlds, Ks):
    render_normals_from_depth = None
    if render_mode in ["ED", "RGB+ED"]:
        render_colors = torch.cat([render_colors[..., :-1],
                                   render_colors[..., -1:] / render_alphas.clamp(min=1e-10)], dim=-1)
    if render_mode in ["RGB+ED", "RGB+D"]:
        depth_for_normal = render_colors[..., -1:] if depth_mode == "expected" else render_median
        render_normals_from_depth = depth_to_normal(depth_for_normal, camtoworlds, Ks).squeeze(0)
    # camera-frame normals -> world frame, n_w[i] = sum_j R[i, j] n[j]: a broadcast product and a
    # sum over j (two elementwise passes; a [H*W,3] x [3,3] GEMM runs far below bandwidth)
    R = camtoworlds[..., :3, :3]
    render_normals = (render_normals[..., None, :] * R[:, None, None]).sum(-1)
    return render_colors, render_normals, render_normals_from_depth


def rasterization_textured_gaussians(
        means: Tensor, quats: Tensor, scales: Tensor, opacities: Tensor, colors: Tensor, textures: Tensor,
        viewmats: Tensor, Ks: Tensor, width: int, height: int, texture_range_x: float = 3.0,
        texture_range_y: float = 3.0, near_plane: float = 0.01, far_plane: float = 1e10,
        radius_clip: float = 0.0, eps2d: float = 0.3, sh_degree: Optional[int] = None, packed: bool = False,
        tile_size: int = 16, backgrounds: Optional[Tensor] = None,
        render_mode: Literal["RGB", "D", "ED", "RGB+D", "RGB+ED"] = "RGB", sparse_grad: bool = False,
        absgrad: bool = False, distloss: bool = False, norm_rot_grad: bool = False,
        depth_mode: Literal["expected", "median"] = "expected", gs_contrib_threshold: float = 0.0,
        filtering: str = "bilinear", g_weight: float = 1.0, grad_mode: Optional[str] = None,
        proj_grad: str = "reference", texture_adam=None
) -> Tuple[Tensor, Tensor, Tensor, Tensor, Tensor, Tensor, Tensor, Tensor, Dict]:
    """Rasterize N textured 2D Gaussians to C image planes.

    means [N,3], quats [N,4] (wxyz, unnormalised), scales [N,3] (two used),
    opacities [N], colors [(C,)N,D] or SH coefficients [N,K,3] with
    ``sh_degree``, textures [N,th,tw,4] (RGB added to the colour, A scales
    alpha), viewmats [C,4,4] world-to-camera, Ks [C,3,3].

    Returns (render_colors [C,H,W,X], render_alphas [C,H,W,1], render_normals
    [C,H,W,3] (world frame), render_normals_from_depth or None, render_distort
    [C,H,W,1], render_median [C,H,W,1], gs_contrib_sum [N], gs_contrib_count
    [N], meta).  ``grad_mode="exact"`` selects true derivatives for the
    projection and the texture lookup.  ``textures`` may be bfloat16 (float32
    inputs otherwise).  ``texture_adam``: a ``BF16TextureAdam`` holding these
    textures; the backward steps it instead of producing a texture gradient.
    """
    N, C = _common_checks(means, quats, scales, opacities, colors, viewmats, Ks, render_mode, distloss, sh_degree)
    assert textures.shape[0] == N and textures.shape[-1] == 4 and textures.dim() == 4, textures.shape
    if grad_mode is not None:
        assert grad_mode in ("reference", "exact"), grad_mode
        if grad_mode == "exact":
            proj_grad = "exact"
            filtering = "bilinear_exact"
    colors, opac_c, densify, meta, camtoworlds = _prepare(
        means, quats, scales, opacities, colors, viewmats, Ks, width, height, near_plane, far_plane, radius_clip,
        eps2d, sh_degree, packed, tile_size, sparse_grad, norm_rot_grad, render_mode, proj_grad)
    (render_colors, render_alphas, render_normals, render_distort, render_median, gs_contrib_sum,
     gs_contrib_count) = rasterize_to_pixels_textured_gaussians(
        meta["means2d"], meta["ray_transforms"], colors, opac_c, textures, texture_range_x, texture_range_y,
        meta["normals"], densify, width, height, tile_size, meta["isect_offsets"], meta["flatten_ids"],
        backgrounds=backgrounds, packed=packed, absgrad=absgrad, distloss=distloss,
        gs_contrib_threshold=gs_contrib_threshold, filtering=filtering, g_weight=g_weight,
        gaussian_ids=meta["gaussian_ids"], texture_adam=texture_adam)
    render_colors, render_normals, render_normals_from_depth = _finish(
        render_colors, render_alphas, render_normals, render_median, render_mode, depth_mode, camtoworlds, Ks)
    meta["render_distort"] = render_distort
    return (render_colors, render_alphas, render_normals, render_normals_from_depth, render_distort,
            render_median, gs_contrib_sum, gs_contrib_count, meta)


def rasterization_2dgs(
        means: Tensor, quats: Tensor, scales: Tensor, opacities: Tensor, colors: Tensor, viewmats: Tensor,
        Ks: Tensor, width: int, height: int, near_plane: float = 0.01, far_plane: float = 1e10,
        radius_clip: float = 0.0, eps2d: float = 0.3, sh_degree: Optional[int] = None, packed: bool = False,
        tile_size: int = 16, backgrounds: Optional[Tensor] = None,
        render_mode: Literal["RGB", "D", "ED", "RGB+D", "RGB+ED"] = "RGB", sparse_grad: bool = False,
        absgrad: bool = False, distloss: bool = False, norm_rot_grad: bool = False,
        depth_mode: Literal["expected", "median"] = "expected", gs_contrib_threshold: float = 0.0,
        grad_mode: Optional[str] = None, proj_grad: str = "reference"
) -> Tuple[Tensor, Tensor, Tensor, Tensor, Tensor, Tensor, Tensor, Tensor, Dict]:
    """Rasterize N plain 2D Gaussians to C image planes (the reference's
    ``rasterization_2dgs``); returns the same 9-tuple as the textured entry."""
    raster_grad = "reference"
    if grad_mode == "exact":
        proj_grad = raster_grad = "exact"
    N, C = _common_checks(means, quats, scales, opacities, colors, viewmats, Ks, render_mode, distloss, sh_degree)
    colors, opac_c, densify, meta, camtoworlds = _prepare(
        means, quats, scales, opacities, colors, viewmats, Ks, width, height, near_plane, far_plane, radius_clip,
        eps2d, sh_degree, packed, tile_size, sparse_grad, norm_rot_grad, render_mode, proj_grad)
    (render_colors, render_alphas, render_normals, render_distort, render_median, gs_contrib_sum,
     gs_contrib_count) = rasterize_to_pixels_2dgs(
        meta["means2d"], meta["ray_transforms"], colors, opac_c, meta["normals"], densify, width, height,
        tile_size, meta["isect_offsets"], meta["flatten_ids"], backgrounds=backgrounds, packed=packed,
        absgrad=absgrad, distloss=distloss, gs_contrib_threshold=gs_contrib_threshold, grad_mode=raster_grad,
        gaussian_ids=meta["gaussian_ids"], n_gaussians=N)
    render_colors, render_normals, render_normals_from_depth = _finish(
        render_colors, render_alphas, render_normals, render_median, render_mode, depth_mode, camtoworlds, Ks)
    meta["render_distort"] = render_distort
    return (render_colors, render_alphas, render_normals, render_normals_from_depth, render_distort,
            render_median, gs_contrib_sum, gs_contrib_count, meta)
