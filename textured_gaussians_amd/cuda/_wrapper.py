"""Python op wrappers and autograd Functions of the textured-2DGS path.

Mirrors the reference's ``textured_gaussians/cuda/_wrapper.py`` for the ops on
this path: same function names, argument names, defaults, returns and
autograd behaviour (``spherical_harmonics`` :54-80 / :1233-1263,
``fully_fused_projection_2dgs`` :1267-1484, ``isect_tiles`` :349-420,
``isect_offset_encode`` :423-440, ``rasterize_to_pixels_2dgs`` :1647-1784 /
:2767-2944, ``rasterize_to_pixels_textured_gaussians`` :1806-2002 /
:2947-3155, ``selective_adam_update`` :26-41).  Native ops are resolved by name
from the op table, like the reference's ``_make_lazy_cuda_func``.

Differences from the reference (each a fix or an addition):
* ``filtering="bilinear_exact"`` and ``proj_grad="exact"`` select true
  derivatives (the reference's rules are kept as the defaults).
* ``v_backgrounds`` is returned when ``backgrounds`` requires grad (the
  reference tests ``needs_input_grad[7]``, a float argument, so never does).
* No ``torch.cuda.synchronize()`` after the backward kernels.
"""
from typing import Callable, Optional, Tuple

import torch
from torch import Tensor

from ._backend import lib
from ._ops import _C


def _make_lazy_cuda_func(name: str) -> Callable:
    def call(*args, **kwargs):
        return getattr(_C, name)(*args, **kwargs)

    return call


def selective_adam_update(param: Tensor, param_grad: Tensor, exp_avg: Tensor, exp_avg_sq: Tensor,
                          tiles_touched: Tensor, lr: float, b1: float, b2: float, eps: float, N: int,
                          M: int) -> None:
    _make_lazy_cuda_func("selective_adam_update")(param, param_grad, exp_avg, exp_avg_sq, tiles_touched, lr, b1,
                                                  b2, eps, N, M)


# ----------------------------------------------------------------------- SH
def spherical_harmonics(degrees_to_use: int, dirs: Tensor, coeffs: Tensor,
                        masks: Optional[Tensor] = None) -> Tensor:
    """Computes spherical harmonics.  dirs [..., 3], coeffs [..., K, 3], masks
    [...] -> colors [..., 3]."""
    assert (degrees_to_use + 1) ** 2 <= coeffs.shape[-2], coeffs.shape
    assert dirs.shape[:-1] == coeffs.shape[:-2], (dirs.shape, coeffs.shape)
    assert dirs.shape[-1] == 3, dirs.shape
    assert coeffs.shape[-1] == 3, coeffs.shape
    if masks is not None:
        assert masks.shape == dirs.shape[:-1], masks.shape
        masks = masks.contiguous()
    return _SphericalHarmonics.apply(degrees_to_use, dirs.contiguous(), coeffs.contiguous(), masks)


def spherical_harmonics_views(degrees_to_use: int, dirs: Tensor, coeffs: Tensor,
                              masks: Optional[Tensor] = None) -> Tensor:
    """Colours of one coefficient set coeffs [N, K, 3] seen from C views, dirs
    [C, N, 3] (masks [C, N]) -> [C, N, 3]: what the reference computes as
    ``spherical_harmonics(d, dirs, coeffs.expand(C, ...), masks)``
    (rendering.py:1573-1583).  float32 with K <= 16 runs one kernel that loops over
    the views (no [C, N, K, 3] copy, gradients summed over the views in the kernel);
    anything else takes that expanded form."""
    C = dirs.shape[0]
    if C > 1 and dirs.is_cuda and dirs.dtype == torch.float32 and _C.sh_cams_supported(coeffs):
        assert (degrees_to_use + 1) ** 2 <= coeffs.shape[-2], coeffs.shape
        assert dirs.shape == (C, coeffs.shape[0], 3), (dirs.shape, coeffs.shape)
        if masks is not None:
            assert masks.shape == dirs.shape[:-1], masks.shape
            masks = masks.contiguous()
        return _SphericalHarmonicsViews.apply(degrees_to_use, dirs.contiguous(), coeffs.contiguous(), masks)
    # one view: a leading axis by view (an expand's backward is a reduction over it)
    views = coeffs[None] if C == 1 else coeffs.expand(C, *([-1] * coeffs.dim()))
    return spherical_harmonics(degrees_to_use, dirs, views, masks)


class _SphericalHarmonicsViews(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sh_degree: int, dirs: Tensor, coeffs: Tensor, masks: Tensor) -> Tensor:
        colors = _C.compute_sh_fwd_cams(sh_degree, dirs, coeffs, masks)
        ctx.save_for_backward(dirs, coeffs, masks)
        ctx.sh_degree = sh_degree
        return colors

    @staticmethod
    def backward(ctx, v_colors: Tensor):
        dirs, coeffs, masks = ctx.saved_tensors
        v_coeffs, v_dirs = _C.compute_sh_bwd_cams(ctx.sh_degree, dirs, coeffs, masks, v_colors.contiguous(),
                                                    ctx.needs_input_grad[1])
        if not ctx.needs_input_grad[2]:
            v_coeffs = None
        return None, v_dirs, v_coeffs, None


class _SphericalHarmonics(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sh_degree: int, dirs: Tensor, coeffs: Tensor, masks: Tensor) -> Tensor:
        colors = _make_lazy_cuda_func("compute_sh_fwd")(sh_degree, dirs, coeffs, masks)
        ctx.save_for_backward(dirs, coeffs, masks)
        ctx.sh_degree = sh_degree
        ctx.num_bases = coeffs.shape[-2]
        return colors

    @staticmethod
    def backward(ctx, v_colors: Tensor):
        dirs, coeffs, masks = ctx.saved_tensors
        compute_v_dirs = ctx.needs_input_grad[1]
        v_coeffs, v_dirs = _make_lazy_cuda_func("compute_sh_bwd")(
            ctx.num_bases, ctx.sh_degree, dirs, coeffs, masks, v_colors.contiguous(), compute_v_dirs)
        if not compute_v_dirs:
            v_dirs = None
        return None, v_dirs, v_coeffs, None


# --------------------------------------------------------------- projection
def fully_fused_projection_2dgs(means: Tensor, quats: Tensor, scales: Tensor, viewmats: Tensor, Ks: Tensor,
                                width: int, height: int, eps2d: float = 0.3, near_plane: float = 0.01,
                                far_plane: float = 1e10, radius_clip: float = 0.0, packed: bool = False,
                                sparse_grad: bool = False, norm_rot_grad: bool = False,
                                proj_grad: str = "reference") -> Tuple[Tensor, Tensor, Tensor, Tensor, Tensor]:
    """Prepare 2D Gaussians for rasterization: returns (radii [C,N] int32,
    means2d [C,N,2], depths [C,N], ray_transforms [C,N,3,3], normals [C,N,3]).

    ``proj_grad="exact"`` makes the backward the true derivative (the
    reference's VJP drops part of the means2d term, ``utils.cuh:684-703``)."""
    C = viewmats.size(0)
    N = means.size(0)
    assert means.size() == (N, 3), means.size()
    assert viewmats.size() == (C, 4, 4), viewmats.size()
    assert Ks.size() == (C, 3, 3), Ks.size()
    assert quats is not None, "quats is required"
    assert scales is not None, "scales is required"
    assert quats.size() == (N, 4), quats.size()
    assert scales.size() == (N, 3), scales.size()
    assert proj_grad in ("reference", "exact"), proj_grad
    if sparse_grad:
        assert packed, "sparse_grad is only supported when packed is True"
    if packed:
        return _FullyFusedProjectionPacked2DGS.apply(means.contiguous(), quats.contiguous(), scales.contiguous(),
                                                     viewmats.contiguous(), Ks.contiguous(), width, height,
                                                     near_plane, far_plane, radius_clip, sparse_grad,
                                                     norm_rot_grad, proj_grad)
    return _FullyFusedProjection2DGS.apply(means.contiguous(), quats.contiguous(), scales.contiguous(),
                                           viewmats.contiguous(), Ks.contiguous(), width, height, eps2d,
                                           near_plane, far_plane, radius_clip, norm_rot_grad, proj_grad)


class _FullyFusedProjection2DGS(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means, quats, scales, viewmats, Ks, width, height, eps2d, near_plane, far_plane,
                radius_clip, norm_rot_grad, proj_grad):
        radii, means2d, depths, ray_transforms, normals = _make_lazy_cuda_func("fully_fused_projection_fwd_2dgs")(
            means, quats, scales, viewmats, Ks, width, height, eps2d, near_plane, far_plane, radius_clip)
        ctx.save_for_backward(means, quats, scales, viewmats, Ks, radii, ray_transforms, normals)
        ctx.width, ctx.height = width, height
        ctx.norm_rot_grad = norm_rot_grad
        ctx.proj_grad = proj_grad
        ctx.mark_non_differentiable(radii)
        return radii, means2d, depths, ray_transforms, normals

    @staticmethod
    def backward(ctx, v_radii, v_means2d, v_depths, v_ray_transforms, v_normals):
        means, quats, scales, viewmats, Ks, radii, ray_transforms, normals = ctx.saved_tensors
        if ctx.proj_grad == "exact":
            name = "fully_fused_projection_bwd_exact_2dgs"
        elif ctx.norm_rot_grad:
            name = "fully_fused_projection_bwd2_2dgs"
        else:
            name = "fully_fused_projection_bwd_2dgs"
        v_means, v_quats, v_scales, v_viewmats = _make_lazy_cuda_func(name)(
            means, quats, scales, viewmats, Ks, ctx.width, ctx.height, radii, ray_transforms,
            v_means2d.contiguous(), v_depths.contiguous(), v_normals.contiguous(), v_ray_transforms.contiguous(),
            ctx.needs_input_grad[3])
        g = ctx.needs_input_grad
        return (v_means if g[0] else None, v_quats if g[1] else None, v_scales if g[2] else None,
                v_viewmats if g[3] else None, None, None, None, None, None, None, None, None, None)


class _FullyFusedProjectionPacked2DGS(torch.autograd.Function):
    """Packed layout ([W]:1487-1644): returns (camera_ids, gaussian_ids, radii, means2d,
    depths, ray_transforms, normals) of the visible entries."""

    @staticmethod
    def forward(ctx, means, quats, scales, viewmats, Ks, width, height, near_plane, far_plane, radius_clip,
                sparse_grad, norm_rot_grad, proj_grad):
        (indptr, camera_ids, gaussian_ids, radii, means2d, depths, ray_transforms,
         normals) = _make_lazy_cuda_func("fully_fused_projection_packed_fwd_2dgs")(
            means, quats, scales, viewmats, Ks, width, height, near_plane, far_plane, radius_clip)
        ctx.save_for_backward(camera_ids, gaussian_ids, means, quats, scales, viewmats, Ks, ray_transforms)
        ctx.width, ctx.height = width, height
        ctx.sparse_grad, ctx.norm_rot_grad, ctx.proj_grad = sparse_grad, norm_rot_grad, proj_grad
        ctx.mark_non_differentiable(camera_ids, gaussian_ids, radii)
        return camera_ids, gaussian_ids, radii, means2d, depths, ray_transforms, normals

    @staticmethod
    def backward(ctx, v_camera_ids, v_gaussian_ids, v_radii, v_means2d, v_depths, v_ray_transforms, v_normals):
        camera_ids, gaussian_ids, means, quats, scales, viewmats, Ks, ray_transforms = ctx.saved_tensors
        if ctx.proj_grad == "exact":
            name = "fully_fused_projection_packed_bwd_exact_2dgs"
        elif ctx.norm_rot_grad:
            name = "fully_fused_projection_packed_bwd2_2dgs"
        else:
            name = "fully_fused_projection_packed_bwd_2dgs"
        z = lambda t, shape: t.contiguous() if t is not None else torch.zeros(shape, device=means.device,  # noqa: E731
                                                                              dtype=means.dtype)
        nnz = camera_ids.numel()
        v_means, v_quats, v_scales, v_viewmats = _make_lazy_cuda_func(name)(
            means, quats, scales, viewmats, Ks, ctx.width, ctx.height, camera_ids, gaussian_ids, ray_transforms,
            z(v_means2d, (nnz, 2)), z(v_depths, (nnz,)), z(v_normals, (nnz, 3)), z(v_ray_transforms, (nnz, 3, 3)),
            ctx.needs_input_grad[3], ctx.sparse_grad)
        if ctx.sparse_grad:
            # as the reference: sparse COO tensors over the entries' Gaussians
            def sp(v, like):
                return torch.sparse_coo_tensor(gaussian_ids[None], v, like.size(), is_coalesced=len(viewmats) == 1)
            v_means, v_quats, v_scales = sp(v_means, means), sp(v_quats, quats), sp(v_scales, scales)
        g = ctx.needs_input_grad
        return (v_means if g[0] else None, v_quats if g[1] else None, v_scales if g[2] else None,
                v_viewmats if g[3] else None, None, None, None, None, None, None, None, None, None)


# -------------------------------------------------------------------- isect
@torch.no_grad()
def isect_tiles(means2d: Tensor, radii: Tensor, depths: Tensor, tile_size: int, tile_width: int,
                tile_height: int, sort: bool = True, packed: bool = False, n_cameras: Optional[int] = None,
                camera_ids: Optional[Tensor] = None, gaussian_ids: Optional[Tensor] = None
                ) -> Tuple[Tensor, Tensor, Tensor]:
    """Maps projected Gaussians to intersecting tiles.  Returns
    (tiles_per_gauss int32 [C,N], isect_ids int64 [I] = cam | tile | depth bits,
    flatten_ids int32 [I]), sorted by isect_ids (stable) when ``sort``."""
    if packed:
        nnz = means2d.size(0)
        assert means2d.shape == (nnz, 2), means2d.size()
        assert camera_ids is not None, "camera_ids is required if packed is True"
        assert gaussian_ids is not None, "gaussian_ids is required if packed is True"
        assert n_cameras is not None, "n_cameras is required if packed is True"
        C = n_cameras
    else:
        C, N, _ = means2d.shape
        assert means2d.shape == (C, N, 2), means2d.size()
        assert radii.shape == (C, N), radii.size()
        assert depths.shape == (C, N), depths.size()
    return _make_lazy_cuda_func("isect_tiles")(means2d.contiguous(), radii.contiguous(), depths.contiguous(),
                                               camera_ids, gaussian_ids, C, tile_size, tile_width, tile_height,
                                               sort, True)


@torch.no_grad()
def isect_offset_encode(isect_ids: Tensor, n_cameras: int, tile_width: int, tile_height: int) -> Tensor:
    """Encodes intersection ids to per-tile offsets [C, tile_height, tile_width]."""
    return _make_lazy_cuda_func("isect_offset_encode")(isect_ids.contiguous(), n_cameras, tile_width, tile_height)


# --------------------------------------------------------------- rasterizer
def _pad_channels(colors: Tensor, backgrounds: Optional[Tensor]):
    """Pad to the nearest supported channel count ([W]:1880-1902)."""
    channels = colors.shape[-1]
    if channels > 512 or channels == 0:
        raise ValueError(f"Unsupported number of color channels: {channels}")
    if channels not in (1, 2, 3, 4, 8, 16, 32, 64, 128, 256, 512):
        padded = (1 << (channels - 1).bit_length()) - channels
        colors = torch.cat([colors, torch.zeros(*colors.shape[:-1], padded, device=colors.device,
                                                dtype=colors.dtype)], dim=-1)
        if backgrounds is not None:
            backgrounds = torch.cat([backgrounds, torch.zeros(*backgrounds.shape[:-1], padded,
                                                              device=backgrounds.device,
                                                              dtype=backgrounds.dtype)], dim=-1)
    else:
        padded = 0
    return colors, backgrounds, padded


PRE = "rasterize_to_pixels_"
TGSPOST = "_textured_gaussians"
tgs_fns = {
    "bilinear": (f"{PRE}fwd{TGSPOST}", f"{PRE}bwd{TGSPOST}"),
    "bilinear_bwd2": (f"{PRE}fwd{TGSPOST}", f"{PRE}bwd2{TGSPOST}"),
    "bilinear_exact": (f"{PRE}fwd{TGSPOST}", f"{PRE}bwd_exact{TGSPOST}"),
}


def rasterize_to_pixels_textured_gaussians(
        means2d: Tensor, ray_transforms: Tensor, colors: Tensor, opacities: Tensor, textures: Tensor,
        texture_range_x: float, texture_range_y: float, normals: Tensor, densify: Tensor, image_width: int,
        image_height: int, tile_size: int, isect_offsets: Tensor, flatten_ids: Tensor,
        backgrounds: Optional[Tensor] = None, masks: Optional[Tensor] = None, packed: bool = False,
        absgrad: bool = False, distloss: bool = False, gs_contrib_threshold: float = 0.0,
        filtering: str = "bilinear", g_weight: float = 1.0, gaussian_ids: Optional[Tensor] = None,
        texture_adam=None):
    """Rasterize textured 2D Gaussians to pixels.

    textures [N, th, tw, 4] (RGBA, bilinear, sampled over splat coordinates
    [-texture_range, texture_range]).  Returns (render_colors [C,H,W,D],
    render_alphas [C,H,W,1], render_normals [C,H,W,3], render_distort
    [C,H,W,1], render_median [C,H,W,1], gs_contrib_sum [N], gs_contrib_count [N]).
    ``packed``: the per-Gaussian inputs are the [nnz, ...] entries of the packed
    projection and ``gaussian_ids`` [nnz] (no reference counterpart) names each
    entry's Gaussian, whose texture it samples (the reference samples the
    texture of the entry index there).
    ``texture_adam`` (a ``BF16TextureAdam`` whose ``textures`` are passed as
    ``textures``; no reference counterpart): the backward applies one selective Adam
    step to the textures instead of returning their gradient.
    """
    C = isect_offsets.size(0)
    N = textures.shape[0]
    if packed:
        if gaussian_ids is None:
            raise NotImplementedError("rasterize_to_pixels_textured_gaussians: packed=True needs gaussian_ids")
        nnz = means2d.size(0)
        assert means2d.shape == (nnz, 2), means2d.shape
        assert ray_transforms.shape == (nnz, 3, 3), ray_transforms.shape
        assert colors.shape[0] == nnz and colors.dim() == 2, colors.shape
        assert opacities.shape == (nnz,), opacities.shape
        assert gaussian_ids.shape == (nnz,), gaussian_ids.shape
        gaussian_ids = gaussian_ids.contiguous()
    else:
        gaussian_ids = None
        assert means2d.shape == (C, N, 2), means2d.shape
        assert ray_transforms.shape == (C, N, 3, 3), ray_transforms.shape
        assert colors.shape[:2] == (C, N), colors.shape
        assert opacities.shape == (C, N), opacities.shape
    assert textures.dim() == 4 and textures.shape[-1] == 4, textures.shape
    if backgrounds is not None:
        assert backgrounds.shape == (C, colors.shape[-1]), backgrounds.shape
        backgrounds = backgrounds.contiguous()
    colors, backgrounds, padded = _pad_channels(colors, backgrounds)
    tile_height, tile_width = isect_offsets.shape[1:3]
    assert tile_height * tile_size >= image_height, f"Assert Failed: {tile_height} * {tile_size} >= {image_height}"
    assert tile_width * tile_size >= image_width, f"Assert Failed: {tile_width} * {tile_size} >= {image_width}"
    if filtering == "mipmapped2":
        # [W]:1944-1989: square power-of-two textures, pyramid built (and differentiated) here
        if texture_adam is not None:
            raise NotImplementedError("mipmapped2: texture_adam not supported")
        assert textures.size(1) == textures.size(2)
        log_res = textures.size(1).bit_length() - 1
        assert textures.size(1) == 1 << log_res
        mip = generate_mipmap(textures.contiguous(), log_res, 4, tile_size)
        out = _RasterizeToPixelsMip2TexturedGaussians.apply(
            means2d.contiguous(), ray_transforms.contiguous(), colors.contiguous(), opacities.contiguous(), mip,
            log_res, texture_range_x, texture_range_y, normals.contiguous(), densify.contiguous(), backgrounds, masks,
            image_width, image_height, tile_size, isect_offsets.contiguous(), flatten_ids.contiguous(), absgrad,
            distloss, gs_contrib_threshold, g_weight, gaussian_ids)
        render_colors = out[0]
        if padded > 0:
            render_colors = render_colors[..., :-padded]
        return (render_colors,) + tuple(out[1:])
    if filtering not in tgs_fns:
        raise Exception(f"Unsupported filter type {filtering}")
    out = _RasterizeToPixelsTexturedGaussians.apply(
        tgs_fns[filtering][0], tgs_fns[filtering][1], means2d.contiguous(), ray_transforms.contiguous(),
        colors.contiguous(), opacities.contiguous(), textures.contiguous(), texture_range_x, texture_range_y,
        normals.contiguous(), densify.contiguous(), backgrounds, masks, image_width, image_height, tile_size,
        isect_offsets.contiguous(), flatten_ids.contiguous(), absgrad, distloss, gs_contrib_threshold, g_weight,
        gaussian_ids, texture_adam)
    render_colors = out[0]
    if padded > 0:
        render_colors = render_colors[..., :-padded]
    return (render_colors,) + tuple(out[1:])


# Texture gradients through per-pair records and a per-row gather (True), float atomics
# into a v_textures whose listed rows the forward zero-fills (False), or whichever is
# faster for the textures at hand (None): records where the library's streaming gather
# takes them (float32 compute, rows up to about 50x50 texels), atomics elsewhere.
# Measured on the 1M-Gaussian scene at 50x50 (one MI355X, same box): atomics 20.7-20.8
# ms/iter, records 19.0 ms/iter.  The backward's float atomics into the 40 GB gradient
# are performed at the memory side (1.8e8 requests, ~5 ms of its 10 ms); the records
# path writes 32-byte records instead and the gather sums each listed row in LDS and
# streams it out once.
TEXTURE_GRAD_RECORDS = None
# Between the two (used when TEXTURE_GRAD_RECORDS is False): the rows with at least
# HEAVY_ROW_PAIRS texture pairs (few rows, most pairs: large splats) take records and
# the gather, the others float atomics.  0 disables it.
HEAVY_ROW_PAIRS = 0
# With records, the rows absent from the tile lists are zero-filled along the
# rasterizer walks: the forward takes the rows whose index mod 8 is in this set
# (bit r), the backward the others; the gather writes the listed rows.  Both walks
# are bound by memory traffic, so the split changes little; the forward alone is
# fastest without the fill (19.0 vs 19.4 ms/iter with all of it in the forward).
FWD_FILL_RESIDUES = 0x00
# The texture gradient's zero-fill rides along the rasterizer walks (True), or is one
# launch before the forward (False; same result, used to time the rasterizer kernels
# without the fill): the whole gradient for float atomics, the rows absent from the
# tile lists for records (the gather writes the others).
FUSED_FILL = True


def _records_fit(textures: Tensor, dtype: torch.dtype) -> bool:
    """The gather keeps one texture row in LDS: at most 80 KB."""
    return textures.shape[1] * textures.shape[2] * 4 * torch.tensor([], dtype=dtype).element_size() <= 80 * 1024


# The forward keeps the texture sample of every composited in-range pair for the
# backward (16 bytes per pair, in slots of 1 KB per wavefront and list entry: 17 GB at
# 1M Gaussians, 1297x840), which then reads no texels.  None (default): when the
# backward will run with texture-gradient records (its lean sample-reading
# instantiation) and the slots fit; True: whenever they apply and fit; False: never.
# Measured at that size (one MI355X, one process): forward +0.1 ms (1.1 GB of samples),
# backward with records -0.4 ms (7.18 -> 6.77 ms), 17.43 -> 17.12 ms per iteration.
PAIR_SAMPLES = None
# at most this fraction of the device memory free when the forward runs
PAIR_SAMPLES_MEMORY_FRACTION = 0.25


def _use_pair_samples(bwd_fn, means2d, colors, textures, flatten_ids, tile_size, records: bool) -> bool:
    """Pair samples apply to the reference ``bilinear`` rule (the other rules need the
    texels' differences) in float32 with at most 3 channels, when the slots fit in a
    quarter of the free device memory (see PAIR_SAMPLES for when they are taken)."""
    if PAIR_SAMPLES is False or (PAIR_SAMPLES is None and not records):
        return False
    if bwd_fn != tgs_fns["bilinear"][1] or means2d.dtype != torch.float32:
        return False
    if colors.shape[-1] > 3 or textures.dim() != 4 or flatten_ids.numel() == 0:
        return False
    free, _ = torch.cuda.mem_get_info(means2d.device)
    return _C.pair_samples_bytes(flatten_ids, tile_size) <= PAIR_SAMPLES_MEMORY_FRACTION * free


def _use_records(textures: Tensor, dtype: torch.dtype) -> bool:
    """Texture gradients through records (see TEXTURE_GRAD_RECORDS)."""
    if TEXTURE_GRAD_RECORDS is None:
        dt = 1 if dtype == torch.float64 else 0
        return bool(lib().tgs_texture_gather_fast(dt, textures.shape[1], textures.shape[2]))
    return bool(TEXTURE_GRAD_RECORDS) and _records_fit(textures, dtype)


class _RasterizeToPixelsTexturedGaussians(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fwd_fn, bwd_fn, means2d, ray_transforms, colors, opacities, textures, texture_range_x,
                texture_range_y, normals, densify, backgrounds, masks, width, height, tile_size, isect_offsets,
                flatten_ids, absgrad, distloss, gs_contrib_threshold, g_weight, gaussian_ids, texture_adam):
        # The texture gradient is as large as the textures (40 GB at 1M x 50x50) and is never
        # zero-filled on its own.  With records (the default), the forward counts per row the
        # pairs that will contribute, the backward writes one record per pair, and the gather
        # sums each listed row's records in LDS and writes the row once; the rows absent from
        # the tile lists are zeroed along the forward and backward walks.  Without records, the
        # forward zeroes the listed rows along its walk, the backward accumulates into them
        # with float atomics and zeroes the other rows.
        ctx.v_tex = None
        ctx.tex_counts = None
        ctx.record_rows = False
        ctx.adam = None
        zero_rows = None
        if texture_adam is not None:
            # the backward steps the textures from per-pair records; the rows in this
            # step's tile lists are the visible ones
            if textures.data_ptr() != texture_adam.textures.data_ptr():
                raise ValueError("texture_adam: pass texture_adam.textures as the textures")
            listed = flatten_ids if gaussian_ids is None else gaussian_ids[flatten_ids.long()].int()
            ctx.adam = (texture_adam, _C.mark_rows(listed, textures.shape[0]))
            ctx.tex_counts = _C.tex_counts_buffer(textures)
        elif ctx.needs_input_grad[6] and textures.is_cuda and textures.shape[0] > 0:
            # float gradient (also for bfloat16 textures; autograd casts it to theirs)
            v_tex = torch.empty(textures.shape, dtype=means2d.dtype, device=textures.device)
            # texture rows present in the tile lists (packed: the entries' Gaussians)
            listed = flatten_ids if gaussian_ids is None else gaussian_ids[flatten_ids.long()].int()
            rows = _C.mark_rows(listed, textures.shape[0])
            if _use_records(textures, means2d.dtype):
                ctx.tex_counts = _C.tex_counts_buffer(textures)
                if not FUSED_FILL:
                    # the unlisted rows by one launch now; the walks fill nothing (mask value 2: no row)
                    _C.zero_fill_rows_(v_tex, rows, 0)
                    ctx.v_tex = (v_tex, rows, 2)
                else:
                    fwd_res = FWD_FILL_RESIDUES & 0xFF
                    if fwd_res:
                        zero_rows = (v_tex, rows, fwd_res << 8)  # mask value 0, these residues
                    bwd_res = 0xFF & ~fwd_res
                    # the backward's share; mask value 2 (no row) when the forward took every residue
                    ctx.v_tex = (v_tex, rows, (bwd_res << 8) if bwd_res else 2)
            elif not FUSED_FILL:  # one dense fill; the backward adds in place
                _C.zero_fill_(v_tex)
                ctx.v_tex = (v_tex, None, None)
            else:  # the forward zero-fills the listed rows and the backward adds in place
                zero_rows = (v_tex, rows, 1)
                ctx.v_tex = (v_tex, rows, 0)
                if HEAVY_ROW_PAIRS > 0 and _records_fit(textures, means2d.dtype):
                    # rows with many pairs through records and the gather (it overwrites them)
                    ctx.tex_counts = _C.tex_counts_buffer(textures)
                    ctx.record_rows = True
        # the forward records which lanes of each wavefront composited each list entry; the
        # backward visits only those entries, and under the reference rule in float32 it reads
        # the forward's texture samples instead of the texels
        ctx.live = _C.live_entries_buffer(flatten_ids, tile_size) if means2d.is_cuda else None
        ctx.samples = None
        records = ctx.tex_counts is not None and not ctx.record_rows
        if (ctx.live is not None and any(ctx.needs_input_grad)
                and _use_pair_samples(bwd_fn, means2d, colors, textures, flatten_ids, tile_size, records)):
            ctx.samples = _C.pair_samples_buffer(flatten_ids, tile_size)
        (render_colors, render_alphas, render_normals, render_distort, render_median, last_ids, median_ids,
         gs_contrib_sum, gs_contrib_count, accum_depth, trans) = _make_lazy_cuda_func(fwd_fn + "_full")(
            means2d, ray_transforms, colors, opacities, textures, texture_range_x, texture_range_y, normals,
            backgrounds, masks, width, height, tile_size, isect_offsets, flatten_ids, gs_contrib_threshold,
            g_weight, zero_rows=zero_rows, live_entries=ctx.live, tex_counts=ctx.tex_counts, gaussian_ids=gaussian_ids,
            pair_samples=ctx.samples)
        if ctx.tex_counts is not None:
            if ctx.record_rows:  # only the heavy rows take records
                ctx.tex_counts = torch.where(ctx.tex_counts >= HEAVY_ROW_PAIRS, ctx.tex_counts, 0)
            # record sizing starts now; the backward reads its total without waiting
            ctx.tex_counts = _C.TexRecordPlan(ctx.tex_counts)
        ctx.save_for_backward(means2d, ray_transforms, colors, opacities, textures, normals, densify, backgrounds,
                              masks, isect_offsets, flatten_ids, render_colors, render_alphas, last_ids,
                              median_ids, accum_depth, gaussian_ids, trans)
        ctx.bwd_fn = bwd_fn
        ctx.texture_range = (texture_range_x, texture_range_y)
        ctx.width, ctx.height, ctx.tile_size = width, height, tile_size
        ctx.absgrad, ctx.g_weight = absgrad, g_weight
        ctx.mark_non_differentiable(gs_contrib_sum, gs_contrib_count)
        return (render_colors, render_alphas, render_normals, render_distort, render_median, gs_contrib_sum,
                gs_contrib_count)

    @staticmethod
    def backward(ctx, v_render_colors, v_render_alphas, v_render_normals, v_render_distort, v_render_median,
                 v_gs_contrib_sum, v_gs_contrib_count):
        (means2d, ray_transforms, colors, opacities, textures, normals, densify, backgrounds, masks, isect_offsets,
         flatten_ids, render_colors, render_alphas, last_ids, median_ids, accum_depth,
         gaussian_ids, trans) = ctx.saved_tensors
        exact = ctx.bwd_fn == tgs_fns["bilinear_exact"][1]
        v_tex_buf, zero_rows = None, None
        if ctx.adam is not None:
            v_tex_buf = False  # records only
        elif ctx.v_tex is not None:
            v_tex_buf = ctx.v_tex[0]
            zero_rows = ctx.v_tex if ctx.v_tex[1] is not None else None
            ctx.v_tex = None
        (v_means2d_abs, v_means2d, v_ray_transforms, v_colors, v_opacities, v_textures, v_normals,
         v_densify) = _make_lazy_cuda_func(ctx.bwd_fn)(
            means2d, ray_transforms, colors, opacities, textures, ctx.texture_range[0], ctx.texture_range[1],
            normals, densify, backgrounds, masks, ctx.width, ctx.height, ctx.tile_size, isect_offsets, flatten_ids,
            ctx.g_weight, render_colors, render_alphas, last_ids, median_ids, v_render_colors.contiguous(),
            v_render_alphas.contiguous(), v_render_normals.contiguous(), v_render_distort.contiguous(),
            v_render_median.contiguous(), ctx.absgrad, render_accum_depth=accum_depth if exact else None,
            render_trans=trans, v_textures_out=v_tex_buf, zero_rows=zero_rows, live_entries=ctx.live,
            tex_counts=ctx.tex_counts, gaussian_ids=gaussian_ids, record_rows=ctx.record_rows,
            pair_samples=ctx.samples)
        ctx.live = None
        ctx.samples = None
        ctx.tex_counts = None
        if ctx.adam is not None:
            opt, rows = ctx.adam
            opt.step_from_records(v_textures, rows)
            v_textures = None
            ctx.adam = None
        if ctx.absgrad:
            means2d.absgrad = v_means2d_abs
        v_backgrounds = None
        if ctx.needs_input_grad[11]:
            v_backgrounds = (v_render_colors * (1.0 - render_alphas)).sum(dim=(1, 2))
        return (None, None, v_means2d, v_ray_transforms, v_colors, v_opacities, v_textures, None, None, v_normals,
                v_densify, v_backgrounds, None, None, None, None, None, None, None, None, None, None, None, None)


class _GenerateMipmap(torch.autograd.Function):
    """Box-filter pyramid of textures [N, R, R, C] -> [N, (4^(L+1) - 1) / 3, C] ([W]:4013-4049)."""

    @staticmethod
    def forward(ctx, textures, log_texture_res, log_reduce, tile_size):
        ctx.N, ctx.channels = textures.shape[0], textures.shape[-1]
        ctx.log_texture_res, ctx.log_reduce, ctx.tile_size = log_texture_res, log_reduce, tile_size
        return _make_lazy_cuda_func("generate_mipmap_fwd")(textures, log_texture_res, log_reduce, tile_size)

    @staticmethod
    def backward(ctx, v_mip_textures):
        v = _make_lazy_cuda_func("generate_mipmap_bwd")(ctx.N, ctx.channels, ctx.log_texture_res, ctx.log_reduce,
                                                        ctx.tile_size, v_mip_textures.contiguous())
        return v, None, None, None


def generate_mipmap(textures: Tensor, log_texture_res: int, log_reduce: int, tile_size: int) -> Tensor:
    return _GenerateMipmap.apply(textures, log_texture_res, log_reduce, tile_size)


class _RasterizeToPixelsMip2TexturedGaussians(torch.autograd.Function):
    """The ``mipmapped2`` filter ([W]:3375-3560): trilinear lookups in the pyramid with
    the level of detail from the pixel footprint; the reference's gradient rule."""

    @staticmethod
    def forward(ctx, means2d, ray_transforms, colors, opacities, mip_textures, log_texture_res, texture_range_x,
                texture_range_y, normals, densify, backgrounds, masks, width, height, tile_size, isect_offsets,
                flatten_ids, absgrad, distloss, gs_contrib_threshold, g_weight, gaussian_ids):
        ctx.live = _C.live_entries_buffer(flatten_ids, tile_size) if means2d.is_cuda else None
        # the pyramid's gradient, zero-filled along the rasterizer walks (listed rows by the
        # forward, the others by the backward), as for the bilinear filter
        ctx.v_tex = None
        if ctx.needs_input_grad[4] and mip_textures.is_cuda:
            v_tex = torch.empty(mip_textures.shape, dtype=means2d.dtype, device=mip_textures.device)
            listed = flatten_ids if gaussian_ids is None else gaussian_ids[flatten_ids.long()].int()
            ctx.v_tex = (v_tex, _C.mark_rows(listed, mip_textures.shape[0]))
        (render_colors, render_alphas, render_normals, render_distort, render_median, last_ids, median_ids,
         gs_contrib_sum, gs_contrib_count, _, trans) = _make_lazy_cuda_func(
            "rasterize_to_pixels_fwd_mip2_textured_gaussians_full")(
            means2d, ray_transforms, colors, opacities, mip_textures, log_texture_res, texture_range_x,
            texture_range_y, normals, backgrounds, masks, width, height, tile_size, isect_offsets, flatten_ids,
            gs_contrib_threshold, g_weight, live_entries=ctx.live, zero_rows=ctx.v_tex, gaussian_ids=gaussian_ids)
        ctx.save_for_backward(means2d, ray_transforms, colors, opacities, mip_textures, normals, densify, backgrounds,
                              masks, isect_offsets, flatten_ids, render_colors, render_alphas, last_ids, median_ids,
                              trans, gaussian_ids)
        ctx.log_texture_res = log_texture_res
        ctx.texture_range = (texture_range_x, texture_range_y)
        ctx.width, ctx.height, ctx.tile_size = width, height, tile_size
        ctx.absgrad, ctx.g_weight = absgrad, g_weight
        ctx.mark_non_differentiable(gs_contrib_sum, gs_contrib_count)
        return (render_colors, render_alphas, render_normals, render_distort, render_median, gs_contrib_sum,
                gs_contrib_count)

    @staticmethod
    def backward(ctx, v_render_colors, v_render_alphas, v_render_normals, v_render_distort, v_render_median,
                 v_gs_contrib_sum, v_gs_contrib_count):
        (means2d, ray_transforms, colors, opacities, mip_textures, normals, densify, backgrounds, masks,
         isect_offsets, flatten_ids, render_colors, render_alphas, last_ids, median_ids, trans,
         gaussian_ids) = ctx.saved_tensors
        (v_means2d_abs, v_means2d, v_ray_transforms, v_colors, v_opacities, v_textures, v_normals,
         v_densify) = _make_lazy_cuda_func("rasterize_to_pixels_bwd_mip2_textured_gaussians")(
            means2d, ray_transforms, colors, opacities, mip_textures, ctx.log_texture_res, ctx.texture_range[0],
            ctx.texture_range[1], normals, densify, backgrounds, masks, ctx.width, ctx.height, ctx.tile_size,
            isect_offsets, flatten_ids, ctx.g_weight, render_colors, render_alphas, last_ids, median_ids,
            v_render_colors.contiguous(), v_render_alphas.contiguous(), v_render_normals.contiguous(),
            v_render_distort.contiguous(), v_render_median.contiguous(), ctx.absgrad, live_entries=ctx.live,
            v_textures_out=ctx.v_tex[0] if ctx.v_tex is not None else None, zero_rows=ctx.v_tex, render_trans=trans,
            gaussian_ids=gaussian_ids)
        ctx.live = None
        ctx.v_tex = None
        if ctx.absgrad:
            means2d.absgrad = v_means2d_abs
        v_backgrounds = None
        if ctx.needs_input_grad[10]:
            v_backgrounds = (v_render_colors * (1.0 - render_alphas)).sum(dim=(1, 2))
        return (v_means2d, v_ray_transforms, v_colors, v_opacities, v_textures, None, None, None, v_normals,
                v_densify, v_backgrounds, None, None, None, None, None, None, None, None, None, None, None)


def rasterize_to_pixels_2dgs(means2d: Tensor, ray_transforms: Tensor, colors: Tensor, opacities: Tensor,
                             normals: Tensor, densify: Tensor, image_width: int, image_height: int,
                             tile_size: int, isect_offsets: Tensor, flatten_ids: Tensor,
                             backgrounds: Optional[Tensor] = None, masks: Optional[Tensor] = None,
                             packed: bool = False, absgrad: bool = False, distloss: bool = False,
                             gs_contrib_threshold: float = 0.0, grad_mode: str = "reference",
                             gaussian_ids: Optional[Tensor] = None, n_gaussians: Optional[int] = None):
    """Rasterize plain 2D Gaussians to pixels (``alpha = min(.999, o exp(-sigma))``).
    ``grad_mode="exact"`` uses the forward's own distortion accumulator in the
    backward (the reference seeds it from ``render_colors``).  ``packed``: [nnz, ...]
    entries with ``gaussian_ids`` [nnz] naming their Gaussians (the contribution
    counters are per Gaussian, ``n_gaussians`` of them)."""
    C = isect_offsets.size(0)
    if packed:
        if gaussian_ids is None:
            raise NotImplementedError("rasterize_to_pixels_2dgs: packed=True needs gaussian_ids")
        nnz = means2d.size(0)
        assert means2d.shape == (nnz, 2), means2d.shape
        assert ray_transforms.shape == (nnz, 3, 3), ray_transforms.shape
        assert colors.shape[0] == nnz and colors.dim() == 2, colors.shape
        assert opacities.shape == (nnz,), opacities.shape
        gaussian_ids = gaussian_ids.contiguous()
    else:
        gaussian_ids, n_gaussians = None, None
        N = means2d.size(1)
        assert means2d.shape == (C, N, 2), means2d.shape
        assert ray_transforms.shape == (C, N, 3, 3), ray_transforms.shape
        assert colors.shape[:2] == (C, N), colors.shape
        assert opacities.shape == (C, N), opacities.shape
    if backgrounds is not None:
        assert backgrounds.shape == (C, colors.shape[-1]), backgrounds.shape
        backgrounds = backgrounds.contiguous()
    colors, backgrounds, padded = _pad_channels(colors, backgrounds)
    tile_height, tile_width = isect_offsets.shape[1:3]
    assert tile_height * tile_size >= image_height
    assert tile_width * tile_size >= image_width
    out = _RasterizeToPixels2DGS.apply(
        means2d.contiguous(), ray_transforms.contiguous(), colors.contiguous(), opacities.contiguous(),
        normals.contiguous(), densify.contiguous(), backgrounds, masks, image_width, image_height, tile_size,
        isect_offsets.contiguous(), flatten_ids.contiguous(), absgrad, distloss, gs_contrib_threshold,
        grad_mode == "exact", gaussian_ids, n_gaussians)
    render_colors = out[0]
    if padded > 0:
        render_colors = render_colors[..., :-padded]
    return (render_colors,) + tuple(out[1:])


class _RasterizeToPixels2DGS(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means2d, ray_transforms, colors, opacities, normals, densify, backgrounds, masks, width,
                height, tile_size, isect_offsets, flatten_ids, absgrad, distloss, gs_contrib_threshold, exact,
                gaussian_ids, n_gaussians):
        ctx.live = _C.live_entries_buffer(flatten_ids, tile_size) if means2d.is_cuda else None
        ctx.packed = dict(gaussian_ids=gaussian_ids, n_gaussians=n_gaussians)
        (render_colors, render_alphas, render_normals, render_distort, render_median, last_ids, median_ids,
         gs_contrib_sum, gs_contrib_count, accum_depth, trans) = _make_lazy_cuda_func(
            "rasterize_to_pixels_fwd_2dgs_full")(
            means2d, ray_transforms, colors, opacities, normals, backgrounds, masks, width, height, tile_size,
            isect_offsets, flatten_ids, gs_contrib_threshold, live_entries=ctx.live, **ctx.packed)
        ctx.save_for_backward(means2d, ray_transforms, colors, opacities, normals, densify, backgrounds, masks,
                              isect_offsets, flatten_ids, render_colors, render_alphas, last_ids, median_ids,
                              accum_depth, trans)
        ctx.width, ctx.height, ctx.tile_size, ctx.absgrad = width, height, tile_size, absgrad
        ctx.exact = exact
        ctx.mark_non_differentiable(gs_contrib_sum, gs_contrib_count)
        return (render_colors, render_alphas, render_normals, render_distort, render_median, gs_contrib_sum,
                gs_contrib_count)

    @staticmethod
    def backward(ctx, v_render_colors, v_render_alphas, v_render_normals, v_render_distort, v_render_median,
                 v_gs_contrib_sum, v_gs_contrib_count):
        (means2d, ray_transforms, colors, opacities, normals, densify, backgrounds, masks, isect_offsets,
         flatten_ids, render_colors, render_alphas, last_ids, median_ids, accum_depth, trans) = ctx.saved_tensors
        (v_means2d_abs, v_means2d, v_ray_transforms, v_colors, v_opacities, v_normals,
         v_densify) = _make_lazy_cuda_func("rasterize_to_pixels_bwd_2dgs")(
            means2d, ray_transforms, colors, opacities, normals, densify, backgrounds, masks, ctx.width,
            ctx.height, ctx.tile_size, isect_offsets, flatten_ids, render_colors, render_alphas, last_ids,
            median_ids, v_render_colors.contiguous(), v_render_alphas.contiguous(), v_render_normals.contiguous(),
            v_render_distort.contiguous(), v_render_median.contiguous(), ctx.absgrad,
            render_accum_depth=accum_depth if ctx.exact else None, live_entries=ctx.live, render_trans=trans,
            **ctx.packed)
        ctx.live = None
        if ctx.absgrad:
            means2d.absgrad = v_means2d_abs
        v_backgrounds = None
        if ctx.needs_input_grad[6]:
            v_backgrounds = (v_render_colors * (1.0 - render_alphas)).sum(dim=(1, 2))
        return (v_means2d, v_ray_transforms, v_colors, v_opacities, v_normals, v_densify, v_backgrounds, None,
                None, None, None, None, None, None, None, None, None, None, None)
