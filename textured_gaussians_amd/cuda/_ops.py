"""The native op table: same names, positional arguments and returns as the
reference's pybind registry (``textured_gaussians/cuda/csrc/ext.cpp`` /
``bindings.h``) for the textured-2DGS render path.

Each op validates its tensors like the reference (``kernel_utils.h:11-17``:
"x must be a CUDA tensor", "x must be contiguous"), allocates its outputs with
torch's caching allocator on the input's device, and calls the C ABI of
libtgs_hip.so on torch's current stream.  Nothing here computes on the CPU.
"""
import ctypes
from typing import Optional, Tuple

import torch
from torch import Tensor

from ._backend import check, lib

_DTYPES = {torch.float32: 0, torch.float64: 1}

UV_NONE, UV_EXACT, UV_DOUBLE = 0, 1, 2
VJP_REFERENCE, VJP_REFERENCE_NORM, VJP_EXACT = 0, 1, 2


def _p(t: Optional[Tensor]):
    return None if t is None else t.data_ptr()


# Optional per-launch timing of the rasterizer kernels: when a dict is
# installed, each rasterizer call records a pair of HIP events on the stream it
# launches on, under "raster_fwd" / "raster_bwd".
_TIMER = None


def set_kernel_timer(store: Optional[dict]) -> None:
    global _TIMER
    _TIMER = store


# library event slots (tgs_set_raster_events) recorded inside an op: the rasterizer
# kernel itself, and for the backward also the texture-gradient gather
_RASTER_EVENTS = {"raster_fwd": (("raster_fwd", 0),), "raster_bwd": (("raster_bwd", 1), ("tex_gather", 2))}


class _Timed:
    """HIP events on an op's stream: around the rasterizer kernel alone for
    "raster_fwd" / "raster_bwd", and around the backward's texture gather as
    "tex_gather" (recorded by the library, tgs_set_raster_events), around the whole
    op otherwise.  Holds the device only: a reference to the op's tensor would outlive
    the op (and a gradient buffer with an extra reference is copied by autograd
    instead of handed over)."""
    __slots__ = ("name", "device", "ev")

    def __init__(self, name, device):
        self.name, self.device = name, device

    def __enter__(self):
        if _TIMER is not None:
            st = torch.cuda.current_stream(self.device)
            slots = _RASTER_EVENTS.get(self.name, ((self.name, None),))
            self.ev = []
            for key, which in slots:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(st)
                if which is not None:
                    # recorded once so that they exist; the library records them again
                    # (a gather that does not run leaves them at ~0 ms)
                    e1.record(st)
                    check(lib().tgs_set_raster_events(which, e0.cuda_event, e1.cuda_event), "set_raster_events")
                self.ev.append((key, which, e0, e1))

    def __exit__(self, *exc):
        if _TIMER is not None:
            for key, which, e0, e1 in self.ev:
                if which is None:
                    e1.record(torch.cuda.current_stream(self.device))
                else:
                    check(lib().tgs_set_raster_events(which, None, None), "set_raster_events")
                _TIMER.setdefault(key, []).append((e0, e1))


def _timed(name, t: Tensor):
    return _Timed(name, t.device)


def _stream(t: Tensor):
    return torch.cuda.current_stream(t.device).cuda_stream


def wait_event(ev) -> None:
    """Wait on the host for a recorded event by polling it.  hipEventSynchronize on
    these (non-blocking) events was measured to return ~400 us after the event
    completed, and a synchronous copy ~70 us after, with the device idle meanwhile
    whenever the host has the next launches to issue; polling returns within a few
    microseconds."""
    while not ev.query():
        pass


def read_scalar(t: Tensor) -> int:
    """int(t) of a one-element device tensor: an asynchronous copy to pinned host
    memory on the current stream, then ``wait_event`` on it."""
    host = torch.empty(1, dtype=t.dtype, pin_memory=True)
    host.copy_(t.reshape(1), non_blocking=True)
    ev = torch.cuda.Event()
    ev.record(torch.cuda.current_stream(t.device))
    wait_event(ev)
    return int(host[0])


def _check_input(t: Tensor, name: str) -> None:
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _dtype_code(t: Tensor, name: str) -> int:
    if t.dtype not in _DTYPES:
        raise RuntimeError(f"{name}: unsupported dtype {t.dtype} (float32 or float64)")
    return _DTYPES[t.dtype]


def _same(dtype: torch.dtype, **ts: Tensor) -> None:
    for n, t in ts.items():
        if t is not None and t.dtype != dtype:
            raise RuntimeError(f"{n} must have dtype {dtype}, got {t.dtype}")


TEX_BF16 = 0x100  # OR-ed into the rasterizers' dtype: bfloat16 textures
TEX_MIP2 = 0x200  # OR-ed into the rasterizers' dtype: the mipmapped2 filter over a pyramid
FWD_NO_CULL = 0x400  # OR-ed into the forward's dtype: visit every list entry (no candidate pre-pass)
BWD_RECORD_ROWS = 0x800  # OR-ed into the backward's dtype: only rows with a nonzero count take records
# The forward's candidate pre-pass (same outputs; off only to compare against the full walk)
FWD_CULL = True
# Fused row fill: stores (1 KB each) per visited list entry and wavefront; 0 takes the
# library's (1 in the forward, 2 in the backward).  The remainder of a wave's rows goes
# out when its walk ends.
FILL_STEPS = 0


def _tex_dtype_flag(textured, textures, means2d):
    """TEX_BF16 for bfloat16 textures (float32 compute only), else 0; checks the dtype."""
    if not textured or textures.dtype == means2d.dtype:
        return 0
    if textures.dtype == torch.bfloat16 and means2d.dtype == torch.float32:
        return TEX_BF16
    raise RuntimeError(f"textures must have dtype {means2d.dtype} (or bfloat16 with float32), got {textures.dtype}")


# ------------------------------------------------------------------ projection
def fully_fused_projection_fwd_2dgs(means, quats, scales, viewmats, Ks, image_width, image_height,
                                    eps2d, near_plane, far_plane, radius_clip):
    for n, t in dict(means=means, quats=quats, scales=scales, viewmats=viewmats, Ks=Ks).items():
        _check_input(t, n)
    dt = _dtype_code(means, "means")
    _same(means.dtype, quats=quats, scales=scales, viewmats=viewmats, Ks=Ks)
    N, C = means.size(0), viewmats.size(0)
    with torch.cuda.device(means.device):
        radii = torch.empty((C, N), dtype=torch.int32, device=means.device)
        means2d = torch.empty((C, N, 2), dtype=means.dtype, device=means.device)
        depths = torch.empty((C, N), dtype=means.dtype, device=means.device)
        rt = torch.empty((C, N, 3, 3), dtype=means.dtype, device=means.device)
        normals = torch.empty((C, N, 3), dtype=means.dtype, device=means.device)
        check(lib().tgs_projection_2dgs_fwd(dt, C, N, _p(means), _p(quats), _p(scales), _p(viewmats), _p(Ks),
                                            int(image_width), int(image_height), float(near_plane),
                                            float(far_plane), float(radius_clip), _p(radii), _p(means2d),
                                            _p(depths), _p(rt), _p(normals), _stream(means)),
              "fully_fused_projection_fwd_2dgs")
    return radii, means2d, depths, rt, normals


def _projection_bwd(mode, means, quats, scales, viewmats, Ks, image_width, image_height, radii, ray_transforms,
                    v_means2d, v_depths, v_normals, v_ray_transforms, viewmats_requires_grad):
    ts = dict(means=means, quats=quats, scales=scales, viewmats=viewmats, Ks=Ks, radii=radii,
              ray_transforms=ray_transforms, v_means2d=v_means2d, v_depths=v_depths, v_normals=v_normals,
              v_ray_transforms=v_ray_transforms)
    for n, t in ts.items():
        _check_input(t, n)
    dt = _dtype_code(means, "means")
    N, C = means.size(0), viewmats.size(0)
    with torch.cuda.device(means.device):
        v_means = torch.zeros_like(means)
        v_quats = torch.zeros_like(quats)
        v_scales = torch.zeros_like(scales)
        v_viewmats = torch.zeros_like(viewmats) if viewmats_requires_grad else None
        vv = v_viewmats if (viewmats_requires_grad and mode == VJP_EXACT) else None
        check(lib().tgs_projection_2dgs_bwd(dt, C, N, _p(means), _p(quats), _p(scales), _p(viewmats), _p(Ks),
                                            _p(radii), _p(ray_transforms), _p(v_means2d), _p(v_depths),
                                            _p(v_normals), _p(v_ray_transforms), mode, _p(v_means), _p(v_quats),
                                            _p(v_scales), _p(vv), _stream(means)),
              "fully_fused_projection_bwd_2dgs")
    return v_means, v_quats, v_scales, v_viewmats


def fully_fused_projection_bwd_2dgs(*args):
    """Reference VJP (``fully_fused_projection_2dgs_bwd.cu``)."""
    return _projection_bwd(VJP_REFERENCE, *args)


def fully_fused_projection_bwd2_2dgs(*args):
    """Reference ``norm_rot_grad`` VJP (``fully_fused_projection_2dgs_bwd2.cu``)."""
    return _projection_bwd(VJP_REFERENCE_NORM, *args)


def fully_fused_projection_bwd_exact_2dgs(*args):
    """True derivative of the forward (no reference counterpart)."""
    return _projection_bwd(VJP_EXACT, *args)


def fully_fused_projection_packed_fwd_2dgs(means, quats, scales, viewmats, Ks, image_width, image_height,
                                           near_plane, far_plane, radius_clip):
    """Packed layout: (indptr [C+1], camera_ids [nnz], gaussian_ids [nnz], radii, means2d, depths,
    ray_transforms, normals), the visible (camera, Gaussian) entries in (camera, Gaussian) order."""
    for n, t in dict(means=means, quats=quats, scales=scales, viewmats=viewmats, Ks=Ks).items():
        _check_input(t, n)
    dt = _dtype_code(means, "means")
    _same(means.dtype, quats=quats, scales=scales, viewmats=viewmats, Ks=Ks)
    N, C = means.size(0), viewmats.size(0)
    dev, fdt = means.device, means.dtype
    L = lib()
    args = (_p(means), _p(quats), _p(scales), _p(viewmats), _p(Ks), int(image_width), int(image_height),
            float(near_plane), float(far_plane), float(radius_clip))
    with torch.cuda.device(dev):
        st = _stream(means)
        flags = torch.empty(C * N, dtype=torch.int32, device=dev)
        incl = torch.empty(C * N, dtype=torch.int64, device=dev)
        temp = torch.empty(max(1, L.tgs_isect_count_temp_bytes(C * N)), dtype=torch.uint8, device=dev)
        check(L.tgs_projection_2dgs_packed_count(dt, C, N, *args, _p(flags), _p(incl), _p(temp), temp.numel(), st),
              "fully_fused_projection_packed_fwd_2dgs")
        nnz = read_scalar(incl[-1]) if C * N > 0 else 0  # device -> host, as the reference's two-pass kernel
        indptr = torch.empty(C + 1, dtype=torch.int64, device=dev)
        camera_ids = torch.empty(nnz, dtype=torch.int64, device=dev)
        gaussian_ids = torch.empty(nnz, dtype=torch.int64, device=dev)
        radii = torch.empty(nnz, dtype=torch.int32, device=dev)
        means2d = torch.empty((nnz, 2), dtype=fdt, device=dev)
        depths = torch.empty(nnz, dtype=fdt, device=dev)
        rt = torch.empty((nnz, 3, 3), dtype=fdt, device=dev)
        normals = torch.empty((nnz, 3), dtype=fdt, device=dev)
        check(L.tgs_projection_2dgs_packed_fill(dt, C, N, *args, _p(incl), _p(indptr), _p(camera_ids),
                                                _p(gaussian_ids), _p(radii), _p(means2d), _p(depths), _p(rt),
                                                _p(normals), st), "fully_fused_projection_packed_fwd_2dgs")
    return indptr, camera_ids, gaussian_ids, radii, means2d, depths, rt, normals


def _projection_packed_bwd(mode, means, quats, scales, viewmats, Ks, image_width, image_height, camera_ids,
                           gaussian_ids, ray_transforms, v_means2d, v_depths, v_normals, v_ray_transforms,
                           viewmats_requires_grad, sparse_grad):
    ts = dict(means=means, quats=quats, scales=scales, viewmats=viewmats, Ks=Ks, camera_ids=camera_ids,
              gaussian_ids=gaussian_ids, ray_transforms=ray_transforms, v_means2d=v_means2d, v_depths=v_depths,
              v_normals=v_normals, v_ray_transforms=v_ray_transforms)
    for n, t in ts.items():
        _check_input(t, n)
    dt = _dtype_code(means, "means")
    N, C, nnz = means.size(0), viewmats.size(0), camera_ids.numel()
    with torch.cuda.device(means.device):
        if sparse_grad:  # per-entry values, as the reference's sparse gradients
            v_means = torch.zeros((nnz, 3), dtype=means.dtype, device=means.device)
            v_quats = torch.zeros((nnz, 4), dtype=means.dtype, device=means.device)
            v_scales = torch.zeros((nnz, 3), dtype=means.dtype, device=means.device)
        else:
            v_means = torch.zeros_like(means)
            v_quats = torch.zeros_like(quats)
            v_scales = torch.zeros_like(scales)
        v_viewmats = torch.zeros_like(viewmats) if viewmats_requires_grad else None
        vv = v_viewmats if (viewmats_requires_grad and mode == VJP_EXACT) else None
        check(lib().tgs_projection_2dgs_packed_bwd(dt, C, N, nnz, _p(camera_ids), _p(gaussian_ids), int(sparse_grad),
                                                   _p(means), _p(quats), _p(scales), _p(viewmats), _p(Ks), None,
                                                   _p(ray_transforms), _p(v_means2d), _p(v_depths), _p(v_normals),
                                                   _p(v_ray_transforms), mode, _p(v_means), _p(v_quats),
                                                   _p(v_scales), _p(vv), _stream(means)),
              "fully_fused_projection_packed_bwd_2dgs")
    return v_means, v_quats, v_scales, v_viewmats


def fully_fused_projection_packed_bwd_2dgs(*args):
    """Reference VJP, packed layout."""
    return _projection_packed_bwd(VJP_REFERENCE, *args)


def fully_fused_projection_packed_bwd2_2dgs(*args):
    """Reference ``norm_rot_grad`` VJP, packed layout."""
    return _projection_packed_bwd(VJP_REFERENCE_NORM, *args)


def fully_fused_projection_packed_bwd_exact_2dgs(*args):
    """True derivative, packed layout (no reference counterpart)."""
    return _projection_packed_bwd(VJP_EXACT, *args)


# ------------------------------------------------------------------- isect
def _bit_length(x: int) -> int:
    return int(x).bit_length()


def isect_tiles(means2d, radii, depths, camera_ids, gaussian_ids, C, tile_size, tile_width, tile_height,
                sort, double_buffer):
    _check_input(means2d, "means2d")
    _check_input(radii, "radii")
    _check_input(depths, "depths")
    packed = means2d.dim() == 2
    if packed:
        if camera_ids is None or gaussian_ids is None:
            raise RuntimeError("When packed is set, camera_ids and gaussian_ids must be provided.")
        _check_input(camera_ids, "camera_ids")
    if means2d.dtype != torch.float32 or depths.dtype != torch.float32:
        # the key layout stores float32 depth bits (isect_tiles.cu:92)
        means2d = means2d.float().contiguous()
        depths = depths.float().contiguous()
    # packed entries are counted as one row of nnz; the camera comes from camera_ids
    C_, N = (1, means2d.size(0)) if packed else (means2d.size(0), means2d.size(1))
    n = C_ * N
    dev = means2d.device
    L = lib()
    n_tiles = tile_width * tile_height
    tile_bits = _bit_length(n_tiles)
    cam_bits = _bit_length(C)
    assert tile_bits + cam_bits <= 32
    with torch.cuda.device(dev):
        st = _stream(means2d)
        tpg = torch.empty((N,) if packed else (C_, N), dtype=torch.int32, device=dev)
        if n == 0:
            return tpg, torch.empty(0, dtype=torch.int64, device=dev), torch.empty(0, dtype=torch.int32, device=dev)
        cum = torch.empty(n, dtype=torch.int64, device=dev)
        tb = L.tgs_isect_count_temp_bytes(n)
        temp = torch.empty(max(tb, 1), dtype=torch.uint8, device=dev)
        check(L.tgs_isect_tiles_count(C_, N, _p(means2d), _p(radii), tile_size, tile_width, tile_height, _p(tpg),
                                      _p(cum), _p(temp), tb, st), "isect_tiles")
        n_isects = read_scalar(cum[-1])  # device -> host (isect_tiles.cu:200-201)
        isect_ids = torch.empty(n_isects, dtype=torch.int64, device=dev)
        flatten_ids = torch.empty(n_isects, dtype=torch.int32, device=dev)
        if n_isects == 0:
            return tpg, isect_ids, flatten_ids
        if packed:
            check(L.tgs_isect_tiles_fill_packed(C, N, _p(camera_ids), _p(means2d), _p(radii), _p(depths), _p(cum),
                                                tile_size, tile_width, tile_height, _p(isect_ids),
                                                _p(flatten_ids), st), "isect_tiles")
        else:
            check(L.tgs_isect_tiles_fill(C_, N, _p(means2d), _p(radii), _p(depths), _p(cum), tile_size,
                                         tile_width, tile_height, _p(isect_ids), _p(flatten_ids), st),
                  "isect_tiles")
        if not sort:
            return tpg, isect_ids, flatten_ids
        keys_alt = torch.empty_like(isect_ids)
        vals_alt = torch.empty_like(flatten_ids)
        sb = L.tgs_sort_pairs_temp_bytes(n_isects)
        stemp = torch.empty(sb, dtype=torch.uint8, device=dev)
        sel = ctypes.c_int32(0)
        check(L.tgs_sort_pairs(n_isects, 32 + tile_bits + cam_bits, _p(isect_ids), _p(flatten_ids), _p(keys_alt),
                               _p(vals_alt), _p(stemp), sb, ctypes.byref(sel), st), "isect_tiles sort")
        if sel.value == 1:
            return tpg, keys_alt, vals_alt
        return tpg, isect_ids, flatten_ids


def isect_offset_encode(isect_ids, C, tile_width, tile_height):
    _check_input(isect_ids, "isect_ids")
    dev = isect_ids.device
    with torch.cuda.device(dev):
        offsets = torch.empty((C, tile_height, tile_width), dtype=torch.int32, device=dev)
        check(lib().tgs_isect_offset_encode(isect_ids.numel(), _p(isect_ids), C, tile_width, tile_height,
                                            _p(offsets), _stream(isect_ids)), "isect_offset_encode")
    return offsets


# -------------------------------------------------------------------- SH
def compute_sh_fwd(degrees_to_use, dirs, coeffs, masks):
    _check_input(dirs, "dirs")
    _check_input(coeffs, "coeffs")
    if masks is not None:
        _check_input(masks, "masks")
    if coeffs.size(-1) != 3:
        raise RuntimeError("coeffs must have last dimension 3")
    if dirs.size(-1) != 3:
        raise RuntimeError("dirs must have last dimension 3")
    dt = _dtype_code(dirs, "dirs")
    K = coeffs.size(-2)
    n = dirs.numel() // 3
    with torch.cuda.device(dirs.device):
        colors = torch.empty_like(dirs)
        check(lib().tgs_sh_fwd(dt, n, K, degrees_to_use, _p(dirs), _p(coeffs),
                               _p(masks.view(torch.uint8)) if masks is not None else None, _p(colors),
                               _stream(dirs)), "compute_sh_fwd")
    return colors


def compute_sh_bwd(K, degrees_to_use, dirs, coeffs, masks, v_colors, compute_v_dirs):
    _check_input(dirs, "dirs")
    _check_input(coeffs, "coeffs")
    _check_input(v_colors, "v_colors")
    if masks is not None:
        _check_input(masks, "masks")
    dt = _dtype_code(dirs, "dirs")
    n = dirs.numel() // 3
    with torch.cuda.device(dirs.device):
        v_coeffs = torch.empty_like(coeffs)
        v_dirs = torch.zeros_like(dirs) if compute_v_dirs else None
        check(lib().tgs_sh_bwd(dt, n, K, degrees_to_use, _p(dirs), _p(coeffs),
                               _p(masks.view(torch.uint8)) if masks is not None else None, _p(v_colors),
                               _p(v_coeffs), _p(v_dirs), _stream(dirs)), "compute_sh_bwd")
    return v_coeffs, v_dirs


def sh_cams_supported(coeffs) -> bool:
    """Whether ``compute_sh_{fwd,bwd}_cams`` take these coefficients (float32, K*3 <= 48)."""
    return coeffs.dtype == torch.float32 and coeffs.dim() == 3 and coeffs.shape[1] * 3 <= 48


def compute_sh_fwd_cams(degrees_to_use, dirs, coeffs, masks):
    """colors [C, N, 3] of dirs [C, N, 3] against one coefficient set coeffs [N, K, 3]
    (masks [C, N] or None): ``compute_sh_fwd`` of the coefficients expanded to every
    camera, without the expanded copy."""
    for t, name in ((dirs, "dirs"), (coeffs, "coeffs")):
        _check_input(t, name)
    if masks is not None:
        _check_input(masks, "masks")
    C, N = dirs.shape[0], dirs.shape[1]
    if dirs.shape != (C, N, 3) or coeffs.dim() != 3 or coeffs.shape[0] != N or coeffs.shape[2] != 3:
        raise RuntimeError(f"sh_fwd_cams: dirs [C, N, 3] and coeffs [N, K, 3] expected ({dirs.shape}, {coeffs.shape})")
    with torch.cuda.device(dirs.device):
        colors = torch.empty_like(dirs)
        check(lib().tgs_sh_fwd_cams(_dtype_code(dirs, "dirs"), C, N, coeffs.shape[1], degrees_to_use, _p(dirs),
                                    _p(coeffs), _p(masks.view(torch.uint8)) if masks is not None else None,
                                    _p(colors), _stream(dirs)), "compute_sh_fwd_cams")
    return colors


def compute_sh_bwd_cams(degrees_to_use, dirs, coeffs, masks, v_colors, compute_v_dirs):
    """(v_coeffs [N, K, 3] summed over the cameras, v_dirs [C, N, 3] or None)."""
    for t, name in ((dirs, "dirs"), (coeffs, "coeffs"), (v_colors, "v_colors")):
        _check_input(t, name)
    if masks is not None:
        _check_input(masks, "masks")
    C, N = dirs.shape[0], dirs.shape[1]
    with torch.cuda.device(dirs.device):
        v_coeffs = torch.empty_like(coeffs)
        v_dirs = torch.empty_like(dirs) if compute_v_dirs else None
        check(lib().tgs_sh_bwd_cams(_dtype_code(dirs, "dirs"), C, N, coeffs.shape[1], degrees_to_use, _p(dirs),
                                    _p(coeffs), _p(masks.view(torch.uint8)) if masks is not None else None,
                                    _p(v_colors), _p(v_coeffs), _p(v_dirs), _stream(dirs)), "compute_sh_bwd_cams")
    return v_coeffs, v_dirs


# --------------------------------------------------------------- rasterizer
def _raster_workspace(dt, C, N, dev, tile_offsets):
    """The rasterizer workspace with room for the tile order after it (the kernels then
    take the tiles longest list first)."""
    L = lib()
    nbytes = L.tgs_rasterize_workspace_bytes(dt, C, N)
    nbytes += L.tgs_rasterize_tile_order_bytes(tile_offsets.size(0), tile_offsets.size(2), tile_offsets.size(1))
    return torch.empty((nbytes,), dtype=torch.uint8, device=dev)


def _zero_rows_args(zero_rows, N, select):
    """(dst ptr, mask ptr, row bytes, selection) for the fused row fill, or nulls.
    ``zero_rows`` = (dst, mask) or (dst, mask, selection): the rows with mask value
    ``selection & 0xff`` are filled, of those only the rows r with bit (r & 7) of
    ``selection >> 8`` set when that byte is nonzero; ``select`` is the default."""
    if zero_rows is None:
        return None, None, 0, 0
    dst, mask = zero_rows[0], zero_rows[1]
    sel = int(zero_rows[2]) if len(zero_rows) > 2 else select
    if FILL_STEPS and not (sel >> 16) & 0xFF:
        sel |= (int(FILL_STEPS) & 0xFF) << 16
    _check_input(dst, "zero_rows dst")
    _check_input(mask, "zero_rows mask")
    assert dst.shape[0] == N and mask.numel() >= N and mask.dtype == torch.uint8
    return _p(dst), _p(mask), dst[0].numel() * dst.element_size(), sel


def _tex_dims(textured, textures, mip_log_res):
    """(tex_h, tex_w, dtype flag): the texture grid, or for the mip2 filter the finest
    level of the pyramid [N, (4^(L+1) - 1) / 3, 4] (L = mip_log_res)."""
    if not textured:
        return 0, 0, 0
    if mip_log_res is None:
        return textures.size(1), textures.size(2), 0
    L = int(mip_log_res)
    if textures.dim() != 3 or textures.size(1) != mipmap_texels(L) or textures.size(2) != 4:
        raise RuntimeError(f"mip2 textures must be [N, {mipmap_texels(L)}, 4] (a pyramid over 2^{L} textures)")
    return 1 << L, 1 << L, TEX_MIP2


def _packed_args(means2d, gaussian_ids, textures, n_gaussians):
    """(gaussian_ids or None, nnz, N) for the grid ([C, N, ...]) or packed ([nnz, ...]) layout."""
    if means2d.dim() == 3:
        return None, 0, means2d.size(1)
    if gaussian_ids is None:
        # the reference indexes the textures with the packed entry index here (Q3); this build
        # needs the Gaussian of each entry
        raise NotImplementedError("rasterize_to_pixels: the packed layout needs gaussian_ids")
    _check_input(gaussian_ids, "gaussian_ids")
    if gaussian_ids.dtype != torch.int64 or gaussian_ids.numel() != means2d.size(0):
        raise RuntimeError("gaussian_ids must be int64 [nnz]")
    if textures is not None:
        N = textures.size(0)
    elif n_gaussians is not None:
        N = int(n_gaussians)
    else:
        N = int(gaussian_ids.max().item()) + 1 if gaussian_ids.numel() else 1
    if gaussian_ids.numel():
        # an id outside [0, N) would address texture rows and counters past their ends
        lo, hi = torch.stack(torch.aminmax(gaussian_ids)).tolist()
        if lo < 0 or hi >= N:
            raise RuntimeError(f"gaussian_ids must lie in [0, {N}) (one per texture row / Gaussian); "
                               f"they span [{lo}, {hi}]")
    if gaussian_ids.numel() == 0:
        # no entries: the C ABI tells the packed layout by a non-null gaussian_ids, and an
        # empty tensor has none; the one placeholder id is never read (nnz = 0)
        gaussian_ids = torch.zeros(1, dtype=torch.int64, device=means2d.device)
    return gaussian_ids, means2d.size(0), N


def _raster_fwd(textured, means2d, ray_transforms, colors, opacities, textures, trx, try_, normals, backgrounds,
                masks, image_width, image_height, tile_size, tile_offsets, flatten_ids, gs_contrib_threshold,
                g_weight, zero_rows=None, live_entries=None, tex_counts=None, gaussian_ids=None, n_gaussians=None,
                mip_log_res=None, noalloc=False, pair_samples=None):
    """noalloc: call tgs_rasterize_fwd_noalloc, whose workspace is the library's static
    per-GPU scratch (the outputs are still allocated here)."""
    ts = dict(means2d=means2d, ray_transforms=ray_transforms, colors=colors, opacities=opacities,
              normals=normals, tile_offsets=tile_offsets, flatten_ids=flatten_ids)
    if textured:
        ts["textures"] = textures
    for n, t in ts.items():
        _check_input(t, n)
    if backgrounds is not None:
        _check_input(backgrounds, "backgrounds")
    if masks is not None:
        _check_input(masks, "masks")
    dt = _dtype_code(means2d, "means2d")
    _same(means2d.dtype, ray_transforms=ray_transforms, colors=colors, opacities=opacities, normals=normals,
          backgrounds=backgrounds)
    tex_flag = _tex_dtype_flag(textured, textures, means2d)
    C = tile_offsets.size(0)
    gids, nnz, N = _packed_args(means2d, gaussian_ids, textures if textured else None, n_gaussians)
    K = colors.size(-1)
    th_t, tw_t = tile_offsets.size(1), tile_offsets.size(2)
    H, W = image_height, image_width
    dev, fdt = means2d.device, means2d.dtype
    tex_h, tex_w, mip_flag = _tex_dims(textured, textures, mip_log_res)
    tex_flag |= mip_flag
    with torch.cuda.device(dev):
        renders = torch.empty((C, H, W, K), dtype=fdt, device=dev)
        alphas = torch.empty((C, H, W, 1), dtype=fdt, device=dev)
        last_ids = torch.empty((C, H, W), dtype=torch.int32, device=dev)
        median_ids = torch.empty((C, H, W), dtype=torch.int32, device=dev)
        rnormals = torch.empty((C, H, W, 3), dtype=fdt, device=dev)
        rdistort = torch.empty((C, H, W, 1), dtype=fdt, device=dev)
        rmedian = torch.empty((C, H, W, 1), dtype=fdt, device=dev)
        csum = torch.zeros((N,), dtype=fdt, device=dev)
        ccnt = torch.zeros((N,), dtype=fdt, device=dev)
        accd = torch.empty((C, H, W), dtype=fdt, device=dev)
        trans = torch.empty((C, H, W), dtype=fdt, device=dev)
        if noalloc:
            ws = None
        else:
            ws = (_raster_workspace(dt, 1, nnz, dev, tile_offsets) if gids is not None
                  else _raster_workspace(dt, C, N, dev, tile_offsets))
        zd, zm, zb, zs = _zero_rows_args(zero_rows, N, 1)
        entry = lib().tgs_rasterize_fwd_noalloc if noalloc else lib().tgs_rasterize_fwd
        ws_args = () if noalloc else (_p(ws), ws.numel())
        with _timed("raster_fwd", means2d):
            check(entry(
              dt | tex_flag | (0 if FWD_CULL else FWD_NO_CULL), C, N, K, _p(means2d), _p(ray_transforms), _p(colors), _p(opacities), _p(normals),
              _p(textures) if textured else None, tex_h, tex_w, float(trx), float(try_), _p(backgrounds),
              _p(masks.view(torch.uint8)) if masks is not None else None, _p(gids), nnz, W, H, tile_size, tw_t, th_t,
              _p(tile_offsets), _p(flatten_ids), flatten_ids.numel(), float(gs_contrib_threshold), float(g_weight),
              _p(renders), _p(alphas), _p(rnormals), _p(rdistort), _p(rmedian), _p(last_ids), _p(median_ids),
              _p(csum), _p(ccnt), _p(accd), _p(trans), _live_arg(live_entries, flatten_ids, tile_size),
              _samples_arg(pair_samples, live_entries), _counts_arg(tex_counts, textures.shape[0] if textured else 0),
              zd, zm, zb, zs, *ws_args,
              _stream(means2d)),
              "rasterize_to_pixels_fwd")
    return renders, alphas, rnormals, rdistort, rmedian, last_ids, median_ids, csum, ccnt, accd, trans


def rasterize_to_pixels_fwd_textured_gaussians(means2d, ray_transforms, colors, opacities, textures,
                                               texture_range_x, texture_range_y, normals, backgrounds, masks,
                                               image_width, image_height, tile_size, tile_offsets, flatten_ids,
                                               gs_contrib_threshold, g_weight):
    return _raster_fwd(True, means2d, ray_transforms, colors, opacities, textures, texture_range_x,
                       texture_range_y, normals, backgrounds, masks, image_width, image_height, tile_size,
                       tile_offsets, flatten_ids, gs_contrib_threshold, g_weight)[:9]


def live_entries_buffer(flatten_ids: Tensor, tile_size: int) -> Tensor:
    """int64 [waves per tile, n_isects] for the forward to record, per wavefront and list
    entry, the mask of the wave's lanes that composited the entry (no reference
    counterpart)."""
    waves = (tile_size * tile_size + 63) // 64
    return torch.empty((waves, flatten_ids.numel()), dtype=torch.int64, device=flatten_ids.device)


def _live_arg(live_entries, flatten_ids, tile_size):
    if live_entries is None:
        return None
    _check_input(live_entries, "live_entries")
    waves = (tile_size * tile_size + 63) // 64
    if live_entries.dtype != torch.int64 or live_entries.numel() != waves * flatten_ids.numel():
        raise RuntimeError(f"live_entries must be int64 [{waves}, {flatten_ids.numel()}]")
    return _p(live_entries)


def pair_samples_bytes(flatten_ids: Tensor, tile_size: int) -> int:
    """Bytes of ``pair_samples_buffer`` (16 per lane of every wavefront and list entry)."""
    waves = (tile_size * tile_size + 63) // 64
    return waves * flatten_ids.numel() * 64 * 16


def pair_samples_buffer(flatten_ids: Tensor, tile_size: int) -> Tensor:
    """float32 [waves per tile, n_isects, 64, 4]: the forward stores the texture sample of
    each composited in-range pair (RGB, alpha) for the backward, which then reads no
    texels (no reference counterpart).  Only the slots of composited pairs are written."""
    waves = (tile_size * tile_size + 63) // 64
    return torch.empty((waves, flatten_ids.numel(), 64, 4), dtype=torch.float32, device=flatten_ids.device)


def _samples_arg(pair_samples, live_entries):
    if pair_samples is None:
        return None
    _check_input(pair_samples, "pair_samples")
    if live_entries is None or pair_samples.dtype != torch.float32 or \
            pair_samples.numel() != live_entries.numel() * 64 * 4:
        raise RuntimeError("pair_samples must be float32 [waves, n_isects, 64, 4] beside live_entries")
    return _p(pair_samples)


def tex_counts_buffer(textures: Tensor) -> Tensor:
    """int32 [N] zeros: the forward counts per texture row the pairs that will add
    texture gradient (no reference counterpart)."""
    return torch.zeros(textures.shape[0], dtype=torch.int32, device=textures.device)


def _counts_arg(tex_counts, N):
    if tex_counts is None:
        return None
    _check_input(tex_counts, "tex_counts")
    if tex_counts.dtype != torch.int32 or tex_counts.numel() != N:
        raise RuntimeError(f"tex_counts must be int32 [{N}]")
    return _p(tex_counts)


class TexRecordPlan:
    """Sizing of the texture-gradient records, started right after the forward: the
    inclusive prefix sum of the forward's per-row pair counts, and its total copied to
    pinned host memory asynchronously, so the backward reads it without waiting for
    the device (no reference counterpart)."""

    def __init__(self, tex_counts: Tensor):
        N = tex_counts.numel()
        dev = tex_counts.device
        self.counts = tex_counts
        self.incl = torch.empty(N, dtype=torch.int64, device=dev)
        temp = torch.empty(max(1, lib().tgs_isect_count_temp_bytes(N)), dtype=torch.uint8, device=dev)
        check(lib().tgs_prefix_sum_i32(N, _p(tex_counts), _p(self.incl), _p(temp), temp.numel(), _stream(tex_counts)),
              "prefix_sum")
        self.total_host = torch.zeros(1, dtype=torch.int64, pin_memory=True)
        if N > 0:
            self.total_host.copy_(self.incl[-1:], non_blocking=True)
        self.done = torch.cuda.Event()
        self.done.record(torch.cuda.current_stream(dev))
        self._temp = temp  # kept until the copy has run

    def total(self) -> int:
        wait_event(self.done)
        return int(self.total_host[0])


def _tex_record_args(tex_counts, dt, N, dev):
    """(counts, inclusive offsets, cursor, records) for the backward's record path;
    ``tex_counts`` is the forward's count tensor or a ``TexRecordPlan`` of it."""
    if tex_counts is None:
        return None, None, None, None, None
    plan = tex_counts if isinstance(tex_counts, TexRecordPlan) else TexRecordPlan(tex_counts)
    _counts_arg(plan.counts, N)
    total = plan.total()
    recs = torch.empty(max(1, total) * lib().tgs_tex_record_bytes(dt), dtype=torch.uint8, device=dev)
    cursor = torch.zeros(N, dtype=torch.int32, device=dev)
    return plan.counts, plan.incl, cursor, recs, None


def texture_adam_rows(records, visibility, tex_hi, tex_lo, exp_avg, exp_avg_sq, lr, b1, b2, eps, seed):
    """Selective Adam step of bfloat16 textures [N, h, w, 4] from the backward's
    texture-gradient records ``(counts, inclusive offsets, cursor, records)``: the
    float32 master is (tex_hi, tex_lo), tex_hi the render copy; moments bfloat16.
    ``visibility`` bool/uint8 [N].  All in place."""
    counts, incl, cursor, recs = records
    for n, t in dict(counts=counts, incl=incl, cursor=cursor, recs=recs, visibility=visibility, tex_hi=tex_hi,
                     tex_lo=tex_lo, exp_avg=exp_avg, exp_avg_sq=exp_avg_sq).items():
        _check_input(t, n)
    N, th, tw = tex_hi.size(0), tex_hi.size(1), tex_hi.size(2)
    for n, t in dict(tex_lo=tex_lo, exp_avg=exp_avg, exp_avg_sq=exp_avg_sq).items():
        if t.shape != tex_hi.shape or t.element_size() != 2:
            raise RuntimeError(f"{n} must be 16-bit with the shape of tex_hi")
    if visibility.numel() != N or visibility.element_size() != 1:
        raise RuntimeError("visibility must be bool/uint8 [N]")
    with torch.cuda.device(tex_hi.device), _timed("texture_adam", tex_hi):
        check(lib().tgs_texture_adam_rows(N, th, tw, _p(counts), _p(incl), _p(cursor), _p(recs),
                                          _p(visibility.view(torch.uint8)), _p(tex_hi), _p(tex_lo), _p(exp_avg),
                                          _p(exp_avg_sq), float(lr), float(b1), float(b2), float(eps),
                                          int(seed) & 0xFFFFFFFF, _stream(tex_hi)), "texture_adam_rows")


def rasterize_to_pixels_fwd_textured_gaussians_full(*args, zero_rows=None, live_entries=None, tex_counts=None,
                                                    gaussian_ids=None, pair_samples=None):
    """As ``rasterize_to_pixels_fwd_textured_gaussians`` plus a 10th output, the
    per-pixel distortion accumulator sum_i w_i colors[i, K-1] [C,H,W], and an 11th,
    the final transmittance T [C,H,W] (for the backward's ``render_trans``).
    ``zero_rows`` = (buffer [N, ...], uint8 mask [N]): rows with mask 1 are
    zero-filled by the forward kernel (see ``mark_rows``).  ``live_entries``
    (from ``live_entries_buffer``) receives the entries each wavefront composited,
    for the backward to visit only those.  ``tex_counts`` (from ``tex_counts_buffer``)
    receives per texture row the pairs that will add texture gradient, for the
    backward's record-and-gather path.  ``pair_samples`` (from ``pair_samples_buffer``,
    with ``live_entries``) receives the texture samples of the composited pairs, for a
    backward under the ``bilinear`` rule to read instead of the texels."""
    return _raster_fwd(True, *args, zero_rows=zero_rows, live_entries=live_entries, tex_counts=tex_counts,
                       gaussian_ids=gaussian_ids, pair_samples=pair_samples)


def rasterize_to_pixels_fwd_2dgs(means2d, ray_transforms, colors, opacities, normals, backgrounds, masks,
                                 image_width, image_height, tile_size, tile_offsets, flatten_ids,
                                 gs_contrib_threshold):
    return rasterize_to_pixels_fwd_2dgs_full(means2d, ray_transforms, colors, opacities, normals, backgrounds,
                                             masks, image_width, image_height, tile_size, tile_offsets,
                                             flatten_ids, gs_contrib_threshold)[:9]


def rasterize_to_pixels_fwd_2dgs_full(means2d, ray_transforms, colors, opacities, normals, backgrounds, masks,
                                      image_width, image_height, tile_size, tile_offsets, flatten_ids,
                                      gs_contrib_threshold, live_entries=None, gaussian_ids=None, n_gaussians=None):
    return _raster_fwd(False, means2d, ray_transforms, colors, opacities, None, 1.0, 1.0, normals, backgrounds,
                       masks, image_width, image_height, tile_size, tile_offsets, flatten_ids,
                       gs_contrib_threshold, 1.0, live_entries=live_entries, gaussian_ids=gaussian_ids,
                       n_gaussians=n_gaussians)


def _raster_bwd(textured, uv_mode, means2d, ray_transforms, colors, opacities, textures, trx, try_, normals,
                densify, backgrounds, masks, image_width, image_height, tile_size, tile_offsets, flatten_ids,
                g_weight, render_colors, render_alphas, last_ids, median_ids, v_render_colors, v_render_alphas,
                v_render_normals, v_render_distort, v_render_median, absgrad, render_accum_depth=None,
                v_textures_out=None, zero_rows=None, live_entries=None, tex_counts=None, gaussian_ids=None,
                n_gaussians=None, mip_log_res=None, render_trans=None, record_rows=False, pair_samples=None):
    ts = dict(means2d=means2d, ray_transforms=ray_transforms, colors=colors, opacities=opacities,
              normals=normals, densify=densify, tile_offsets=tile_offsets, flatten_ids=flatten_ids,
              render_colors=render_colors, render_alphas=render_alphas, last_ids=last_ids,
              median_ids=median_ids, v_render_colors=v_render_colors, v_render_alphas=v_render_alphas,
              v_render_normals=v_render_normals, v_render_distort=v_render_distort,
              v_render_median=v_render_median)
    if textured:
        ts["textures"] = textures
    for n, t in ts.items():
        _check_input(t, n)
    if backgrounds is not None:
        _check_input(backgrounds, "backgrounds")
    if masks is not None:
        _check_input(masks, "masks")
    dt = _dtype_code(means2d, "means2d")
    tex_flag = _tex_dtype_flag(textured, textures, means2d)
    C = tile_offsets.size(0)
    gids, nnz, N = _packed_args(means2d, gaussian_ids, textures if textured else None, n_gaussians)
    K = colors.size(-1)
    th_t, tw_t = tile_offsets.size(1), tile_offsets.size(2)
    dev = means2d.device
    records_only = textured and v_textures_out is False
    tex_h, tex_w, mip_flag = _tex_dims(textured, textures, mip_log_res)
    tex_flag |= mip_flag
    if record_rows:
        tex_flag |= BWD_RECORD_ROWS
    with torch.cuda.device(dev):
        v_means2d = torch.zeros_like(means2d)
        v_rt = torch.zeros_like(ray_transforms)
        v_colors = torch.zeros_like(colors)
        if records_only:
            # the texture gradient leaves as records only (for texture_adam_rows)
            if tex_counts is None:
                raise RuntimeError("rasterize_to_pixels_bwd: v_textures_out=False needs the forward's tex_counts")
            v_textures = None
        elif textured and v_textures_out is not None:
            # caller-provided, already zero-filled (see the textured autograd Function)
            assert v_textures_out.shape == textures.shape and v_textures_out.dtype == means2d.dtype
            v_textures = v_textures_out
        else:
            # float gradient also for bfloat16 textures
            v_textures = torch.zeros(textures.shape, dtype=means2d.dtype, device=dev) if textured else None
        v_normals = torch.zeros_like(normals)
        v_opacities = torch.zeros_like(opacities)
        v_means2d_abs = torch.zeros_like(means2d) if absgrad else None
        v_densify = torch.zeros_like(densify)
        ws = (_raster_workspace(dt, 1, nnz, dev, tile_offsets) if gids is not None
              else _raster_workspace(dt, C, N, dev, tile_offsets))
        zd, zm, zb, zs = _zero_rows_args(zero_rows, N, 0)
        if render_trans is not None:
            _check_input(render_trans, "render_trans")
        rc_, ri_, rcur_, rrec_, _tmp = _tex_record_args(tex_counts if textured else None, dt, textures.shape[0]
                                                        if textured else 0, dev)
        with _timed("raster_bwd", means2d):
            check(lib().tgs_rasterize_bwd(
              dt | tex_flag, C, N, K, _p(means2d), _p(ray_transforms), _p(colors), _p(opacities), _p(normals),
              _p(textures) if textured else None, tex_h, tex_w, float(trx), float(try_), _p(backgrounds),
              _p(masks.view(torch.uint8)) if masks is not None else None, _p(gids), nnz, image_width,
              image_height, tile_size, tw_t, th_t, _p(tile_offsets), _p(flatten_ids), flatten_ids.numel(),
              float(g_weight),
              _p(render_colors), _p(render_alphas), _p(last_ids), _p(median_ids), _p(render_accum_depth),
              _p(render_trans), _live_arg(live_entries, flatten_ids, tile_size),
              _samples_arg(pair_samples, live_entries), _p(rc_), _p(ri_), _p(rcur_),
              _p(rrec_),
              _p(v_render_colors),
              _p(v_render_alphas), _p(v_render_normals), _p(v_render_distort), _p(v_render_median), uv_mode,
              _p(v_means2d_abs), _p(v_means2d), _p(v_rt), _p(v_colors), _p(v_opacities), _p(v_textures),
              _p(v_normals), _p(v_densify) if densify.shape == means2d.shape else None,
              zd, zm, zb, zs, _p(ws), ws.numel(), _stream(means2d)), "rasterize_to_pixels_bwd")
    if records_only:
        # (counts, inclusive offsets, cursor, records) in place of v_textures
        return v_means2d_abs, v_means2d, v_rt, v_colors, v_opacities, (rc_, ri_, rcur_, rrec_), v_normals, v_densify
    if textured:
        return v_means2d_abs, v_means2d, v_rt, v_colors, v_opacities, v_textures, v_normals, v_densify
    return v_means2d_abs, v_means2d, v_rt, v_colors, v_opacities, v_normals, v_densify


# Keyword options of the textured backward ops (no reference counterparts):
# render_accum_depth, render_trans, v_textures_out, zero_rows, live_entries,
# pair_samples, tex_counts, gaussian_ids (see _raster_bwd and include/tgs_hip.h).
def rasterize_to_pixels_bwd_textured_gaussians(*args, **opts):
    """Reference ``bilinear`` backward rule (no d(texture sample)/ds)."""
    return _raster_bwd(True, UV_NONE, *args, **opts)


def rasterize_to_pixels_bwd2_textured_gaussians(*args, **opts):
    """Reference ``bilinear_bwd2`` rule (du/ds = (res-1)/range)."""
    return _raster_bwd(True, UV_DOUBLE, *args, **opts)


def rasterize_to_pixels_bwd_exact_textured_gaussians(*args, **opts):
    """True derivative: du/ds = (res-1)/(2 range), and the forward's own
    distortion accumulator when ``render_accum_depth`` is given (no reference
    counterpart)."""
    return _raster_bwd(True, UV_EXACT, *args, **opts)


def rasterize_to_pixels_bwd_2dgs(means2d, ray_transforms, colors, opacities, normals, densify, backgrounds, masks,
                                 image_width, image_height, tile_size, tile_offsets, flatten_ids, render_colors,
                                 render_alphas, last_ids, median_ids, v_render_colors, v_render_alphas,
                                 v_render_normals, v_render_distort, v_render_median, absgrad,
                                 render_accum_depth=None, live_entries=None, gaussian_ids=None, n_gaussians=None,
                                 render_trans=None):
    return _raster_bwd(False, UV_NONE, means2d, ray_transforms, colors, opacities, None, 1.0, 1.0, normals,
                       densify, backgrounds, masks, image_width, image_height, tile_size, tile_offsets,
                       flatten_ids, 1.0, render_colors, render_alphas, last_ids, median_ids, v_render_colors,
                       v_render_alphas, v_render_normals, v_render_distort, v_render_median, absgrad,
                       render_accum_depth=render_accum_depth, live_entries=live_entries, gaussian_ids=gaussian_ids,
                       n_gaussians=n_gaussians, render_trans=render_trans)


# ---------------------------------------------------------------- mipmaps
def mipmap_texels(log_texture_res: int) -> int:
    """Texels of the pyramid over 2^L x 2^L textures: (4^(L+1) - 1) / 3."""
    return ((1 << (2 * (int(log_texture_res) + 1))) - 1) // 3


def generate_mipmap_fwd(textures, log_texture_res, log_reduce=4, tile_size=16):
    """[N, R, R, C] (R = 2^L) -> [N, (4^(L+1) - 1) / 3, C], coarsest level first
    (``generate_mipmap_fwd``, ext.cpp:188; ``log_reduce`` and ``tile_size`` only
    shape the reference's launches)."""
    _check_input(textures, "textures")
    dt = _dtype_code(textures, "textures")
    N, R, R2, C = textures.shape
    L = int(log_texture_res)
    if R != R2 or R != 1 << L:
        raise RuntimeError(f"generate_mipmap: textures must be [N, {1 << L}, {1 << L}, C]")
    if C > 5:
        raise RuntimeError(f"Unsupported number of channels: {C}")
    with torch.cuda.device(textures.device):
        mip = torch.empty((N, mipmap_texels(L), C), dtype=textures.dtype, device=textures.device)
        check(lib().tgs_generate_mipmap_fwd(dt, N, C, L, _p(textures), _p(mip), _stream(textures)),
              "generate_mipmap_fwd")
    return mip


def generate_mipmap_bwd(N, channels, log_texture_res, log_reduce, tile_size, v_mip_textures):
    """Adjoint of ``generate_mipmap_fwd`` (ext.cpp:192): -> v_textures [N, R, R, C]."""
    _check_input(v_mip_textures, "v_mip_textures")
    dt = _dtype_code(v_mip_textures, "v_mip_textures")
    L = int(log_texture_res)
    R = 1 << L
    with torch.cuda.device(v_mip_textures.device):
        v = torch.empty((N, R, R, channels), dtype=v_mip_textures.dtype, device=v_mip_textures.device)
        check(lib().tgs_generate_mipmap_bwd(dt, N, channels, L, _p(v_mip_textures), _p(v),
                                            _stream(v_mip_textures)), "generate_mipmap_bwd")
    return v


def rasterize_to_pixels_fwd_mip2_textured_gaussians(means2d, ray_transforms, colors, opacities, textures,
                                                    log_texture_res, texture_range_x, texture_range_y, normals,
                                                    backgrounds, masks, image_width, image_height, tile_size,
                                                    tile_offsets, flatten_ids, gs_contrib_threshold, g_weight,
                                                    live_entries=None, zero_rows=None):
    """The ``mipmapped2`` forward (ext.cpp:136-137) over the pyramid ``textures``
    [N, (4^(L+1) - 1) / 3, 4]; returns the 9-tuple of the bilinear forward."""
    return rasterize_to_pixels_fwd_mip2_textured_gaussians_full(
        means2d, ray_transforms, colors, opacities, textures, log_texture_res, texture_range_x, texture_range_y,
        normals, backgrounds, masks, image_width, image_height, tile_size, tile_offsets, flatten_ids,
        gs_contrib_threshold, g_weight, live_entries=live_entries, zero_rows=zero_rows)[:9]


def rasterize_to_pixels_fwd_mip2_textured_gaussians_full(means2d, ray_transforms, colors, opacities, textures,
                                                         log_texture_res, texture_range_x, texture_range_y, normals,
                                                         backgrounds, masks, image_width, image_height, tile_size,
                                                         tile_offsets, flatten_ids, gs_contrib_threshold, g_weight,
                                                         live_entries=None, zero_rows=None, gaussian_ids=None):
    """As above plus the distortion accumulator and the final transmittance (11 outputs,
    as ``rasterize_to_pixels_fwd_textured_gaussians_full``); ``gaussian_ids``: the packed
    layout."""
    return _raster_fwd(True, means2d, ray_transforms, colors, opacities, textures, texture_range_x,
                       texture_range_y, normals, backgrounds, masks, image_width, image_height, tile_size,
                       tile_offsets, flatten_ids, gs_contrib_threshold, g_weight, live_entries=live_entries,
                       zero_rows=zero_rows, mip_log_res=log_texture_res, gaussian_ids=gaussian_ids)


def rasterize_to_pixels_bwd_mip2_textured_gaussians(means2d, ray_transforms, colors, opacities, textures,
                                                    log_texture_res, texture_range_x, texture_range_y, normals,
                                                    densify, backgrounds, masks, image_width, image_height,
                                                    tile_size, tile_offsets, flatten_ids, g_weight, render_colors,
                                                    render_alphas, last_ids, median_ids, v_render_colors,
                                                    v_render_alphas, v_render_normals, v_render_distort,
                                                    v_render_median, absgrad, live_entries=None, v_textures_out=None,
                                                    zero_rows=None, render_trans=None, gaussian_ids=None):
    """The ``mipmapped2`` backward (ext.cpp:139-140): the reference's rule (no
    d(sample)/d(position) term); v_textures has the pyramid's shape."""
    return _raster_bwd(True, UV_NONE, means2d, ray_transforms, colors, opacities, textures, texture_range_x,
                       texture_range_y, normals, densify, backgrounds, masks, image_width, image_height, tile_size,
                       tile_offsets, flatten_ids, g_weight, render_colors, render_alphas, last_ids, median_ids,
                       v_render_colors, v_render_alphas, v_render_normals, v_render_distort, v_render_median,
                       absgrad, live_entries=live_entries, v_textures_out=v_textures_out, zero_rows=zero_rows,
                       mip_log_res=log_texture_res, render_trans=render_trans,
                       gaussian_ids=gaussian_ids)


# ------------------------------------------------------------------- SSIM
def ssim_fwd(x, y):
    """Fused SSIM of channels-last images [B, H, W, C] (the trainer's torchmetrics
    settings): returns (ssim, work) with ``work`` kept for ``ssim_bwd``."""
    for n, t in dict(x=x, y=y).items():
        _check_input(t, n)
    dt = _dtype_code(x, "x")
    _same(x.dtype, y=y)
    if x.dim() != 4 or x.shape != y.shape:
        raise RuntimeError("ssim: x and y must be [B, H, W, C] of one shape")
    B, H, W, C = x.shape
    L = lib()
    with torch.cuda.device(x.device):
        work = torch.empty(max(1, L.tgs_ssim_work_bytes(dt, B, H, W, C)), dtype=torch.uint8, device=x.device)
        part = torch.empty(max(1, L.tgs_ssim_partials(B, H, W, C)), dtype=x.dtype, device=x.device)
        check(L.tgs_ssim_fwd(dt, B, H, W, C, _p(x), _p(y), _p(work), _p(part), _stream(x)), "ssim_fwd")
    M = B * C * (H - 10) * (W - 10)
    return part.sum() / M, work


def ssim_bwd(x, y, work):
    """d(SSIM)/dx for the ``work`` of ``ssim_fwd(x, y)`` (y held constant)."""
    dt = _dtype_code(x, "x")
    B, H, W, C = x.shape
    M = B * C * (H - 10) * (W - 10)
    with torch.cuda.device(x.device):
        dx = torch.empty_like(x)
        check(lib().tgs_ssim_bwd(dt, B, H, W, C, _p(x), _p(y), _p(work), 1.0 / M, _p(dx), _stream(x)), "ssim_bwd")
    return dx


# ------------------------------------------------------------------ Adam
def selective_adam_update(param, param_grad, exp_avg, exp_avg_sq, tiles_touched, lr, b1, b2, eps, N, M):
    for n, t in dict(param=param, param_grad=param_grad, exp_avg=exp_avg, exp_avg_sq=exp_avg_sq,
                     tiles_touched=tiles_touched).items():
        _check_input(t, n)
    if param.dtype != torch.float32:
        raise RuntimeError("selective_adam_update: param must be float32")
    with torch.cuda.device(param.device):
        check(lib().tgs_selective_adam_update(N, M, _p(param), _p(param_grad), _p(exp_avg), _p(exp_avg_sq),
                                              _p(tiles_touched.view(torch.uint8)), float(lr), float(b1),
                                              float(b2), float(eps), _stream(param)), "selective_adam_update")


def set_texel_marks(marks: Optional[Tensor]) -> None:
    """Measurement only: while set (a zero-filled uint8 device tensor of at least
    N * tex_h * tex_w bytes), the float-texture forward marks every texel it reads;
    None clears it."""
    if marks is None:
        check(lib().tgs_set_texel_marks(None, 0), "set_texel_marks")
        return
    _check_input(marks, "marks")
    if marks.dtype != torch.uint8:
        raise RuntimeError("marks must be uint8")
    check(lib().tgs_set_texel_marks(_p(marks), marks.numel()), "set_texel_marks")


def lds_atomic_rates(device=None) -> dict:
    """LDS atomic-add throughput of the GPU, lanes per clock per CU, by access kind
    (measurement hook, no reference counterpart; see tgs_lds_atomic_rates)."""
    import ctypes
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    out = (ctypes.c_double * 4)()
    with torch.cuda.device(dev):
        check(lib().tgs_lds_atomic_rates(ctypes.cast(out, ctypes.c_void_p), 4,
                                         torch.cuda.current_stream(dev).cuda_stream), "lds_atomic_rates")
    return {"atomic_add_f32": out[0], "atomic_add_f64": out[1], "atomic_add_u32": out[2],
            "read_add_write_f32": out[3]}


def zero_fill_(t: Tensor) -> Tensor:
    """Zero-fills a contiguous device tensor in place with one native launch on
    the current stream (no reference counterpart; used for the texture
    gradient, which the backward accumulates into)."""
    _check_input(t, "t")
    with torch.cuda.device(t.device), _timed("grad_fill", t):
        check(lib().tgs_zero_fill(_p(t), t.numel() * t.element_size(), _stream(t)), "zero_fill")
    return t


def mark_rows(flatten_ids: Tensor, N: int) -> Tensor:
    """uint8 mask [N]: 1 for each Gaussian id % N present in ``flatten_ids``
    (the texture rows the backward can touch), on the current stream."""
    _check_input(flatten_ids, "flatten_ids")
    mask = torch.empty((N,), dtype=torch.uint8, device=flatten_ids.device)
    with torch.cuda.device(flatten_ids.device):
        check(lib().tgs_mark_rows(flatten_ids.numel(), _p(flatten_ids), N, _p(mask), _stream(flatten_ids)),
              "mark_rows")
    return mask


def zero_fill_rows_(t: Tensor, mask: Tensor, select: int) -> Tensor:
    """Zero-fills the rows t[n] (t contiguous, first dim N) with mask[n] == select."""
    _check_input(t, "t")
    _check_input(mask, "mask")
    row_bytes = t[0].numel() * t.element_size() if t.shape[0] > 0 else 0
    if t.shape[0] == 0:
        return t
    with torch.cuda.device(t.device), _timed("grad_fill", t):
        check(lib().tgs_zero_fill_rows(_p(t), row_bytes, t.shape[0], _p(mask), int(select), _stream(t)),
              "zero_fill_rows")
    return t


# ------------------------------------------------- neural-texture decoder
EPI_NONE, EPI_RELU, EPI_SIGMOID, EPI_RELU_MASK, EPI_ATOMIC = 0, 1, 2, 3, 4
GEMM_A_BF16, GEMM_B_BF16 = 0x10, 0x20


def _f32_2d(t: Tensor, name: str) -> None:
    _check_input(t, name)
    if t.dtype != torch.float32 or t.dim() != 2:
        raise RuntimeError(f"{name} must be a 2-D float32 tensor")


def gemm_bf16(a: Tensor, b: Tensor, trans_a: bool = False, trans_b: bool = False, epilogue: int = EPI_NONE,
              bias: Optional[Tensor] = None, mask: Optional[Tensor] = None, out: Optional[Tensor] = None,
              k_splits: int = 1) -> Tensor:
    """C = epilogue(op(a) @ op(b) + bias) on the matrix cores, bfloat16 operands and
    float32 accumulation (op = transpose when trans_*).  ``a`` and ``b`` are float32
    (rounded to nearest even when staged) or bfloat16 (staged as is).  With
    ``EPI_ATOMIC`` the product is added into ``out`` (which must be given),
    optionally with K split over ``k_splits`` workgroup slices."""
    flags = 0
    for t, name, flag in ((a, "a", GEMM_A_BF16), (b, "b", GEMM_B_BF16)):
        _check_input(t, name)
        if t.dim() != 2 or t.dtype not in (torch.float32, torch.bfloat16) or t.stride(1) != 1:
            raise RuntimeError(f"{name} must be a 2-D float32 or bfloat16 tensor with unit column stride")
        if t.dtype == torch.bfloat16:
            flags |= flag
    M, K = (a.shape[1], a.shape[0]) if trans_a else (a.shape[0], a.shape[1])
    Kb, N = (b.shape[1], b.shape[0]) if trans_b else (b.shape[0], b.shape[1])
    if K != Kb:
        raise RuntimeError(f"gemm_bf16: inner sizes differ ({K} vs {Kb})")
    if bias is not None:
        _check_input(bias, "bias")
        assert bias.dtype == torch.float32 and bias.numel() == N
    if mask is not None:
        _f32_2d(mask, "mask")
        assert mask.shape == (M, N)
    if out is None:
        if epilogue == EPI_ATOMIC:
            raise RuntimeError("gemm_bf16: the atomic epilogue accumulates into a given out")
        out = torch.empty((M, N), dtype=torch.float32, device=a.device)
    else:
        _f32_2d(out, "out")
        assert out.shape == (M, N)
    with torch.cuda.device(a.device):
        check(lib().tgs_gemm_bf16(int(trans_a), int(trans_b), int(epilogue) | flags, M, N, K, _p(a), a.stride(0), _p(b),
                                  b.stride(0), _p(bias), _p(out), out.stride(0), _p(mask),
                                  mask.stride(0) if mask is not None else 0, int(k_splits), _stream(a)),
              "gemm_bf16")
    return out


def sigmoid_bwd(g_out: Tensor, out: Tensor) -> Tensor:
    _check_input(g_out, "g_out")
    _check_input(out, "out")
    g = torch.empty_like(out)
    with torch.cuda.device(out.device):
        check(lib().tgs_sigmoid_bwd(out.numel(), _p(g_out), _p(out), _p(g), _stream(out)), "sigmoid_bwd")
    return g


def sigmoid_bwd_colsum(g_out: Tensor, out: Tensor, g_dtype=torch.float32):
    """(g_out * out * (1 - out), its column sums) for 2-D float32 [M, N] (N % 4 == 0),
    in one pass; g stored as ``g_dtype`` (float32 or bfloat16, nearest even), the
    sums taken over the float32 values."""
    _f32_2d(out, "out")
    _check_input(g_out, "g_out")
    if g_out.shape != out.shape or g_out.dtype != torch.float32 or out.shape[1] % 4:
        raise RuntimeError("sigmoid_bwd_colsum: g_out and out must be float32 [M, N] with N % 4 == 0")
    if g_dtype not in (torch.float32, torch.bfloat16):
        raise RuntimeError("sigmoid_bwd_colsum: g_dtype must be float32 or bfloat16")
    g = torch.empty(out.shape, dtype=g_dtype, device=out.device)
    db = torch.zeros(out.shape[1], dtype=torch.float32, device=out.device)
    with torch.cuda.device(out.device):
        check(lib().tgs_sigmoid_bwd_colsum(out.shape[0], out.shape[1], _p(g_out), _p(out), _p(g),
                                           int(g_dtype == torch.bfloat16), _p(db), _stream(out)),
              "sigmoid_bwd_colsum")
    return g, db


def colsum_(g: Tensor, db: Tensor) -> Tensor:
    """db += g.sum(0) for a 2-D float32 g (float32 accumulation)."""
    _f32_2d(g, "g")
    _check_input(db, "db")
    with torch.cuda.device(g.device):
        check(lib().tgs_colsum(g.shape[0], g.shape[1], _p(g), g.stride(0), _p(db), _stream(g)), "colsum")
    return db


DECODER_SIZES = (32, 128, 128, 1024)


def decoder_pack(W1: Tensor, W2: Tensor, W3: Tensor, packed: Optional[Tensor] = None) -> Tensor:
    """bfloat16 MFMA fragments of the decoder weights ([in, out] float32) for
    :func:`decode_textures`; repack after every weight change."""
    i, h1, h2, o = DECODER_SIZES
    for w, shape, name in ((W1, (i, h1), "W1"), (W2, (h1, h2), "W2"), (W3, (h2, o), "W3")):
        _f32_2d(w, name)
        if tuple(w.shape) != shape:
            raise RuntimeError(f"decoder_pack: {name} must be {shape}, got {tuple(w.shape)}")
    nbytes = lib().tgs_decoder_pack_bytes()
    if packed is None:
        packed = torch.empty(nbytes // 2, dtype=torch.int16, device=W1.device)
    assert packed.numel() * packed.element_size() >= nbytes and packed.is_contiguous()
    with torch.cuda.device(W1.device):
        check(lib().tgs_decoder_pack(_p(W1), _p(W2), _p(W3), _p(packed), _stream(W1)), "decoder_pack")
    return packed


def decode_textures(z: Tensor, packed: Tensor, b1: Tensor, b2: Tensor, b3: Tensor,
                    out: Optional[Tensor] = None) -> Tensor:
    """Fused decoder inference: [n, 32] latents -> [n, 16, 16, 4] float32 textures."""
    _f32_2d(z, "z")
    if z.shape[1] != DECODER_SIZES[0]:
        raise RuntimeError(f"decode_textures: latents must be [n, {DECODER_SIZES[0]}]")
    for b, k, name in ((b1, DECODER_SIZES[1], "b1"), (b2, DECODER_SIZES[2], "b2"), (b3, DECODER_SIZES[3], "b3")):
        _check_input(b, name)
        if b.dtype != torch.float32 or b.numel() != k:
            raise RuntimeError(f"decode_textures: {name} must be {k} float32 values")
    n = z.shape[0]
    if out is None:
        out = torch.empty((n, 16, 16, 4), dtype=torch.float32, device=z.device)
    assert out.is_contiguous() and out.dtype == torch.float32 and out.numel() == n * DECODER_SIZES[3]
    with torch.cuda.device(z.device):
        check(lib().tgs_decode_textures(n, _p(z), _p(packed), _p(b1), _p(b2), _p(b3), _p(out), _stream(z)),
              "decode_textures")
    return out


class _OpTable:
    """Attribute access by op name, like the reference's ``_C`` module."""

    def __getattr__(self, name):
        fn = globals().get(name)
        if fn is None or name.startswith("_"):
            raise AttributeError(name)
        return fn


_C = _OpTable()
