"""Loader for libtgs_hip.so, the gfx950 kernels behind the C ABI declared in
``include/tgs_hip.h``.

The reference JIT-compiles a pybind11 module (``textured_gaussians/cuda/_backend.py``);
here the kernels are a plain shared library built ahead of time by
``python -m textured_gaussians_amd.build`` and bound with ctypes.  There is no
fallback: if the library is missing or cannot be loaded, every op raises.
"""
import ctypes
import os
import threading

import torch  # loads the HIP runtime the library binds to (same soname)

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_HERE, "libtgs_hip.so")

_lock = threading.Lock()
_lib = None

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_size = ctypes.c_size_t
vp = ctypes.c_void_p

# name -> (restype, argtypes), mirroring include/tgs_hip.h
_PROTOS = {
    "tgs_last_error": (ctypes.c_char_p, []),
    "tgs_abi_version": (c_int, []),
    "tgs_zero_fill": (c_int, [vp, c_size, vp]),
    "tgs_mark_rows": (c_int, [c_i64, vp, c_i64, vp, vp]),
    "tgs_zero_fill_rows": (c_int, [vp, c_size, c_i64, vp, c_int, vp]),
    "tgs_projection_2dgs_fwd": (c_int, [c_int, c_i64, c_i64, vp, vp, vp, vp, vp, c_i32, c_i32, c_dbl, c_dbl,
                                        c_dbl, vp, vp, vp, vp, vp, vp]),
    "tgs_projection_2dgs_bwd": (c_int, [c_int, c_i64, c_i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, c_int,
                                        vp, vp, vp, vp, vp]),
    "tgs_isect_count_temp_bytes": (c_size, [c_i64]),
    "tgs_isect_tiles_count": (c_int, [c_i64, c_i64, vp, vp, c_i32, c_i32, c_i32, vp, vp, vp, c_size, vp]),
    "tgs_isect_tiles_fill": (c_int, [c_i64, c_i64, vp, vp, vp, vp, c_i32, c_i32, c_i32, vp, vp, vp]),
    "tgs_sort_pairs_temp_bytes": (c_size, [c_i64]),
    "tgs_sort_pairs": (c_int, [c_i64, c_i32, vp, vp, vp, vp, vp, c_size, ctypes.POINTER(c_i32), vp]),
    "tgs_isect_offset_encode": (c_int, [c_i64, vp, c_i64, c_i32, c_i32, vp, vp]),
    "tgs_isect_tiles_fill_packed": (c_int, [c_i64, c_i64, vp, vp, vp, vp, vp, c_i32, c_i32, c_i32, vp, vp, vp]),
    "tgs_projection_2dgs_packed_count": (c_int, [c_int, c_i64, c_i64, vp, vp, vp, vp, vp, c_i32, c_i32, c_dbl,
                                                 c_dbl, c_dbl, vp, vp, vp, c_size, vp]),
    "tgs_projection_2dgs_packed_fill": (c_int, [c_int, c_i64, c_i64, vp, vp, vp, vp, vp, c_i32, c_i32, c_dbl, c_dbl,
                                                c_dbl, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "tgs_projection_2dgs_packed_bwd": (c_int, [c_int, c_i64, c_i64, c_i64, vp, vp, c_int, vp, vp, vp, vp, vp, vp,
                                               vp, vp, vp, vp, vp, c_int, vp, vp, vp, vp, vp]),
    "tgs_sh_fwd": (c_int, [c_int, c_i64, c_i32, c_i32, vp, vp, vp, vp, vp]),
    "tgs_sh_bwd": (c_int, [c_int, c_i64, c_i32, c_i32, vp, vp, vp, vp, vp, vp, vp]),
    "tgs_sh_fwd_cams": (c_int, [c_int, c_i64, c_i64, c_i32, c_i32, vp, vp, vp, vp, vp]),
    "tgs_sh_bwd_cams": (c_int, [c_int, c_i64, c_i64, c_i32, c_i32, vp, vp, vp, vp, vp, vp, vp]),
    "tgs_rasterize_fwd": (c_int, [c_int, c_i64, c_i64, c_i32, vp, vp, vp, vp, vp, vp, c_i64, c_i64, c_dbl, c_dbl,
                                  vp, vp, vp, c_i64, c_i32, c_i32, c_i32, c_i32, c_i32, vp, vp, c_i64, c_dbl, c_dbl,
                                  vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, c_size, c_int,
                                  vp, c_size, vp]),
    "tgs_rasterize_fwd_noalloc": (c_int, [c_int, c_i64, c_i64, c_i32, vp, vp, vp, vp, vp, vp, c_i64, c_i64, c_dbl,
                                          c_dbl, vp, vp, vp, c_i64, c_i32, c_i32, c_i32, c_i32, c_i32, vp, vp, c_i64,
                                          c_dbl, c_dbl, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                          vp, c_size, c_int, vp]),
    "tgs_set_raster_events": (c_int, [c_int, vp, vp]),
    "tgs_set_texel_marks": (c_int, [vp, c_i64]),
    "tgs_lds_atomic_rates": (c_int, [vp, c_i32, vp]),
    "tgs_rasterize_workspace_bytes": (c_size, [c_int, c_i64, c_i64]),
    "tgs_rasterize_tile_order_bytes": (c_size, [c_i64, c_i32, c_i32]),
    "tgs_tex_record_bytes": (c_size, [c_int]),
    "tgs_texture_gather_fast": (c_int, [c_int, c_i64, c_i64]),
    "tgs_mipmap_texels": (c_i64, [c_i32]),
    "tgs_sigmoid_bwd_colsum": (c_int, [c_i64, c_i64, vp, vp, vp, c_int, vp, vp]),
    "tgs_ssim_work_bytes": (c_size, [c_int, c_i32, c_i32, c_i32, c_i32]),
    "tgs_ssim_partials": (c_i64, [c_i32, c_i32, c_i32, c_i32]),
    "tgs_ssim_fwd": (c_int, [c_int, c_i32, c_i32, c_i32, c_i32, vp, vp, vp, vp, vp]),
    "tgs_ssim_bwd": (c_int, [c_int, c_i32, c_i32, c_i32, c_i32, vp, vp, vp, c_dbl, vp, vp]),
    "tgs_generate_mipmap_fwd": (c_int, [c_int, c_i64, c_i32, c_i32, vp, vp, vp]),
    "tgs_generate_mipmap_bwd": (c_int, [c_int, c_i64, c_i32, c_i32, vp, vp, vp]),
    "tgs_texture_adam_rows": (c_int, [c_i64, c_i32, c_i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, c_dbl, c_dbl, c_dbl,
                                      c_dbl, ctypes.c_uint32, vp]),
    "tgs_prefix_sum_i32": (c_int, [c_i64, vp, vp, vp, c_size, vp]),
    "tgs_rasterize_bwd": (c_int, [c_int, c_i64, c_i64, c_i32, vp, vp, vp, vp, vp, vp, c_i64, c_i64, c_dbl, c_dbl,
                                  vp, vp, vp, c_i64, c_i32, c_i32, c_i32, c_i32, c_i32, vp, vp, c_i64, c_dbl,
                                  vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, c_int,
                                  vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, c_size, c_int, vp, c_size, vp]),
    "tgs_gemm_bf16": (c_int, [c_int, c_int, c_int, c_i64, c_i64, c_i64, vp, c_i64, vp, c_i64, vp, vp, c_i64, vp,
                              c_i64, c_i32, vp]),
    "tgs_sigmoid_bwd": (c_int, [c_i64, vp, vp, vp, vp]),
    "tgs_colsum": (c_int, [c_i64, c_i64, vp, c_i64, vp, vp]),
    "tgs_decoder_pack_bytes": (c_size, []),
    "tgs_decoder_pack": (c_int, [vp, vp, vp, vp, vp]),
    "tgs_decode_textures": (c_int, [c_i64, vp, vp, vp, vp, vp, vp, vp]),
    "tgs_selective_adam_update": (c_int, [c_i64, c_i64, vp, vp, vp, vp, vp, c_dbl, c_dbl, c_dbl, c_dbl, vp]),
}


def lib() -> ctypes.CDLL:
    """The loaded library (loaded once per process).  Raises if not built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(
                        f"{LIB_PATH} is missing: build it with `python -m textured_gaussians_amd.build`")
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in _PROTOS.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def check(rc: int, op: str) -> None:
    if rc != 0:
        msg = lib().tgs_last_error().decode(errors="replace")
        raise RuntimeError(f"{op}: {msg} (code {rc})")
