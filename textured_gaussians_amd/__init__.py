"""MI355X (gfx950) implementation of the textured 2D Gaussian splatting render
path: 2DGS projection, tile intersection with a stable radix sort, a
front-to-back textured rasterizer and its backward, SH colours and a fused
Adam step, as hand-written HIP kernels behind a C ABI (``include/tgs_hip.h``,
``libtgs_hip.so``), with the reference's Python API on top.
"""
from .rendering import rasterization_2dgs, rasterization_textured_gaussians  # noqa: F401
from .cuda import (  # noqa: F401
    fully_fused_projection_2dgs,
    isect_offset_encode,
    isect_tiles,
    rasterize_to_pixels_2dgs,
    rasterize_to_pixels_textured_gaussians,
    selective_adam_update,
    spherical_harmonics,
)
from .optimizers import SelectiveAdam  # noqa: F401
from . import checkpoint  # noqa: F401

__version__ = "0.1.0"
