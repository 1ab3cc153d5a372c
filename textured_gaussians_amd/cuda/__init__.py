"""Native (gfx950) ops of the textured-2DGS render path and their autograd wrappers."""
from ._wrapper import (  # noqa: F401
    fully_fused_projection_2dgs,
    generate_mipmap,
    isect_offset_encode,
    isect_tiles,
    rasterize_to_pixels_2dgs,
    rasterize_to_pixels_textured_gaussians,
    selective_adam_update,
    spherical_harmonics,
    spherical_harmonics_views,
    tgs_fns,
)
