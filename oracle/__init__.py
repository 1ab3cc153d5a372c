"""CPU oracle for the textured-2DGS render path.

TEST INFRASTRUCTURE ONLY.  This package is a restatement, on the CPU, of what
the reference computes on its GPU path (Camak2441/textured_gaussians,
``textured_gaussians/cuda/csrc/*.cu`` and ``textured_gaussians/cuda/_wrapper.py``).
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline.  The product package ``textured_gaussians_amd`` never imports it and
has no CPU fallback.

Modules
-------
``projection``  2DGS ray-splat projection (fwd, reference VJPs).
``isect``       tile intersection, stable key sort, offset encode (bit-exact,
                numpy float32 arithmetic where the reference is float32).
``sh``          real spherical harmonics up to degree 4.
``raster``      textured / plain front-to-back compositing, differentiable
                (torch autograd gives the exact gradient; ``uv_grad`` selects
                the reference's ``bilinear`` / ``bilinear_bwd2`` gradient rules).
``render``      the full pipeline (projection -> isect -> SH -> raster).
``adam``        the reference's selective Adam update (no bias correction).
``ssim``        the trainer's SSIM (torchmetrics' published algorithm), checker of the
                fused SSIM kernels.

Parity pins (what checks each function, besides the GPU comparison)
-------------------------------------------------------------------
* projection: closed forms for an on-axis / rotated Gaussian, and
  the row-major layout of ``ray_transforms`` on an asymmetric case
  (tests/test_oracle_projection.py); the reference VJP restatement is pinned
  against the exact autograd VJP on its exact terms and against a
  hand-derived formula of the one term it drops.
* isect / offsets: brute-force python double loop (reference loop semantics),
  ``numpy.lexsort`` stable order, offsets via ``searchsorted``, invariants
  (tests/test_oracle_isect.py).
* sh: closed-form degree-0 value, orthonormality of the basis under
  Monte-Carlo integration over the sphere, autograd vs central differences.
* raster: ``grid_sample(align_corners=True)`` equals the bilinear lookup
  inside the texture range; a white/zero texture reduces to plain 2DGS with
  the textured alpha law; a single camera-facing Gaussian gives the analytic
  ``s``; ``render_alpha = 1 - T``; background linearity; brute-force
  sequential per-pixel loop on tiny inputs; autograd vs float64 central
  differences (tests/test_oracle_raster.py).
* adam: closed form per element vs ``torch.optim.Adam`` after undoing the
  bias correction.
"""
