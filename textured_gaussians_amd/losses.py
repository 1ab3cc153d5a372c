"""Photometric loss of the reference trainer: ``(1 - lambda) * L1 + lambda * (1 - SSIM)``
(examples/simple_trainer_textured_gaussians.py:1424-1429, lambda = ssim_lambda = 0.2).

The trainer's SSIM is torchmetrics' ``StructuralSimilarityIndexMeasure(data_range=1.0)``
([TR]:430): an 11-tap Gaussian window (sigma 1.5) over x, y, x^2, y^2, xy per channel,
c1 = (0.01 R)^2, c2 = (0.03 R)^2, the map averaged over the windows inside the image,
variances clamped at 0.  Here it runs as the fused gfx950 kernels (``tgs_ssim_fwd`` /
``tgs_ssim_bwd``); there is no other implementation in the package.
"""
import torch
from torch import Tensor


class _FusedSSIM(torch.autograd.Function):
    """SSIM of channels-last [B, H, W, C] images by the gfx950 kernels.  The kernels
    differentiate in the first image; SSIM is symmetric, so the second image's gradient
    is the first-image gradient of SSIM(y, x)."""

    @staticmethod
    def forward(ctx, x, y):
        from .cuda import _ops
        s, work = _ops.ssim_fwd(x, y)
        ctx.save_for_backward(x, y, work)
        return s

    @staticmethod
    def backward(ctx, g):
        from .cuda import _ops
        x, y, work = ctx.saved_tensors
        gx = _ops.ssim_bwd(x, y, work) * g if ctx.needs_input_grad[0] else None
        gy = None
        if ctx.needs_input_grad[1]:
            _, work_yx = _ops.ssim_fwd(y, x)
            gy = _ops.ssim_bwd(y, x, work_yx) * g
        return gx, gy


def ssim_hwc(x: Tensor, y: Tensor) -> Tensor:
    """Mean SSIM of channels-last [B, H, W, C] images (C <= 4, float32 or float64, on
    the GPU), differentiable in both images."""
    if not (x.is_cuda and y.is_cuda):
        raise RuntimeError("ssim_hwc: the fused SSIM kernels need GPU tensors")
    if x.dtype not in (torch.float32, torch.float64) or y.dtype != x.dtype:
        raise RuntimeError(f"ssim_hwc: float32 or float64 images of one dtype expected, got {x.dtype}, {y.dtype}")
    if x.shape != y.shape or x.dim() != 4 or not 1 <= x.shape[-1] <= 4:
        raise RuntimeError(f"ssim_hwc: [B, H, W, C] images with C <= 4 expected, got {tuple(x.shape)}, "
                           f"{tuple(y.shape)}")
    return _FusedSSIM.apply(x.contiguous(), y.contiguous())


def l1_dssim_loss(render: Tensor, target: Tensor, ssim_lambda: float = 0.2) -> Tensor:
    """The trainer's loss on [B, H, W, C] images: (1 - l) L1 + l (1 - SSIM)."""
    l1 = (render - target).abs().mean()
    d = 1.0 - ssim_hwc(render, target)  # SSIM is symmetric in its two images
    return l1 * (1.0 - ssim_lambda) + d * ssim_lambda
