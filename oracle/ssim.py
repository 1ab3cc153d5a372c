"""SSIM of the reference trainer's loss, restated on the CPU (test infrastructure only).

The trainer uses torchmetrics' ``StructuralSimilarityIndexMeasure(data_range=1.0)``
(examples/simple_trainer_textured_gaussians.py:430, 1424-1429), a third-party module
that is not vendored; parity with it is unpinned.  Its published algorithm, restated
here: an 11-tap Gaussian window (sigma 1.5, normalised) applied to x, y, x^2, y^2, xy
per channel; c1 = (0.01 R)^2, c2 = (0.03 R)^2;
SSIM = (2 mu_x mu_y + c1)(2 s_xy + c2) / ((mu_x^2 + mu_y^2 + c1)(s_x^2 + s_y^2 + c2));
the map is averaged over the windows that lie inside the image (torchmetrics pads by
reflection and then crops the padded border away, which leaves exactly those), then
over images.  Variances are clamped at 0.  Pinned against a brute-force window sum
(tests/test_losses_texture_state.py).
"""
import torch
import torch.nn.functional as F
from torch import Tensor


def gaussian_window(size: int = 11, sigma: float = 1.5, dtype=torch.float32, device=None) -> Tensor:
    d = torch.arange((1 - size) / 2, (1 + size) / 2, 1.0, dtype=dtype, device=device)
    g = torch.exp(-0.5 * (d / sigma) ** 2)
    return g / g.sum()


def ssim(preds: Tensor, target: Tensor, data_range: float = 1.0, kernel_size: int = 11, sigma: float = 1.5,
         k1: float = 0.01, k2: float = 0.03) -> Tensor:
    """Mean SSIM of [B, C, H, W] images (H, W >= kernel_size), differentiable."""
    assert preds.shape == target.shape and preds.dim() == 4, (preds.shape, target.shape)
    B, C, H, W = preds.shape
    assert H >= kernel_size and W >= kernel_size, "image smaller than the SSIM window"
    g = gaussian_window(kernel_size, sigma, preds.dtype, preds.device)
    kv = g.view(1, 1, -1, 1).expand(C, 1, kernel_size, 1).contiguous()
    kh = g.view(1, 1, 1, -1).expand(C, 1, 1, kernel_size).contiguous()
    x = torch.cat([preds, target, preds * preds, target * target, preds * target])  # [5B, C, H, W]
    y = F.conv2d(F.conv2d(x, kv, groups=C), kh, groups=C)  # separable window, valid positions only
    mu_p, mu_t, e_pp, e_tt, e_pt = y.split(B)
    c1, c2 = (k1 * data_range) ** 2, (k2 * data_range) ** 2
    mu_pp, mu_tt, mu_pt = mu_p * mu_p, mu_t * mu_t, mu_p * mu_t
    s_pp = (e_pp - mu_pp).clamp_min(0.0)
    s_tt = (e_tt - mu_tt).clamp_min(0.0)
    s_pt = e_pt - mu_pt
    ssim_map = ((2 * mu_pt + c1) * (2 * s_pt + c2)) / ((mu_pp + mu_tt + c1) * (s_pp + s_tt + c2))
    return ssim_map.reshape(B, -1).mean(-1).mean()
