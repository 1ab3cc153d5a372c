"""CPU checks of the training-step helpers around the hot path: the oracle's SSIM
(the checker of the fused SSIM kernels) and the float32-master split of the bfloat16
texture state; the product's SSIM raises on CPU tensors.

SSIM (torchmetrics, not vendored: parity with it is unpinned) is pinned against a
brute-force float64 restatement of its published algorithm (a direct 2-D window sum
at every valid position) and its fixed points (SSIM(x, x) = 1, symmetry)."""
import math

import numpy as np
import pytest
import torch

from oracle import ssim as O
from textured_gaussians_amd import losses
from textured_gaussians_amd.texture_adam import join_master, split_master


def _ssim_bruteforce(x, y, size=11, sigma=1.5, c1=1e-4, c2=9e-4):
    """float64 numpy, [C, H, W] images: mean over channels and valid windows."""
    d = np.arange(size) - (size - 1) / 2
    g = np.exp(-0.5 * (d / sigma) ** 2)
    g /= g.sum()
    w = np.outer(g, g)
    C, H, W = x.shape
    vals = []
    for c in range(C):
        for i in range(H - size + 1):
            for j in range(W - size + 1):
                a = x[c, i:i + size, j:j + size]
                b = y[c, i:i + size, j:j + size]
                mx, my = (w * a).sum(), (w * b).sum()
                sxx = max((w * a * a).sum() - mx * mx, 0.0)
                syy = max((w * b * b).sum() - my * my, 0.0)
                sxy = (w * a * b).sum() - mx * my
                vals.append((2 * mx * my + c1) * (2 * sxy + c2) / ((mx * mx + my * my + c1) * (sxx + syy + c2)))
    return float(np.mean(vals))


def test_gaussian_window_is_normalised_and_symmetric():
    g = O.gaussian_window(11, 1.5, torch.float64)
    assert g.shape == (11,) and abs(float(g.sum()) - 1) < 1e-15
    assert torch.equal(g, g.flip(0)) and int(g.argmax()) == 5
    assert abs(float(g[5] / g[6]) - math.exp(0.5 / 1.5 ** 2)) < 1e-12


def test_ssim_matches_bruteforce_window_sum():
    gen = torch.Generator().manual_seed(0)
    x = torch.rand(2, 3, 17, 21, generator=gen, dtype=torch.float64)
    y = (x + 0.2 * torch.randn(2, 3, 17, 21, generator=gen, dtype=torch.float64)).clamp(0, 1)
    got = float(O.ssim(x, y))
    want = np.mean([_ssim_bruteforce(x[b].numpy(), y[b].numpy()) for b in range(2)])
    assert abs(got - want) < 1e-12, (got, want)
    assert abs(float(O.ssim(y, x)) - got) < 1e-14  # symmetric


def test_ssim_fixed_points():
    gen = torch.Generator().manual_seed(1)
    x = torch.rand(1, 3, 16, 16, generator=gen, dtype=torch.float64)
    assert abs(float(O.ssim(x, x)) - 1.0) < 1e-12
    assert float(O.ssim(x, 1 - x)) < 0.5


def test_ssim_gradient_is_the_derivative():
    gen = torch.Generator().manual_seed(2)
    x = torch.rand(1, 2, 12, 13, generator=gen, dtype=torch.float64, requires_grad=True)
    y = torch.rand(1, 2, 12, 13, generator=gen, dtype=torch.float64)
    assert torch.autograd.gradcheck(lambda a: O.ssim(a, y), (x,), eps=1e-6, atol=1e-8)


def test_product_ssim_has_no_cpu_path():
    x = torch.rand(1, 12, 12, 3)
    with pytest.raises(RuntimeError, match="GPU tensors"):
        losses.ssim_hwc(x, x)
    with pytest.raises(RuntimeError, match="GPU tensors"):
        losses.l1_dssim_loss(x, x)


def test_master_split_round_trips_bit_exactly():
    gen = torch.Generator().manual_seed(4)
    vals = [torch.randn(100_000, generator=gen) * 10.0 ** e for e in (-30, -3, 0, 3, 30)]
    special = torch.tensor([0.0, -0.0, 1.0, -1.0, 3.4028235e38, -3.4028235e38, 1e-45, -1e-45, 1.17549435e-38,
                            float("inf"), float("-inf")])
    # exact ties of the bfloat16 rounding (low 16 bits 0x8000) and their neighbours
    ties = (torch.randint(0, 1 << 15, (1000,), generator=gen, dtype=torch.int32) << 16 | 0x8000).view(torch.float32)
    x = torch.cat(vals + [special, ties, ties.neg()])
    x = x[torch.isfinite(x) | torch.isinf(x)]
    hi, lo = split_master(x)
    assert hi.dtype == torch.bfloat16 and lo.dtype == torch.int16
    assert torch.equal(join_master(hi, lo).view(torch.int32), x.view(torch.int32))


def test_master_split_render_copy_is_round_to_nearest():
    gen = torch.Generator().manual_seed(5)
    x = torch.randn(200_000, generator=gen)
    hi, _ = split_master(x)
    low = x.view(torch.int32) & 0xFFFF
    not_tie = low != 0x8000
    assert torch.equal(hi[not_tie], x[not_tie].to(torch.bfloat16))  # torch rounds to nearest even
    err = (hi.float() - x).abs()
    ulp = (x.abs().view(torch.int32) & 0x7F800000).view(torch.float32) * 2.0 ** -7
    assert bool((err <= 0.5 * ulp).all())
