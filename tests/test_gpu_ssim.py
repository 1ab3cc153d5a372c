"""The fused SSIM kernels against the oracle's restatement of torchmetrics' algorithm
(``oracle.ssim``, itself pinned against a brute-force window sum on the CPU): value
and gradients in both images, float64 and float32, including constant regions
(clamped variances); the trainer's loss formula at 1920x1080; unsupported inputs raise."""
import pytest
import torch

from oracle import ssim as O
from textured_gaussians_amd import losses

from gpu_util import DEV, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-5)])
@pytest.mark.parametrize("shape", [(1, 37, 53, 3), (2, 11, 16, 4), (1, 64, 96, 1)])
def test_fused_ssim_matches_torch(dtype, tol, shape):
    g = torch.Generator().manual_seed(sum(shape))
    x = torch.rand(shape, generator=g, dtype=torch.float64)
    y = (x + 0.3 * torch.randn(shape, generator=g, dtype=torch.float64)).clamp(0, 1)
    y[:, : shape[1] // 3] = 0.5  # a constant band
    x[:, -3:, : shape[2] // 2] = 0.25
    xd = x.to(DEV, dtype).requires_grad_(True)
    s = losses.ssim_hwc(xd, y.to(DEV, dtype))
    s.backward()
    xr = x.to(dtype).requires_grad_(True)
    sr = O.ssim(xr.permute(0, 3, 1, 2), y.to(dtype).permute(0, 3, 1, 2))
    sr.backward()
    assert abs(float(s.detach()) - float(sr.detach())) < tol * 10
    assert rel_err(xd.grad, xr.grad) < tol * 100


def test_fused_ssim_in_the_trainer_loss_at_full_hd():
    g = torch.Generator().manual_seed(7)
    r = torch.rand(1, 1080, 1920, 3, generator=g).to(DEV).requires_grad_(True)
    t = torch.rand(1, 1080, 1920, 3, generator=g).to(DEV)
    loss = losses.l1_dssim_loss(r, t)
    loss.backward()
    r64 = r.detach().double().cpu().requires_grad_(True)
    ref = 0.8 * (r64 - t.double().cpu()).abs().mean() + 0.2 * (1 - O.ssim(r64.permute(0, 3, 1, 2),
                                                                               t.double().cpu().permute(0, 3, 1, 2)))
    ref.backward()
    assert abs(float(loss.detach()) - float(ref.detach())) < 1e-5
    assert rel_err(r.grad, r64.grad) < 1e-4


def test_fused_ssim_gradient_in_both_images():
    g = torch.Generator().manual_seed(11)
    x = torch.rand(1, 29, 31, 3, generator=g, dtype=torch.float64)
    y = (x + 0.2 * torch.randn(1, 29, 31, 3, generator=g, dtype=torch.float64)).clamp(0, 1)
    xd, yd = x.to(DEV).requires_grad_(True), y.to(DEV).requires_grad_(True)
    losses.ssim_hwc(xd, yd).backward()
    xr, yr = x.clone().requires_grad_(True), y.clone().requires_grad_(True)
    O.ssim(xr.permute(0, 3, 1, 2), yr.permute(0, 3, 1, 2)).backward()
    assert rel_err(xd.grad, xr.grad) < 1e-10 and rel_err(yd.grad, yr.grad) < 1e-10


def test_fused_ssim_rejects_what_it_does_not_run():
    x = torch.rand(1, 16, 16, 5, device=DEV)
    with pytest.raises(RuntimeError, match="C <= 4"):
        losses.ssim_hwc(x, x)
    with pytest.raises(RuntimeError, match="one dtype"):
        losses.ssim_hwc(x[..., :3], x[..., :3].double())
