"""SH colours and the fused selective Adam step on the GPU vs the oracle."""
import pytest
import torch

from oracle.adam import selective_adam_update as adam_ref
from oracle.sh import spherical_harmonics as sh_ref
from textured_gaussians_amd.cuda import _ops
from textured_gaussians_amd.optimizers import SelectiveAdam

from gpu_util import DEV, cpu64, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("degree", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_sh_fwd_bwd(degree, dtype):
    g = torch.Generator().manual_seed(degree)
    n, K = 3001, (degree + 1) ** 2 + 1  # one unused band
    dirs = torch.randn(n, 3, generator=g, dtype=torch.float64)
    coeffs = torch.randn(n, K, 3, generator=g, dtype=torch.float64)
    masks = torch.rand(n, generator=g) > 0.2
    vcol = torch.randn(n, 3, generator=g, dtype=torch.float64)
    dv = lambda t: t.to(DEV, dtype).contiguous()  # noqa: E731
    col = _ops.compute_sh_fwd(degree, dv(dirs), dv(coeffs), masks.to(DEV))
    vco, vdi = _ops.compute_sh_bwd(K, degree, dv(dirs), dv(coeffs), masks.to(DEV), dv(vcol), True)
    d = dirs.clone().requires_grad_(True)
    c = coeffs.clone().requires_grad_(True)
    ref = sh_ref(degree, d, c, masks)
    (ref * vcol).sum().backward()
    tol = 1e-5 if dtype == torch.float32 else 1e-12
    assert rel_err(col, ref) < tol
    assert rel_err(vco, c.grad) < tol
    if degree == 0:
        assert torch.all(vdi == 0)
    else:
        assert rel_err(vdi, d.grad) < (1e-4 if dtype == torch.float32 else 1e-11)
    assert torch.all(col[~masks.to(DEV)] == 0)


@pytest.mark.parametrize("degree", [1, 3])
@pytest.mark.parametrize("offset", [0, 1])
def test_sh_bwd_exact_band_count_and_unaligned_buffers(degree, offset):
    """K = (degree + 1)^2 exactly (K * 3 a multiple of 4: the float32 kernel moves the
    coefficient blocks as 16-byte vectors), a ragged last wave, and coefficient
    buffers one float off 16-byte alignment (element-wise path)."""
    g = torch.Generator().manual_seed(10 + degree)
    n, K = 3001, (degree + 1) ** 2
    dirs = torch.randn(n, 3, generator=g, dtype=torch.float64)
    coeffs = torch.randn(n, K, 3, generator=g, dtype=torch.float64)
    masks = torch.rand(n, generator=g) > 0.2
    vcol = torch.randn(n, 3, generator=g, dtype=torch.float64)
    buf = torch.zeros(n * K * 3 + offset, device=DEV)
    cg = buf[offset:].view(n, K, 3)
    cg.copy_(coeffs)
    vco, vdi = _ops.compute_sh_bwd(K, degree, dirs.to(DEV, torch.float32), cg, masks.to(DEV),
                                   vcol.to(DEV, torch.float32), True)
    d = dirs.clone().requires_grad_(True)
    c = coeffs.clone().requires_grad_(True)
    (sh_ref(degree, d, c, masks) * vcol).sum().backward()
    assert rel_err(vco, c.grad) < 1e-5
    assert rel_err(vdi, d.grad) < 1e-4


@pytest.mark.parametrize("degree,extra", [(0, 0), (1, 1), (2, 0), (3, 0), (3, 1)])
def test_sh_views_kernel_equals_the_expanded_form(degree, extra):
    """One coefficient set seen from C views: the views kernel against
    ``spherical_harmonics`` of the coefficients expanded to every view (what the
    reference computes) and against the oracle.  Colours are bit-equal (same
    per-(view, Gaussian) arithmetic); gradients agree to float32 rounding (the
    views are summed in the kernel instead of by the expand's backward)."""
    from textured_gaussians_amd.cuda import spherical_harmonics, spherical_harmonics_views
    g = torch.Generator().manual_seed(20 + degree)
    C, n, K = 3, 2049, (degree + 1) ** 2 + extra
    dirs = torch.randn(C, n, 3, generator=g, dtype=torch.float64)
    coeffs = torch.randn(n, K, 3, generator=g, dtype=torch.float64)
    masks = torch.rand(C, n, generator=g) > 0.2
    vcol = torch.randn(C, n, 3, generator=g, dtype=torch.float64)
    outs = {}
    for fused in (True, False):
        d = dirs.to(DEV, torch.float32).requires_grad_(True)
        c = coeffs.to(DEV, torch.float32).requires_grad_(True)
        if fused:
            col = spherical_harmonics_views(degree, d, c, masks.to(DEV))
        else:
            col = spherical_harmonics(degree, d, c.expand(C, -1, -1, -1), masks.to(DEV))
        (col * vcol.to(DEV, torch.float32)).sum().backward()
        outs[fused] = (col.detach(), c.grad, d.grad)
    assert torch.equal(outs[True][0], outs[False][0])
    assert rel_err(outs[True][1], outs[False][1]) < 1e-6
    assert rel_err(outs[True][2], outs[False][2]) < 1e-6
    dd = dirs.clone().requires_grad_(True)
    cc = coeffs.clone().requires_grad_(True)
    ref = sh_ref(degree, dd, cc.expand(C, -1, -1, -1), masks)
    (ref * vcol).sum().backward()
    assert rel_err(outs[True][0], ref) < 1e-5
    assert rel_err(outs[True][1], cc.grad) < 1e-5
    if degree > 0:
        assert rel_err(outs[True][2], dd.grad) < 1e-4
    else:
        assert torch.all(outs[True][2] == 0)


def test_selective_adam_matches_oracle_and_torch():
    g = torch.Generator().manual_seed(0)
    N, M = 1000, 48  # M % 4 == 0: the vectorised path
    for m_cols in (M, 7):  # and the scalar path
        p = torch.randn(N, m_cols, generator=g)
        gr = torch.randn(N, m_cols, generator=g)
        m = torch.randn(N, m_cols, generator=g).abs() * 0.1
        v = torch.randn(N, m_cols, generator=g).abs() * 0.1
        vis = torch.rand(N, generator=g) > 0.3
        pd, md, vd = p.to(DEV), m.to(DEV), v.to(DEV)
        _ops.selective_adam_update(pd, gr.to(DEV), md, vd, vis.to(DEV), 1e-2, 0.9, 0.999, 1e-15, N, m_cols)
        pe, me, ve = adam_ref(p, gr, m, v, vis, 1e-2, 0.9, 0.999, 1e-15)
        # float32 with FMA contraction on the GPU: last-bit differences
        assert torch.allclose(pd.cpu(), pe, rtol=1e-5, atol=1e-6)
        assert torch.allclose(md.cpu(), me, rtol=1e-5, atol=1e-7)
        assert torch.allclose(vd.cpu(), ve, rtol=1e-5, atol=1e-7)
        assert torch.equal(pd.cpu()[~vis], p[~vis])


def test_selective_adam_optimizer_first_step_equals_torch_adam():
    # no bias correction: step 1 of the reference rule equals torch Adam with
    # lr' = lr (1-b1)/sqrt(1-b2) and eps' = eps/sqrt(1-b2), on all-visible rows
    g = torch.Generator().manual_seed(1)
    p0 = torch.randn(500, 8, generator=g)
    grad = torch.randn(500, 8, generator=g)
    lr, b1, b2, eps = 1e-2, 0.9, 0.99, 1e-8
    a = torch.nn.Parameter(p0.clone().to(DEV))
    opt = SelectiveAdam([a], eps=eps, betas=(b1, b2))
    opt.param_groups[0]["lr"] = lr
    a.grad = grad.to(DEV)
    opt.step(torch.ones(500, dtype=torch.bool, device=DEV))
    b = torch.nn.Parameter(p0.clone())
    tor = torch.optim.Adam([b], lr=lr * (1 - b1) / (1 - b2) ** 0.5, betas=(b1, b2), eps=eps / (1 - b2) ** 0.5)
    b.grad = grad.clone()
    tor.step()
    assert torch.allclose(a.detach().cpu(), b.detach(), rtol=1e-5, atol=1e-6)
