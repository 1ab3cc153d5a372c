"""2DGS projection forward / VJPs on the GPU vs the oracle."""
import math

import pytest
import torch

from oracle import projection as OP
from textured_gaussians_amd import scenes
from textured_gaussians_amd.cuda import _ops

from gpu_util import DEV, cpu64, rel_err

pytestmark = pytest.mark.gpu


def _scene(n=4000, C=2):
    cams = [(4 * math.cos(a), 4 * math.sin(a), 0.7) for a in [0.3 * k for k in range(C)]]
    return scenes.make_scene(n, 160, 120, 55.0, cams, None, logscale_mu=-3.0, sh_degree=0)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_projection_fwd(dtype):
    sc = _scene()
    s = sc.to(DEV, dtype)
    radii, m2, dep, rt, nrm = _ops.fully_fused_projection_fwd_2dgs(s.means, s.quats, s.scales, s.viewmats, s.Ks,
                                                                   s.width, s.height, 0.3, 0.01, 1e10, 0.0)
    d = sc.to(dtype=torch.float64)
    r_o, m2_o, dep_o, rt_o, nrm_o = OP.project_2dgs(d.means, d.quats, d.scales, d.viewmats, d.Ks, d.width,
                                                    d.height)
    both = (radii.cpu() > 0) & (r_o > 0)
    # radii are ceil() of a float; float32 may land on the other side of an integer
    mism = (radii.cpu() != r_o)
    tol_frac = 1e-3 if dtype == torch.float32 else 0.0
    assert mism.float().mean().item() <= tol_frac
    assert both.sum() > 100
    tol = 2e-5 if dtype == torch.float32 else 1e-12
    for a, b in [(m2, m2_o), (dep, dep_o), (rt, rt_o), (nrm, nrm_o)]:
        assert rel_err(cpu64(a)[both], b[both]) < tol


@pytest.mark.parametrize("mode", ["reference", "reference_norm", "exact"])
def test_projection_bwd(mode):
    sc = _scene(n=3000)
    d = sc.to(dtype=torch.float64)
    s = sc.to(DEV, torch.float64)
    radii, m2, dep, rt, nrm = _ops.fully_fused_projection_fwd_2dgs(s.means, s.quats, s.scales, s.viewmats, s.Ks,
                                                                   s.width, s.height, 0.3, 0.01, 1e10, 0.0)
    g = torch.Generator().manual_seed(5)
    C, N = radii.shape
    vm2 = torch.randn(C, N, 2, generator=g, dtype=torch.float64)
    vd = torch.randn(C, N, generator=g, dtype=torch.float64)
    vn = torch.randn(C, N, 3, generator=g, dtype=torch.float64)
    vrt = torch.randn(C, N, 3, 3, generator=g, dtype=torch.float64)
    fn = {"reference": _ops.fully_fused_projection_bwd_2dgs, "reference_norm": _ops.fully_fused_projection_bwd2_2dgs,
          "exact": _ops.fully_fused_projection_bwd_exact_2dgs}[mode]
    vmeans, vquats, vscales, vview = fn(s.means, s.quats, s.scales, s.viewmats, s.Ks, s.width, s.height, radii, rt,
                                        vm2.to(DEV), vd.to(DEV), vn.to(DEV), vrt.to(DEV), mode == "exact")
    if mode == "exact":
        leaves = [d.means.clone().requires_grad_(True), d.quats.clone().requires_grad_(True),
                  d.scales.clone().requires_grad_(True), d.viewmats.clone().requires_grad_(True)]
        r_o, m2_o, dep_o, rt_o, nrm_o = OP.project_2dgs(leaves[0], leaves[1], leaves[2], leaves[3], d.Ks, d.width,
                                                        d.height)
        assert torch.equal(r_o, radii.cpu())
        L = (m2_o * vm2).sum() + (dep_o * vd).sum() + (nrm_o * vn).sum() + (rt_o * vrt).sum()
        L.backward()
        refs = [l.grad for l in leaves]
        outs = [vmeans, vquats, vscales, vview]
    else:
        refs = OP.projection_vjp_reference(d.means, d.quats, d.scales, d.viewmats, d.Ks, radii.cpu(), cpu64(rt),
                                           vm2, vd, vn, vrt, norm_rot_grad=(mode == "reference_norm"))
        outs = [vmeans, vquats, vscales]
    for o, r in zip(outs, refs):
        assert rel_err(o, r) < 1e-10, (mode, rel_err(o, r))
