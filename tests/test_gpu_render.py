"""End-to-end parity of the public API on the GPU against the CPU oracle, and
FP64 central finite differences of the exact-gradient mode.

All comparisons here run the GPU kernels in float64, so they hold to ~1e-9:
the kernels implement the same arithmetic as the oracle, and every pixel
counts (no float32 threshold flips)."""
import math

import pytest
import torch

import textured_gaussians_amd as tg
from oracle.render import render as oracle_render
from textured_gaussians_amd import scenes
from textured_gaussians_amd.cuda import _ops

from gpu_util import DEV, cpu64, rel_err

pytestmark = pytest.mark.gpu

F64 = torch.float64


def _scene(n=600, w=64, h=48, cams=None, tex=10, sh_degree=1, seed=0):
    cams = cams or [(3.5, 0.3, 0.4)]
    return scenes.make_scene(n, w, h, 50.0, cams, tex, logscale_mu=-3.0, logscale_sigma=0.4,
                             sh_degree=sh_degree, seed=seed).to(dtype=F64)


def _leaves(sc, dev):
    names = ["means", "quats", "scales", "opacities", "colors", "textures"]
    return {k: getattr(sc, k).detach().to(dev).clone().requires_grad_(True) for k in names}


def _gpu_render(sc, L, **kw):
    return tg.rasterization_textured_gaussians(
        L["means"], L["quats"], L["scales"], L["opacities"], L["colors"], L["textures"], sc.viewmats.to(DEV),
        sc.Ks.to(DEV), sc.width, sc.height, sh_degree=sc.sh_degree, backgrounds=sc.backgrounds.to(DEV), **kw)


@pytest.mark.parametrize("render_mode", ["RGB", "RGB+D", "RGB+ED", "D"])
def test_api_forward_matches_oracle(render_mode):
    sc = _scene()
    L = _leaves(sc, DEV)
    bg = sc.backgrounds if render_mode == "RGB" else None
    out = tg.rasterization_textured_gaussians(
        L["means"], L["quats"], L["scales"], L["opacities"], L["colors"], L["textures"], sc.viewmats.to(DEV),
        sc.Ks.to(DEV), sc.width, sc.height, sh_degree=1, render_mode=render_mode,
        backgrounds=None if bg is None else bg.to(DEV)[:, :3])
    ref = oracle_render(sc.means, sc.quats, sc.scales, sc.opacities, sc.colors, sc.viewmats, sc.Ks, sc.width,
                        sc.height, textures=sc.textures, sh_degree=1, backgrounds=bg, render_mode=render_mode)
    assert torch.equal(out[-1]["flatten_ids"].cpu(), ref["meta"]["flatten_ids"])
    for a, b in [(out[0], ref["render_colors"]), (out[1], ref["render_alphas"]), (out[2], ref["render_normals"]),
                 (out[4], ref["render_distort"]), (out[5], ref["render_median"])]:
        assert (cpu64(a) - b).abs().max().item() < 1e-9
    if render_mode in ("RGB+D", "RGB+ED"):
        assert out[3] is not None and out[3].shape == (sc.height, sc.width, 3)


def test_api_exact_gradients_match_oracle_autograd():
    sc = _scene(n=500, sh_degree=1, seed=2)
    L = _leaves(sc, DEV)
    out = _gpu_render(sc, L, grad_mode="exact")
    g = torch.Generator().manual_seed(9)
    wc = torch.randn(out[0].shape, generator=g, dtype=F64)
    wa = torch.randn(out[1].shape, generator=g, dtype=F64)
    ((out[0] * wc.to(DEV)).sum() + (out[1] * wa.to(DEV)).sum()).backward()
    R = {k: getattr(sc, k).clone().requires_grad_(True) for k in L}
    ref = oracle_render(R["means"], R["quats"], R["scales"], R["opacities"], R["colors"], sc.viewmats, sc.Ks,
                        sc.width, sc.height, textures=R["textures"], sh_degree=1, backgrounds=sc.backgrounds,
                        uv_grad="exact")
    ((ref["render_colors"] * wc).sum() + (ref["render_alphas"] * wa).sum()).backward()
    for k in L:
        assert rel_err(L[k].grad, R[k].grad) < 1e-8, (k, rel_err(L[k].grad, R[k].grad))


def test_multicamera_forward_and_texture_indexing():
    # textures and contribution counters are indexed by the Gaussian id, not the
    # flatten id, so cameras >= 1 stay in bounds and use the right texture
    cams = [(3.5 * math.cos(a), 3.5 * math.sin(a), 0.3) for a in (0.0, 1.5, 3.0)]
    sc = _scene(n=500, cams=cams, seed=4)
    L = _leaves(sc, DEV)
    out = _gpu_render(sc, L)
    ref = oracle_render(sc.means, sc.quats, sc.scales, sc.opacities, sc.colors, sc.viewmats, sc.Ks, sc.width,
                        sc.height, textures=sc.textures, sh_degree=1, backgrounds=sc.backgrounds)
    assert out[0].shape[0] == 3
    assert (cpu64(out[0]) - ref["render_colors"]).abs().max().item() < 1e-9
    assert rel_err(out[6], ref["contrib_sum"]) < 1e-10
    assert torch.equal(cpu64(out[7]), ref["contrib_count"])


def test_central_differences_fp64():
    """FP64 central differences (eps = 1e-6) of a fixed linear functional of the
    rendered image, for entries of every parameter, against the exact-mode
    gradient of the GPU path (200 Gaussians, 64x64 pixels)."""
    sc = _scene(n=200, w=64, h=64, tex=8, sh_degree=1, seed=7)
    L = _leaves(sc, DEV)
    g = torch.Generator().manual_seed(1)
    wc = torch.randn(1, 64, 64, 3, generator=g, dtype=F64).to(DEV)

    def f(P):
        with torch.no_grad():
            o = tg.rasterization_textured_gaussians(
                P["means"], P["quats"], P["scales"], P["opacities"], P["colors"], P["textures"],
                sc.viewmats.to(DEV), sc.Ks.to(DEV), 64, 64, sh_degree=1, backgrounds=sc.backgrounds.to(DEV))
        return (o[0] * wc).sum().item()

    out = _gpu_render(sc, L, grad_mode="exact")
    (out[0] * wc).sum().backward()
    vis = torch.nonzero(out[-1]["radii"][0] > 0).flatten().tolist()
    # Gaussians that cover many pixels carry informative gradients
    cnt = out[7].detach()
    picks = sorted(vis, key=lambda i: -float(cnt[i]))[:6]
    eps = 1e-6
    checked = 0
    for name in ["means", "quats", "scales", "opacities", "colors", "textures"]:
        P = {k: v.detach().clone() for k, v in L.items()}
        for gi in picks[:3]:
            flat_idx = [(gi,) + tuple(0 for _ in P[name].shape[1:])]
            if name == "textures":
                flat_idx = [(gi, 4, 4, 0), (gi, 3, 5, 3)]
            elif P[name].dim() > 1:
                flat_idx = [(gi,) + (j,) + tuple(0 for _ in P[name].shape[2:]) for j in range(min(3, P[name].shape[1]))]
            for idx in flat_idx:
                Pp = {k: v.clone() for k, v in P.items()}
                Pm = {k: v.clone() for k, v in P.items()}
                Pp[name][idx] += eps
                Pm[name][idx] -= eps
                fd = (f(Pp) - f(Pm)) / (2 * eps)
                an = L[name].grad[idx].item()
                scale = max(abs(fd), abs(an), 1e-3)
                assert abs(fd - an) / scale < 1e-4, (name, idx, fd, an)
                checked += 1
    assert checked >= 20


def test_texture_gradient_rows_are_filled_by_the_kernels():
    """The texture gradient starts as uninitialised memory (here: NaN left in the
    caching allocator) and is zero-filled inside the rasterizer kernels: every row
    must end finite, rows of Gaussians absent from every tile list exactly 0."""
    sc = _scene(n=400, tex=20, seed=11).to(dtype=torch.float32)  # 2.5 MB of texture: large-block pool
    sc.means[:50] = torch.tensor([3.5, 0.3, 0.4]) * 3.0  # behind the camera: in no tile list
    L = _leaves(sc, DEV)
    junk = torch.full(sc.textures.shape, float("nan"), device=DEV)
    del junk  # the next large allocation of this size (the gradient buffer) reuses the NaN block
    out = _gpu_render(sc, L)
    (out[0].sum() + out[1].sum()).backward()
    g = L["textures"].grad
    assert torch.isfinite(g).all()
    listed = torch.zeros(sc.means.shape[0], dtype=torch.bool)
    listed[out[-1]["flatten_ids"].long().cpu() % sc.means.shape[0]] = True
    assert (~listed).any() and listed.any()
    assert torch.all(g.cpu()[~listed] == 0)
    assert g.cpu()[listed].abs().sum() > 0


@pytest.mark.parametrize("records", [False, True])
def test_texture_gradient_paths_through_the_api(records):
    """Both texture-gradient paths of the autograd Function (float atomics into rows the
    forward zero-fills, or per-pair records summed per row in LDS) start from
    uninitialised memory and give the same gradients."""
    import textured_gaussians_amd.cuda._wrapper as W
    sc = _scene(n=400, tex=20, seed=12).to(dtype=torch.float32)
    sc.means[:50] = torch.tensor([3.5, 0.3, 0.4]) * 3.0  # behind the camera: in no tile list
    grads = {}
    saved = W.TEXTURE_GRAD_RECORDS
    try:
        for mode in (False, records):
            W.TEXTURE_GRAD_RECORDS = mode
            L = _leaves(sc, DEV)
            junk = torch.full(sc.textures.shape, float("nan"), device=DEV)
            del junk  # the gradient buffer reuses this NaN block
            out = _gpu_render(sc, L)
            (out[0].sum() + out[1].sum()).backward()
            grads[mode] = {k: v.grad.detach().clone() for k, v in L.items() if v.grad is not None}
    finally:
        W.TEXTURE_GRAD_RECORDS = saved
    for k, a in grads[False].items():
        b = grads[records][k]
        assert torch.isfinite(b).all(), k
        assert rel_err(b, a) < 1e-5, k


def test_masks_render_background():
    sc = _scene(n=300, seed=5).to(dtype=torch.float32)
    s = sc.to(DEV)
    radii, m2, dep, rt, nrm = _ops.fully_fused_projection_fwd_2dgs(s.means, s.quats, s.scales, s.viewmats, s.Ks,
                                                                   s.width, s.height, 0.3, 0.01, 1e10, 0.0)
    tw, th = math.ceil(s.width / 16), math.ceil(s.height / 16)
    _, ids, flat = _ops.isect_tiles(m2, radii, dep, None, None, 1, 16, tw, th, True, True)
    offs = _ops.isect_offset_encode(ids, 1, tw, th)
    masks = torch.ones(1, th, tw, dtype=torch.bool, device=DEV)
    masks[0, 0, 0] = False
    cols = torch.rand(1, s.means.shape[0], 3, device=DEV, dtype=torch.float32)
    op = s.opacities[None].contiguous()
    out = _ops.rasterize_to_pixels_fwd_textured_gaussians(m2, rt, cols, op, s.textures, 3.0, 3.0, nrm,
                                                          s.backgrounds, masks, s.width, s.height, 16, offs,
                                                          flat, 0.0, 1.0)
    assert torch.all(out[0][0, :16, :16] == 1.0) and torch.all(out[1][0, :16, :16] == 0.0)


def test_singular_view_matrix_raises():
    """The camera-to-world matrices come from ``torch.linalg.inv_ex`` (the routine
    of ``torch.inverse``) with the singularity check read after the intersection
    count: a singular view matrix still raises, as ``torch.inverse`` does in the
    reference.  The projection reads only the top three rows, so it runs normally."""
    sc = _scene(n=50)
    L = _leaves(sc, DEV)
    vm = sc.viewmats.to(DEV).clone()
    vm[0, 3, :] = 0.0
    with pytest.raises(RuntimeError, match="singular"):
        tg.rasterization_textured_gaussians(
            L["means"], L["quats"], L["scales"], L["opacities"], L["colors"], L["textures"], vm, sc.Ks.to(DEV),
            sc.width, sc.height, sh_degree=sc.sh_degree, backgrounds=sc.backgrounds.to(DEV))


@pytest.mark.parametrize("packed", [False, True])
def test_all_gaussians_culled(packed):
    """Every Gaussian behind the camera: no intersections.  The image is the
    background with zero alpha, and every gradient is zero."""
    sc = _scene(n=80).to(dtype=torch.float32)
    sc.means[:] = torch.tensor([3.5, 0.3, 0.4]) * 3.0
    L = _leaves(sc, DEV)
    out = _gpu_render(sc, L, packed=packed)
    assert out[-1]["isect_ids"].numel() == 0
    bg = sc.backgrounds.to(DEV)
    assert torch.equal(out[0][0], bg[0].expand_as(out[0][0]))
    assert torch.all(out[1] == 0)
    (out[0].sum() + out[1].sum()).backward()
    for k, v in L.items():
        assert v.grad is None or torch.all(v.grad == 0), k


@pytest.mark.parametrize("packed", [False, True])
def test_no_gaussians(packed):
    """N = 0: the image is the background, alphas are zero, the gradients are empty."""
    sc = _scene(n=20).to(dtype=torch.float32)
    L = {k: v[:0].contiguous().detach().requires_grad_(True) for k, v in _leaves(sc, DEV).items()}
    out = _gpu_render(sc, L, packed=packed)
    bg = sc.backgrounds.to(DEV)
    assert torch.equal(out[0][0], bg[0].expand_as(out[0][0]))
    assert torch.all(out[1] == 0)
    (out[0].sum() + out[1].sum()).backward()
    for k, v in L.items():
        assert v.grad is None or v.grad.numel() == 0, k


@pytest.mark.parametrize("packed", [False, True])
def test_projection_culling_options_match_oracle(packed):
    """near_plane, far_plane and radius_clip culling (the reference's
    fully_fused_projection_2dgs_fwd.cu:122-125,199-210) through the API, grid and
    packed layouts, against the oracle: the same Gaussians survive, and images and
    exact-mode gradients agree to float64 rounding."""
    sc = _scene(n=900, seed=11)
    kw = dict(near_plane=3.25, far_plane=3.85, radius_clip=3.0)
    L = _leaves(sc, DEV)
    out = _gpu_render(sc, L, grad_mode="exact", packed=packed, **kw)
    base = _gpu_render(sc, _leaves(sc, DEV), grad_mode="exact")
    R = {k: getattr(sc, k).clone().requires_grad_(True) for k in L}
    ref = oracle_render(R["means"], R["quats"], R["scales"], R["opacities"], R["colors"], sc.viewmats, sc.Ks,
                        sc.width, sc.height, textures=R["textures"], sh_degree=1, backgrounds=sc.backgrounds,
                        uv_grad="exact", **kw)
    r_ref = ref["meta"]["radii"]
    n_all, n_kept = int((base[-1]["radii"] > 0).sum()), int((r_ref > 0).sum())
    # the depth range keeps about a third, the radius clip fewer again; enough stay to test
    assert 0.1 * n_all < n_kept < 0.3 * n_all, (n_all, n_kept)
    if packed:
        kept = torch.zeros_like(r_ref, dtype=torch.bool)
        m = out[-1]
        kept[m["camera_ids"].cpu(), m["gaussian_ids"].cpu()] = True
        assert torch.equal(kept, r_ref > 0)
    else:
        assert torch.equal(out[-1]["radii"].cpu(), r_ref)
    assert (cpu64(out[0]) - ref["render_colors"]).abs().max().item() < 1e-9
    assert (cpu64(out[1]) - ref["render_alphas"]).abs().max().item() < 1e-9
    g = torch.Generator().manual_seed(3)
    wc = torch.randn(out[0].shape, generator=g, dtype=F64)
    (out[0] * wc.to(DEV)).sum().backward()
    (ref["render_colors"] * wc).sum().backward()
    for k in L:
        assert rel_err(L[k].grad, R[k].grad) < 1e-8, (k, rel_err(L[k].grad, R[k].grad))
