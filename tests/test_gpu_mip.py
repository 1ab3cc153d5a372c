"""The ``mipmapped2`` filter on the GPU against the oracle: the pyramid (bit-exact in
the same arithmetic), its adjoint, and the trilinear rasterizer forward and backward
(the reference's gradient rule) in float64; the whole frame of the CPU-runnable workload
through the pyramid and the rasterizer's autograd in float32 and float64; the filter
through the public API."""
import pytest
import torch

import textured_gaussians_amd as tg
from oracle.raster import generate_mipmap as oracle_mipmap
from oracle.raster import rasterize_to_pixels
from textured_gaussians_amd import scenes
from textured_gaussians_amd.cuda import _ops

from gpu_util import DEV, all_tiles, cpu64, gpu_stages, gpu_stages_packed, raster_tile_parity, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("C,L", [(4, 4), (3, 5), (5, 0), (1, 6)])
def test_generate_mipmap_matches_oracle(dtype, C, L):
    g = torch.Generator().manual_seed(C * 10 + L)
    t = torch.rand(37, 1 << L, 1 << L, C, generator=g, dtype=dtype)
    got = _ops.generate_mipmap_fwd(t.to(DEV), L)
    want = oracle_mipmap(t)
    assert got.shape == want.shape == (37, _ops.mipmap_texels(L), C)
    assert torch.equal(got.cpu(), want)  # ((a + b) + c) + d then / 4 in both


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_generate_mipmap_backward_is_the_adjoint(dtype):
    L, C = 5, 4
    g = torch.Generator().manual_seed(1)
    t = torch.rand(9, 1 << L, 1 << L, C, generator=g, dtype=torch.float64, requires_grad=True)
    v = torch.randn(9, _ops.mipmap_texels(L), C, generator=g, dtype=torch.float64)
    (oracle_mipmap(t) * v).sum().backward()
    got = _ops.generate_mipmap_bwd(9, C, L, 4, 16, v.to(DEV, dtype))
    assert rel_err(got, t.grad) < (1e-14 if dtype == torch.float64 else 1e-6)


def test_mip2_raster_forward_and_backward_match_oracle():
    dtype = torch.float64
    sc = scenes.config1(n=2500, width=80, height=64, tex_res=16, seed=2, sh_degree=None)
    st = gpu_stages(sc, dtype)
    s = st["scene"]
    C, N = st["radii"].shape
    H, W = s.height, s.width
    g = torch.Generator().manual_seed(9)
    colors = torch.rand(C, N, 3, generator=g, dtype=torch.float64)
    opac = sc.opacities.double()[None].repeat(C, 1)
    bg = torch.rand(C, 3, generator=g, dtype=torch.float64)
    mip = oracle_mipmap(sc.textures.double())
    L = 4
    dv = lambda t: t.to(DEV, dtype).contiguous()  # noqa: E731
    out = _ops.rasterize_to_pixels_fwd_mip2_textured_gaussians(
        st["means2d"], st["rt"], dv(colors), dv(opac), dv(mip), L, 3.0, 3.0, st["normals"], dv(bg), None, W, H, 16,
        st["offsets"], st["flatten_ids"], 0.0, 1.0)
    vc = torch.randn(C, H, W, 3, generator=g, dtype=torch.float64)
    va = torch.randn(C, H, W, 1, generator=g, dtype=torch.float64)
    vn = torch.randn(C, H, W, 3, generator=g, dtype=torch.float64)
    vd = torch.randn(C, H, W, 1, generator=g, dtype=torch.float64)
    vm = torch.randn(C, H, W, 1, generator=g, dtype=torch.float64)
    res = _ops.rasterize_to_pixels_bwd_mip2_textured_gaussians(
        st["means2d"], st["rt"], dv(colors), dv(opac), dv(mip), L, 3.0, 3.0, st["normals"],
        torch.zeros_like(st["means2d"]), dv(bg), None, W, H, 16, st["offsets"], st["flatten_ids"], 1.0, out[0], out[1],
        out[5], out[6], dv(vc), dv(va), dv(vn), dv(vd), dv(vm), False)
    leaves = dict(means2d=cpu64(st["means2d"]).requires_grad_(True), rt=cpu64(st["rt"]).requires_grad_(True),
                  colors=colors.clone().requires_grad_(True), opac=opac.clone().requires_grad_(True),
                  normals=cpu64(st["normals"]).requires_grad_(True), tex=mip.clone().requires_grad_(True))
    r = rasterize_to_pixels(leaves["means2d"], leaves["rt"], leaves["colors"], leaves["opac"], leaves["normals"], W,
                            H, 16, st["offsets"].cpu(), st["flatten_ids"].cpu(), textures=leaves["tex"],
                            backgrounds=bg, uv_grad="none", distort_grad="reference", mip_log_res=L)
    names = ["render_colors", "render_alphas", "render_normals", "render_distort", "render_median"]
    for nm, o in zip(names, out[:5]):
        assert rel_err(o, r[nm]) < 1e-9, nm
    assert torch.equal(out[5].cpu(), r["last_ids"]) and torch.equal(out[6].cpu(), r["median_ids"])
    # the textures contribute (some pairs sample a coarser level than 0)
    plain = rasterize_to_pixels(cpu64(st["means2d"]), cpu64(st["rt"]), colors, opac, cpu64(st["normals"]), W, H, 16,
                                st["offsets"].cpu(), st["flatten_ids"].cpu(), textures=None, backgrounds=bg)
    assert rel_err(r["render_colors"], plain["render_colors"]) > 1e-2
    loss = sum((r[k] * v).sum() for k, v in (("render_colors", vc), ("render_alphas", va), ("render_normals", vn),
                                             ("render_distort", vd), ("render_median", vm)))
    loss.backward()
    got = dict(means2d=res[1], rt=res[2], colors=res[3], opac=res[4], tex=res[5], normals=res[6])
    for k, v in got.items():
        assert rel_err(v, leaves[k].grad) < 1e-8, k
    assert float(res[5].abs().sum()) > 0


def test_mipmapped2_through_the_api():
    """filtering="mipmapped2": the API builds the pyramid and differentiates through it;
    same result as the ops composed by hand."""
    sc = scenes.config1(n=1500, width=64, height=48, tex_res=16, seed=4, sh_degree=None).to(DEV, torch.float64)
    tex = sc.textures.clone().requires_grad_(True)
    out = tg.rasterization_textured_gaussians(sc.means, sc.quats, sc.scales, sc.opacities, sc.colors, tex, sc.viewmats,
                                              sc.Ks, sc.width, sc.height, backgrounds=sc.backgrounds,
                                              filtering="mipmapped2")
    (out[0] * sc.target).sum().backward()
    base = tg.rasterization_textured_gaussians(sc.means, sc.quats, sc.scales, sc.opacities, sc.colors, sc.textures,
                                               sc.viewmats, sc.Ks, sc.width, sc.height, backgrounds=sc.backgrounds)
    assert rel_err(out[0], base[0]) > 1e-4  # a different filter
    assert tex.grad is not None and float(tex.grad.abs().sum()) > 0
    # the texture gradient is the pyramid adjoint of the rasterizer's pyramid gradient
    m = out[-1]
    mip = _ops.generate_mipmap_fwd(sc.textures, 4)
    fwd = _ops.rasterize_to_pixels_fwd_mip2_textured_gaussians(
        m["means2d"], m["ray_transforms"], sc.colors[None].expand(1, -1, -1).contiguous(),
        m["opacities"], mip, 4, 3.0, 3.0, m["normals"], sc.backgrounds, None, sc.width, sc.height, 16,
        m["isect_offsets"], m["flatten_ids"], 0.0, 1.0)
    assert rel_err(fwd[0], out[0]) < 1e-12


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_mip2_whole_frame_matches_oracle(dtype):
    """N = 10,000, 16x16 textures (5 levels), 256x256, every tile: the pyramid built on
    both sides from the base textures, the texture gradient taken back through it.
    Tolerances are the ones stated in gpu_util; fragile pairs include those within
    rounding of a level's edge or of an integer level of detail.  There a flip moves the
    pair's alpha between op*vis*a_s and op*vis, which can move the pixel's median entry to
    another Gaussian, so those pixels (under 1%) get no upstream gradient."""
    sc = scenes.config1()
    st = gpu_stages(sc, dtype)
    s = st["scene"]
    colors = torch.clamp_min(0.28209479177387814 * s.colors[None, :, 0, :] + 0.5, 0.0).contiguous()
    stats = raster_tile_parity(st, colors, s.opacities[None].contiguous(), s.textures, s.backgrounds, dtype,
                               all_tiles(st), filtering="mipmapped2", zero_fragile_cotangents=True)
    assert stats["cotangent_zeroed_pixels"] < 0.01 * 256 * 256
    print("mip2", dtype, stats)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_mip2_packed_matches_oracle_on_every_tile(dtype):
    """mipmapped2 in the packed layout: two cameras, entry i sampling the pyramid of
    Gaussian gaussian_ids[i]; every tile against the oracle (same rules as above; a
    texture row takes the fragility flags of all its cameras' entries, so more cameras
    put more rows under the norm-only rule)."""
    cams = [(3.5, 0.3, 0.4), (-3.0, 1.5, 0.8)]
    sc = scenes.make_scene(2500, 96, 80, 50.0, cams, 16, logscale_mu=-3.0, logscale_sigma=0.4, sh_degree=None,
                           seed=12)
    st = gpu_stages_packed(sc, dtype)
    gids = st["gaussian_ids"]
    nnz, C = gids.numel(), len(cams)
    assert nnz > sc.means.shape[0]
    g = torch.Generator().manual_seed(4)
    colors = torch.rand(nnz, 3, generator=g, dtype=dtype).to(DEV)
    opac = st["scene"].opacities[gids].contiguous()
    bg = torch.rand(C, 3, generator=g, dtype=dtype).to(DEV)
    stats = raster_tile_parity(st, colors, opac, st["scene"].textures, bg, dtype, all_tiles(st),
                               filtering="mipmapped2", gaussian_ids=gids, zero_fragile_cotangents=True)
    print("mip2 packed", dtype, stats)


def test_mipmapped2_packed_through_the_api_equals_grid():
    sc = scenes.make_scene(800, 64, 48, 50.0, [(3.5, 0.3, 0.4), (-3.0, 1.5, 0.8)], 16, logscale_mu=-3.0,
                           logscale_sigma=0.4, sh_degree=1, seed=5).to(DEV, torch.float64)
    res = []
    for packed in (False, True):
        tex = sc.textures.clone().requires_grad_(True)
        m = sc.means.clone().requires_grad_(True)
        out = tg.rasterization_textured_gaussians(m, sc.quats, sc.scales, sc.opacities, sc.colors, tex, sc.viewmats,
                                                  sc.Ks, sc.width, sc.height, sh_degree=1,
                                                  backgrounds=sc.backgrounds, filtering="mipmapped2", packed=packed)
        (out[0] * sc.target).sum().backward()
        res.append((out[0].detach(), tex.grad, m.grad))
    assert rel_err(res[1][0], res[0][0]) < 1e-12
    assert rel_err(res[1][1], res[0][1]) < 1e-10 and rel_err(res[1][2], res[0][2]) < 1e-10
