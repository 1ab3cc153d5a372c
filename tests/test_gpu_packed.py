"""Packed layout (the reference's ``packed=True``): the packed projection is the grid's
visible entries in (camera, Gaussian) order, the packed tile lists are the grid's under
the entry -> (camera, Gaussian) map (bit-exact), renders and gradients through the public
API agree with the grid layout's (float64: to rounding of the atomics' order), and the
packed rasterizer matches the oracle directly on every tile, entry by entry, with entry i
sampling texture row gaussian_ids[i].  Unlike the reference, which indexes the textures
with the entry index there, each packed entry samples its own Gaussian's texture."""
import pytest
import torch

import textured_gaussians_amd as tg
from textured_gaussians_amd import scenes
from textured_gaussians_amd.cuda import _ops

from gpu_util import DEV, all_tiles, gpu_stages_packed, raster_tile_parity, rel_err

pytestmark = pytest.mark.gpu

CAMS = [(3.5, 0.3, 0.4), (-3.0, 1.5, 0.8), (0.5, -3.6, -0.6)]


def _scene(dtype, n=700, tex=10, sh_degree=1, seed=3):
    return scenes.make_scene(n, 64, 48, 50.0, CAMS, tex, logscale_mu=-3.0, logscale_sigma=0.4,
                             sh_degree=sh_degree, seed=seed).to(DEV, dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_packed_projection_is_the_visible_grid_entries(dtype):
    s = _scene(dtype)
    args = (s.means, s.quats, s.scales, s.viewmats, s.Ks, s.width, s.height)
    radii, m2, dep, rt, nrm = _ops.fully_fused_projection_fwd_2dgs(*args, 0.3, 0.01, 1e10, 0.0)
    indptr, cid, gid, pr, pm2, pdep, prt, pnrm = _ops.fully_fused_projection_packed_fwd_2dgs(*args, 0.01, 1e10, 0.0)
    vis = (radii > 0).nonzero()
    assert vis.shape[0] > 0 and torch.equal(cid, vis[:, 0]) and torch.equal(gid, vis[:, 1])
    C = s.viewmats.shape[0]
    counts = (radii > 0).sum(1)
    assert torch.equal(indptr, torch.cat([torch.zeros(1, device=DEV, dtype=torch.int64), counts.cumsum(0)]))
    for a, b in ((pr, radii), (pm2, m2), (pdep, dep), (prt, rt), (pnrm, nrm)):
        assert torch.equal(a, b[cid, gid])
    assert indptr.numel() == C + 1


def test_packed_tile_lists_are_the_grid_lists():
    s = _scene(torch.float32)
    N = s.means.shape[0]
    C = s.viewmats.shape[0]
    args = (s.means, s.quats, s.scales, s.viewmats, s.Ks, s.width, s.height)
    radii, m2, dep, rt, nrm = _ops.fully_fused_projection_fwd_2dgs(*args, 0.3, 0.01, 1e10, 0.0)
    _, cid, gid, pr, pm2, pdep, _, _ = _ops.fully_fused_projection_packed_fwd_2dgs(*args, 0.01, 1e10, 0.0)
    tw, th = 4, 3
    tpg, ids, flat = _ops.isect_tiles(m2, radii, dep, None, None, C, 16, tw, th, True, True)
    ptpg, pids, pflat = _ops.isect_tiles(pm2, pr, pdep, cid, gid, C, 16, tw, th, True, True)
    assert ids.numel() > 100
    assert torch.equal(pids, ids)
    assert torch.equal((cid * N + gid)[pflat.long()].int(), flat)
    assert torch.equal(ptpg, tpg[cid, gid])


def _render(s, L, packed, textured=True, **kw):
    bg = s.backgrounds
    if textured:
        return tg.rasterization_textured_gaussians(L["means"], L["quats"], L["scales"], L["opacities"], L["colors"],
                                                   L["textures"], s.viewmats, s.Ks, s.width, s.height,
                                                   sh_degree=s.sh_degree, backgrounds=bg, packed=packed, **kw)
    return tg.rasterization_2dgs(L["means"], L["quats"], L["scales"], L["opacities"], L["colors"], s.viewmats, s.Ks,
                                 s.width, s.height, sh_degree=s.sh_degree, backgrounds=bg, packed=packed, **kw)


@pytest.mark.parametrize("textured", [True, False])
@pytest.mark.parametrize("grad_mode,sparse", [("reference", False), ("exact", False), ("reference", True)])
def test_packed_render_and_gradients_match_grid(textured, grad_mode, sparse):
    s = _scene(torch.float64)
    names = ["means", "quats", "scales", "opacities", "colors"] + (["textures"] if textured else [])
    res = {}
    for packed in (False, True):
        L = {k: getattr(s, k).detach().clone().requires_grad_(True) for k in names}
        out = _render(s, L, packed, textured, grad_mode=grad_mode, sparse_grad=sparse and packed)
        g = torch.Generator(device=DEV).manual_seed(1)
        loss = (out[0] * torch.rand(out[0].shape, device=DEV, dtype=out[0].dtype, generator=g)).sum() + \
            out[1].sum() + (out[2] * 0.3).sum()
        loss.backward()
        grads = {k: (v.grad.to_dense() if v.grad.is_sparse else v.grad) for k, v in L.items()}
        res[packed] = (out, grads)
    (o0, g0), (o1, g1) = res[False], res[True]
    for i in (0, 1, 2, 4, 5, 6, 7):
        assert rel_err(o1[i], o0[i]) < 1e-12, i
    for k in g0:
        assert rel_err(g1[k], g0[k]) < 1e-10, k


@pytest.mark.parametrize("kind", ["rgb", "per_view"])
def test_packed_plain_and_per_view_colours_match_grid(kind):
    """Colours without SH, [N, 3] or per view [C, N, 3]: the packed layout gathers
    them per entry (and scatters their gradients back), the grid layout expands them."""
    s = _scene(torch.float64)
    C, N = s.viewmats.shape[0], s.means.shape[0]
    g = torch.Generator().manual_seed(5)
    cols = torch.rand((N, 3) if kind == "rgb" else (C, N, 3), generator=g, dtype=torch.float64).to(DEV)
    res = {}
    for packed in (False, True):
        L = {k: getattr(s, k).detach().clone().requires_grad_(True)
             for k in ["means", "quats", "scales", "opacities", "textures"]}
        L["colors"] = cols.clone().requires_grad_(True)
        out = tg.rasterization_textured_gaussians(L["means"], L["quats"], L["scales"], L["opacities"], L["colors"],
                                                  L["textures"], s.viewmats, s.Ks, s.width, s.height,
                                                  backgrounds=s.backgrounds, packed=packed)
        gw = torch.Generator(device=DEV).manual_seed(1)
        (out[0] * torch.rand(out[0].shape, device=DEV, dtype=out[0].dtype, generator=gw)).sum().backward()
        res[packed] = (out[0].detach(), {k: v.grad for k, v in L.items()})
    (o0, g0), (o1, g1) = res[False], res[True]
    assert rel_err(o1, o0) < 1e-12
    for k in g0:
        assert rel_err(g1[k], g0[k]) < 1e-10, k


def test_packed_textures_sample_each_entrys_gaussian():
    """A Gaussian visible only in the second camera: its packed entry index differs from its
    Gaussian id, and its texture (not the texture at its entry index) must colour it."""
    s = _scene(torch.float64, n=300)
    L = {k: getattr(s, k).detach().clone() for k in ["means", "quats", "scales", "opacities", "colors", "textures"]}
    a = _render(s, L, False)[0]
    b = _render(s, L, True)[0]
    assert rel_err(b, a) < 1e-12
    L["textures"] = L["textures"].flip(0).contiguous()  # any mix-up of rows now shows
    a2 = _render(s, L, False)[0]
    b2 = _render(s, L, True)[0]
    assert rel_err(b2, a2) < 1e-12 and rel_err(a2, a) > 1e-3


def test_packed_viewmat_gradient_matches_grid():
    s = _scene(torch.float64)
    res = []
    for packed in (False, True):
        vm = s.viewmats.clone().requires_grad_(True)
        out = tg.rasterization_textured_gaussians(s.means, s.quats, s.scales, s.opacities, s.colors, s.textures, vm,
                                                  s.Ks, s.width, s.height, sh_degree=s.sh_degree,
                                                  backgrounds=s.backgrounds, packed=packed, grad_mode="exact")
        out[0].sum().backward()
        res.append(vm.grad)
    assert float(res[0].abs().max()) > 0 and rel_err(res[1], res[0]) < 1e-10


@pytest.mark.parametrize("dtype,textured,uv", [(torch.float64, True, "none"), (torch.float64, True, "exact"),
                                               (torch.float32, True, "none"), (torch.float32, True, "double"),
                                               (torch.float64, False, "none"), (torch.float32, False, "exact")])
def test_packed_rasterizer_matches_oracle_on_every_tile(dtype, textured, uv):
    """The packed rasterizer against the oracle over every tile of three cameras: the
    oracle's compact problem holds the packed entries the tile lists reference, each with
    its Gaussian's texture row; a Gaussian seen by several cameras has one texture row,
    whose gradient sums its entries'.  Tolerances are the ones stated in gpu_util."""
    sc = scenes.make_scene(3000, 96, 80, 50.0, CAMS, 10, logscale_mu=-3.2, logscale_sigma=0.4, sh_degree=None,
                           seed=7)
    st = gpu_stages_packed(sc, dtype)
    gids = st["gaussian_ids"]
    nnz, N, C = gids.numel(), sc.means.shape[0], len(CAMS)
    # entries of one Gaussian in several cameras, and entry indices that are not Gaussian ids
    assert nnz > 1.5 * N and bool((gids != torch.arange(nnz, device=DEV) % N).any())
    g = torch.Generator().manual_seed(3)
    colors = torch.rand(nnz, 3, generator=g, dtype=dtype).to(DEV)
    opac = st["scene"].opacities[gids].contiguous()
    bg = torch.rand(C, 3, generator=g, dtype=dtype).to(DEV)
    tex = st["scene"].textures if textured else None
    stats = raster_tile_parity(st, colors, opac, tex, bg, dtype, all_tiles(st), uv=uv, gaussian_ids=gids,
                               n_gaussians=N)
    print("packed", dtype, textured, uv, stats)


def test_packed_gaussian_ids_out_of_range_are_refused():
    """An entry naming a Gaussian past the texture rows (or a negative one) is refused on
    the host before any kernel runs."""
    sc = scenes.make_scene(400, 64, 48, 50.0, CAMS, 8, logscale_mu=-3.0, logscale_sigma=0.4, sh_degree=None, seed=1)
    st = gpu_stages_packed(sc, torch.float32)
    gids = st["gaussian_ids"]
    nnz, N = gids.numel(), sc.means.shape[0]
    colors = torch.rand(nnz, 3, device=DEV)
    opac = st["scene"].opacities[gids].contiguous()
    for bad in (N, -1):
        ids = gids.clone()
        ids[nnz // 2] = bad
        with pytest.raises(RuntimeError, match="gaussian_ids must lie in"):
            _ops._raster_fwd(True, st["means2d"], st["rt"], colors, opac, st["scene"].textures, 3.0, 3.0,
                             st["normals"], None, None, 64, 48, 16, st["offsets"], st["flatten_ids"], 0.0, 1.0,
                             gaussian_ids=ids)
    # in range: runs
    _ops._raster_fwd(True, st["means2d"], st["rt"], colors, opac, st["scene"].textures, 3.0, 3.0, st["normals"],
                     None, None, 64, 48, 16, st["offsets"], st["flatten_ids"], 0.0, 1.0, gaussian_ids=gids)
    torch.cuda.synchronize()
