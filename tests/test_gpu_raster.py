"""Rasterizer forward/backward on the GPU vs the oracle, fed identical inputs
(the GPU's projection outputs and tile lists).

Tolerances: float32 kernels - rendered colour/alpha/normal within 1e-4 max-abs,
gradients within 1e-3 relative (max |d| / max |ref| per tensor); float64
kernels - 1e-9, i.e. the same arithmetic as the oracle."""
import pytest
import torch

from oracle.raster import rasterize_to_pixels
from textured_gaussians_amd import scenes
from textured_gaussians_amd.cuda import _ops

from gpu_util import DEV, cpu64, gpu_stages, rel_err

pytestmark = pytest.mark.gpu

UV = {"none": _ops.rasterize_to_pixels_bwd_textured_gaussians,
      "exact": _ops.rasterize_to_pixels_bwd_exact_textured_gaussians,
      "double": _ops.rasterize_to_pixels_bwd2_textured_gaussians}


def _inputs(dtype, K=3, n=2500, w=80, h=64, tex=12, seed=0):
    sc = scenes.config1(n=n, width=w, height=h, tex_res=tex, seed=seed, sh_degree=None)
    st = gpu_stages(sc, dtype)
    C, N = st["radii"].shape
    g = torch.Generator().manual_seed(seed + 11)
    colors = torch.rand(C, N, K, generator=g, dtype=torch.float64)
    opac = sc.opacities.double()[None].repeat(C, 1)
    bg = torch.rand(C, K, generator=g, dtype=torch.float64)
    return sc, st, colors, opac, bg


def _oracle_fwd(st, colors, opac, textures, bg, uv="exact", g_weight=1.0, thr=0.0, distort="exact"):
    s = st["scene"]
    return rasterize_to_pixels(cpu64(st["means2d"]), cpu64(st["rt"]), colors, opac, cpu64(st["normals"]), s.width,
                               s.height, st["tile_size"], st["offsets"].cpu(), st["flatten_ids"].cpu(),
                               textures=textures, backgrounds=bg, g_weight=g_weight, gs_contrib_threshold=thr,
                               uv_grad=uv, distort_grad=distort)


def _gpu_fwd(st, colors, opac, textures, bg, dtype, g_weight=1.0, thr=0.0):
    s = st["scene"]
    dv = lambda t: None if t is None else t.to(DEV, dtype).contiguous()  # noqa: E731
    if textures is not None:
        return _ops.rasterize_to_pixels_fwd_textured_gaussians_full(
            st["means2d"], st["rt"], dv(colors), dv(opac), dv(textures), 3.0, 3.0, st["normals"], dv(bg), None,
            s.width, s.height, st["tile_size"], st["offsets"], st["flatten_ids"], thr, g_weight)
    return _ops.rasterize_to_pixels_fwd_2dgs_full(st["means2d"], st["rt"], dv(colors), dv(opac), st["normals"],
                                                  dv(bg), None, s.width, s.height, st["tile_size"], st["offsets"],
                                                  st["flatten_ids"], thr)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("textured", [True, False])
@pytest.mark.parametrize("K", [1, 3, 4])
@pytest.mark.parametrize("size", [(80, 64), (75, 53)])  # the second is ragged: partial edge tiles
def test_raster_fwd(dtype, textured, K, size):
    sc, st, colors, opac, bg = _inputs(dtype, K=K, w=size[0], h=size[1])
    tex = sc.textures.double() if textured else None
    out = _gpu_fwd(st, colors, opac, tex, bg, dtype)
    ref = _oracle_fwd(st, colors, opac, tex, bg)
    names = ["render_colors", "render_alphas", "render_normals", "render_distort", "render_median"]
    tol = 1e-4 if dtype == torch.float32 else 1e-9
    # float32: compare away from pixels with an entry within rounding margin of a
    # discrete decision (those must stay rare); float64: every pixel
    keep = torch.ones_like(ref["fragile"])
    if dtype == torch.float32:
        # exclude pixels that are ill-conditioned in float32 arithmetic itself: an entry
        # within rounding margin of a discrete decision, or the float32 oracle departing
        # from the float64 oracle by more than a tenth of the tolerance
        r32 = rasterize_to_pixels(*[t.float() if torch.is_tensor(t) and t.is_floating_point() else t for t in (
            cpu64(st["means2d"]), cpu64(st["rt"]), colors, opac, cpu64(st["normals"]))], st["scene"].width,
            st["scene"].height, st["tile_size"], st["offsets"].cpu(), st["flatten_ids"].cpu(),
            textures=None if tex is None else tex.float(), backgrounds=bg.float())
        dev32 = (r32["render_colors"].double() - ref["render_colors"]).abs().amax(-1)
        dev32 = torch.maximum(dev32, (r32["render_alphas"].double() - ref["render_alphas"]).abs()[..., 0])
        keep = ~ref["fragile"] & (dev32 < 5e-5)
    assert keep.float().mean().item() > 0.99
    for nm, o in zip(names, out[:5]):
        d = (cpu64(o) - ref[nm]).abs().amax(-1) if o.dim() == 4 else (cpu64(o) - ref[nm]).abs()
        bad = (d > tol) & keep
        assert not bad.any(), (nm, d[keep].max().item(), int(bad.sum()), bad.nonzero()[:4].tolist(),
                               ref["fragile"].float().mean().item())
    assert torch.equal(out[5].cpu()[keep], ref["last_ids"][keep])
    assert torch.equal(out[6].cpu()[keep], ref["median_ids"][keep])
    assert torch.equal(out[8].cpu().double(), ref["contrib_count"]) or dtype == torch.float32
    if dtype == torch.float32:
        # a pair whose float32 alpha lands on the other side of 1/255 moves one Gaussian's sum by
        # ~0.004, so the per-Gaussian sums are compared in the L2 norm
        cs = cpu64(out[7])
        assert (cs - ref["contrib_sum"]).norm() / ref["contrib_sum"].norm() < 1e-4
    else:
        assert rel_err(out[7], ref["contrib_sum"]) < 1e-10


def test_raster_fwd_contrib_threshold_and_gweight():
    sc, st, colors, opac, bg = _inputs(torch.float64)
    tex = sc.textures.double()
    out = _gpu_fwd(st, colors, opac, tex, bg, torch.float64, g_weight=0.7, thr=0.3)
    ref = _oracle_fwd(st, colors, opac, tex, bg, g_weight=0.7, thr=0.3)
    assert (cpu64(out[0]) - ref["render_colors"]).abs().max().item() < 1e-9
    assert torch.equal(out[8].cpu().double(), ref["contrib_count"])


def _bwd_case(dtype, uv, textured=True, K=3, use_bg=True, seed=0, size=(80, 64)):
    sc, st, colors, opac, bg = _inputs(dtype, K=K, seed=seed, w=size[0], h=size[1])
    if not use_bg:
        bg = None
    s = st["scene"]
    tex = sc.textures.double() if textured else None
    out = _gpu_fwd(st, colors, opac, tex, bg, dtype)
    C, H, W = st["C"], s.height, s.width
    g = torch.Generator().manual_seed(seed + 7)
    vc = torch.randn(C, H, W, K, generator=g, dtype=torch.float64)
    va = torch.randn(C, H, W, 1, generator=g, dtype=torch.float64)
    vn = torch.randn(C, H, W, 3, generator=g, dtype=torch.float64)
    vd = torch.randn(C, H, W, 1, generator=g, dtype=torch.float64)
    vm = torch.randn(C, H, W, 1, generator=g, dtype=torch.float64)
    dv = lambda t: None if t is None else t.to(DEV, dtype).contiguous()  # noqa: E731
    densify = torch.zeros_like(st["means2d"])
    args = [st["means2d"], st["rt"], dv(colors), dv(opac)]
    if textured:
        args += [dv(tex), 3.0, 3.0]
    args += [st["normals"], densify, dv(bg), None, W, H, st["tile_size"], st["offsets"], st["flatten_ids"]]
    if textured:
        args += [1.0]
    args += [out[0], out[1], out[5], out[6], dv(vc), dv(va), dv(vn), dv(vd), dv(vm), True]
    exact = uv == "exact"
    accd = out[9] if exact else None
    if textured:
        res = UV[uv](*args, render_accum_depth=accd)
    else:
        res = _ops.rasterize_to_pixels_bwd_2dgs(*args, render_accum_depth=accd)
    # oracle: autograd of the restated forward
    leaves = dict(means2d=cpu64(st["means2d"]).requires_grad_(True), rt=cpu64(st["rt"]).requires_grad_(True),
                  colors=colors.clone().requires_grad_(True), opac=opac.clone().requires_grad_(True),
                  normals=cpu64(st["normals"]).requires_grad_(True))
    if textured:
        leaves["tex"] = tex.clone().requires_grad_(True)
    pairs = []

    def oracle_grads(lv, pm, dt):
        r = rasterize_to_pixels(lv["means2d"], lv["rt"], lv["colors"], lv["opac"], lv["normals"],
                                W, H, st["tile_size"], st["offsets"].cpu(), st["flatten_ids"].cpu(),
                                textures=lv.get("tex"), backgrounds=None if bg is None else bg.to(dt), uv_grad=uv,
                                pair_means=pm, distort_grad="exact" if exact else "reference")
        loss = sum((r[k] * v.to(dt)).sum() for k, v in (("render_colors", vc), ("render_alphas", va),
                                                          ("render_normals", vn), ("render_distort", vd),
                                                          ("render_median", vm)))
        loss.backward()
        return r

    ref = oracle_grads(leaves, pairs, torch.float64)
    absg = torch.zeros_like(leaves["means2d"]).reshape(-1, 2)
    for c, gid, mp in pairs:
        if mp.grad is not None:
            absg.index_add_(0, gid, mp.grad.abs().sum(0))
    if textured:
        v_abs, v_m2, v_rt, v_col, v_op, v_tex, v_nrm, v_dens = res
    else:
        v_abs, v_m2, v_rt, v_col, v_op, v_nrm, v_dens = res
        v_tex = None
    got = dict(means2d=v_m2, rt=v_rt, colors=v_col, opac=v_op, normals=v_nrm)
    if textured:
        got["tex"] = v_tex
    # float32: relative error in the L2 norm (single ill-conditioned pairs, e.g. splats
    # seen edge-on, dominate the max); float64: max-abs relative
    tol = 1e-3 if dtype == torch.float32 else 1e-8
    # float32 with the exact texture-coordinate rule: rows of Gaussians whose splat
    # coordinate is ill-conditioned somewhere (cond > 1e3, edge-on splats) are left out
    # of the geometry gradients; they must be rare
    well = torch.ones(st["C"], st["radii"].shape[1], dtype=torch.bool)
    if dtype == torch.float32 and uv == "exact":
        # ... and rows with a kept pair at a texel edge, where the exact bilinear derivative
        # jumps and a float32 coordinate may fall in the neighbouring cell, and rows where the
        # same formula evaluated in float32 on the CPU departs from float64 by more than a
        # tenth of the tolerance (a float32 cancellation in one dominant pair)
        l32 = {k: v.detach().float().clone().requires_grad_(True) for k, v in leaves.items()}
        oracle_grads(l32, None, torch.float32)
        dev = torch.zeros_like(well, dtype=torch.float64)
        for k in ("rt", "means2d"):
            d = (l32[k].grad.double() - leaves[k].grad).reshape(*well.shape, -1).norm(dim=-1)
            dev = torch.maximum(dev, d / max(leaves[k].grad.norm().item(), 1e-30))
        well = (ref["cond_max"] < 1e3) & ~ref["texel_edge"] & (dev < 1e-4)
        assert well.float().mean().item() > 0.95
    for k, v in got.items():
        ref_g = leaves[k].grad
        vv = cpu64(v)
        if k in ("rt", "means2d"):
            vv, ref_g = vv[well], ref_g[well]
        if dtype == torch.float32:
            e = (vv - ref_g).norm().item() / max(ref_g.norm().item(), 1e-30)
        else:
            e = rel_err(vv, ref_g)
        assert e < tol, (k, e)
    if dtype == torch.float32:
        assert (cpu64(v_abs).reshape(-1, 2) - absg).norm() / absg.norm() < tol
    else:
        assert rel_err(v_abs.reshape(-1, 2), absg) < tol
    exp_dens = torch.stack([cpu64(v_rt)[..., 0, 2], cpu64(v_rt)[..., 1, 2]], -1) * cpu64(st["rt"])[..., 2, 2, None]
    assert rel_err(v_dens, exp_dens) < (1e-6 if dtype == torch.float32 else 1e-12)


@pytest.mark.parametrize("uv", ["none", "exact", "double"])
def test_raster_bwd_textured_f64(uv):
    _bwd_case(torch.float64, uv)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_raster_bwd_textured_ragged_image(dtype):
    _bwd_case(dtype, "none", size=(75, 53))


@pytest.mark.parametrize("uv", ["none", "exact"])
def test_raster_bwd_textured_f32(uv):
    _bwd_case(torch.float32, uv)


@pytest.mark.parametrize("uv,textured,K,use_bg", [("none", True, 3, True), ("double", True, 3, True),
                                                  ("exact", True, 3, True), ("none", True, 2, False),
                                                  ("none", False, 3, True), ("exact", False, 4, True)])
def test_raster_bwd_f32_per_element(uv, textured, K, use_bg):
    """The float32 scenes of the tests above, element by element over every tile of the
    frame instead of in the L2 norm: every gradient element within 1e-3 relative (with a
    floor of 1e-4 of its row's largest element, plus three times the float32 oracle's own
    error on it), the rows whose pairs sit on float32 decision thresholds (1/255, the
    texture-range edge, the 1e-4 stop, the median) at the looser bound stated in gpu_util,
    the background gradient too when there is a background.  K = 1 stays with the float64
    kernels (test_raster_bwd_channels_nobg): its colour is the depth channel itself, a
    one-element row whose value can cancel to a few parts in 1e4 of its distortion terms,
    which the backward rebuilds by subtraction from the forward's totals in float32 (seen:
    3.140e-5 against 3.162e-5, the float32 oracle 3.164e-5), so the row floor is no floor."""
    from gpu_util import all_tiles, raster_tile_parity
    sc, st, colors, opac, bg = _inputs(torch.float32, K=K)
    dv = lambda t: t.to(DEV, torch.float32).contiguous()  # noqa: E731
    tex = dv(sc.textures) if textured else None
    stats = raster_tile_parity(st, dv(colors), dv(opac), tex, dv(bg) if use_bg else None, torch.float32,
                               all_tiles(st), uv=uv)
    print("per element", uv, textured, K, use_bg, stats)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("rule", ["none", "exact"])
def test_raster_bwd_plain(dtype, rule):
    _bwd_case(dtype, rule, textured=False)


@pytest.mark.parametrize("K", [1, 4])
def test_raster_bwd_channels_nobg(K):
    _bwd_case(torch.float64, "exact", K=K, use_bg=False, seed=3)


@pytest.mark.parametrize("textured", [True, False])
@pytest.mark.parametrize("tile_size", [16, 8])
def test_live_entries_skip_only_unused_entries(textured, tile_size):
    """The forward records, per wavefront and list entry, the mask of the wave's lanes
    that composited the entry; the backward then visits only entries with a nonzero
    mask.  Every pixel's last composited entry has the pixel's bit set in its wave's
    mask, and the backward over live entries gives the full walk's gradients (float64,
    so only the atomics' summation order differs)."""
    dtype = torch.float64
    sc = scenes.config1(n=2500, width=80, height=64, tex_res=12, seed=3, sh_degree=None)
    st = gpu_stages(sc, dtype, tile_size=tile_size)
    s = st["scene"]
    C, N = st["radii"].shape
    g = torch.Generator().manual_seed(4)
    colors = torch.rand(C, N, 3, generator=g, dtype=torch.float64).to(DEV)
    opac = sc.opacities.double()[None].repeat(C, 1).to(DEV)
    tex = sc.textures.to(DEV, dtype) if textured else None
    flat = st["flatten_ids"]
    waves = (tile_size * tile_size + 63) // 64
    live = torch.zeros(waves, flat.numel(), dtype=torch.int64, device=DEV)
    if textured:
        fwd = _ops.rasterize_to_pixels_fwd_textured_gaussians_full(
            st["means2d"], st["rt"], colors, opac, tex, 3.0, 3.0, st["normals"], None, None, s.width, s.height,
            tile_size, st["offsets"], flat, 0.0, 1.0, live_entries=live)
    else:
        fwd = _ops.rasterize_to_pixels_fwd_2dgs_full(st["means2d"], st["rt"], colors, opac, st["normals"], None,
                                                     None, s.width, s.height, tile_size, st["offsets"], flat, 0.0,
                                                     live_entries=live)
    lv = live.cpu()
    assert bool((lv != 0).any()) and bool((lv == 0).any())
    alpha, last = fwd[1][0, ..., 0].cpu(), fwd[5][0].cpu()
    ii, jj = torch.meshgrid(torch.arange(s.height), torch.arange(s.width), indexing="ij")
    pix = (ii % tile_size) * tile_size + jj % tile_size  # pixel of its tile
    wave, bit = pix // 64, pix % 64
    hit = alpha > 0
    assert hit.any() and bool((((lv[wave[hit], last[hit].long()] >> bit[hit]) & 1) == 1).all())
    # a pixel's bit is set only for entries up to its last composited one
    w_, i_ = torch.nonzero(lv != 0, as_tuple=True)
    offs = st["offsets"].reshape(-1).cpu().long()
    tile = torch.searchsorted(offs, i_, right=True) - 1
    tw = (s.width + tile_size - 1) // tile_size
    for b in range(64):
        sel = ((lv[w_, i_] >> b) & 1) == 1
        p = w_[sel] * 64 + b
        y = (tile[sel] // tw) * tile_size + p // tile_size
        x = (tile[sel] % tw) * tile_size + p % tile_size
        assert bool(((y < s.height) & (x < s.width)).all())
        assert bool((last[y, x] >= i_[sel]).all())

    gen = torch.Generator().manual_seed(5)
    H, W = s.height, s.width
    vc = torch.randn(C, H, W, 3, generator=gen, dtype=torch.float64).to(DEV)
    va = torch.randn(C, H, W, 1, generator=gen, dtype=torch.float64).to(DEV)
    vn = torch.randn(C, H, W, 3, generator=gen, dtype=torch.float64).to(DEV)
    vd = torch.randn(C, H, W, 1, generator=gen, dtype=torch.float64).to(DEV)
    vm = torch.randn(C, H, W, 1, generator=gen, dtype=torch.float64).to(DEV)
    densify = torch.zeros_like(st["means2d"])
    args = [st["means2d"], st["rt"], colors, opac]
    if textured:
        args += [tex, 3.0, 3.0]
    args += [st["normals"], densify, None, None, W, H, tile_size, st["offsets"], flat]
    if textured:
        args += [1.0]
    args += [fwd[0], fwd[1], fwd[5], fwd[6], vc, va, vn, vd, vm, False]
    bwd = _ops.rasterize_to_pixels_bwd_exact_textured_gaussians if textured else _ops.rasterize_to_pixels_bwd_2dgs
    full = bwd(*args, render_accum_depth=fwd[9])
    part = bwd(*args, render_accum_depth=fwd[9], live_entries=live)
    for a, b in zip(full[1:], part[1:]):
        assert rel_err(b, a) < 1e-12


@pytest.mark.parametrize("dtype,tex_res", [(torch.float32, 12), (torch.float64, 12), (torch.float32, 130)])
def test_texture_gradient_records_match_direct_accumulation(dtype, tex_res):
    """Texture gradients through per-pair records and the per-row LDS gather equal
    direct accumulation (the forward's pair counts size the records); rows absent
    from the tile lists end zero.  130x130 float32 rows (270 KB) exceed the gather's
    LDS and take the direct path inside the same call."""
    sc = scenes.config1(n=2500, width=80, height=64, tex_res=tex_res, seed=6, sh_degree=None)
    st = gpu_stages(sc, dtype)
    s = st["scene"]
    C, N = st["radii"].shape
    g = torch.Generator().manual_seed(8)
    colors = torch.rand(C, N, 3, generator=g, dtype=torch.float64).to(DEV, dtype)
    opac = sc.opacities.double()[None].repeat(C, 1).to(DEV, dtype)
    tex = sc.textures.to(DEV, dtype)
    flat = st["flatten_ids"]
    H, W = s.height, s.width
    fargs = (st["means2d"], st["rt"], colors, opac, tex, 3.0, 3.0, st["normals"], None, None, W, H, 16,
             st["offsets"], flat, 0.0, 1.0)
    cnt = _ops.tex_counts_buffer(tex)
    fwd = _ops.rasterize_to_pixels_fwd_textured_gaussians_full(*fargs, tex_counts=cnt)
    assert int(cnt.sum()) > 0
    vc = torch.randn(C, H, W, 3, generator=g, dtype=torch.float64).to(DEV, dtype)
    va = torch.randn(C, H, W, 1, generator=g, dtype=torch.float64).to(DEV, dtype)
    z3 = torch.zeros(C, H, W, 3, device=DEV, dtype=dtype)
    z1 = torch.zeros(C, H, W, 1, device=DEV, dtype=dtype)
    bargs = (st["means2d"], st["rt"], colors, opac, tex, 3.0, 3.0, st["normals"], torch.zeros_like(st["means2d"]),
             None, None, W, H, 16, st["offsets"], flat, 1.0, fwd[0], fwd[1], fwd[5], fwd[6], vc, va, z3, z1, z1,
             False)
    direct = _ops.rasterize_to_pixels_bwd_exact_textured_gaussians(*bargs)
    mask = _ops.mark_rows(flat, N)
    vt = torch.full_like(tex, float("nan"))  # every row must be written by the call
    rec = _ops.rasterize_to_pixels_bwd_exact_textured_gaussians(*bargs, v_textures_out=vt, zero_rows=(vt, mask),
                                                                tex_counts=cnt)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert bool(torch.isfinite(rec[5]).all())
    for a, b in zip(direct[1:], rec[1:]):
        assert rel_err(b, a) < tol
    assert float(rec[5][mask == 0].abs().max()) == 0.0
    # without a row mask the gather adds each row's sums to the given (zeroed) gradient
    va = torch.zeros_like(tex)
    acc = _ops.rasterize_to_pixels_bwd_exact_textured_gaussians(*bargs, v_textures_out=va, tex_counts=cnt)
    assert rel_err(acc[5], direct[5]) < tol


def _fwd_pair(st, colors, opac, tex, monkeypatch, tile_size):
    """The forward with and without the candidate pre-pass; the composited-lane masks
    start as all ones so entries a wave never reached compare equal too."""
    s = st["scene"]
    waves = (tile_size * tile_size + 63) // 64
    res = {}
    for cull in (True, False):
        monkeypatch.setattr(_ops, "FWD_CULL", cull)
        live = torch.full((waves, st["flatten_ids"].numel()), -1, dtype=torch.int64, device=DEV)
        if tex is not None:
            cnt = _ops.tex_counts_buffer(tex)
            fo = _ops.rasterize_to_pixels_fwd_textured_gaussians_full(
                st["means2d"], st["rt"], colors, opac, tex, 3.0, 3.0, st["normals"], None, None, s.width, s.height,
                tile_size, st["offsets"], st["flatten_ids"], 0.0, 1.0, live_entries=live, tex_counts=cnt)
        else:
            cnt = None
            fo = _ops.rasterize_to_pixels_fwd_2dgs_full(st["means2d"], st["rt"], colors, opac, st["normals"], None,
                                                        None, s.width, s.height, tile_size, st["offsets"],
                                                        st["flatten_ids"], 0.0, live_entries=live)
        res[cull] = (fo, live, cnt)
    return res


def assert_prepass_exact(res):
    (fa, la, ca), (fb, lb, cb) = res[True], res[False]
    for i, (x, y) in enumerate(zip(fa, fb)):
        if i == 7:  # contribution sums: float atomics from several waves, in any order
            assert rel_err(x, y) < 1e-6
        else:
            assert torch.equal(x, y), i
    assert torch.equal(la, lb)
    if ca is not None:
        assert torch.equal(ca, cb)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("textured", [True, False])
@pytest.mark.parametrize("tile_size", [16, 8])
def test_forward_candidate_prepass_is_exact(dtype, textured, tile_size, monkeypatch):
    """The forward's candidate pre-pass skips only list entries no pixel of a wave can
    keep: every output, the composited-lane masks and the texture pair counts are bit-identical
    to the walk over every entry (the contribution sums to float atomic ordering)."""
    sc = scenes.config1(n=2500, width=80, height=64, tex_res=12, seed=3, sh_degree=None)
    st = gpu_stages(sc, dtype, tile_size=tile_size)
    C, N = st["radii"].shape
    g = torch.Generator().manual_seed(4)
    colors = torch.rand(C, N, 3, generator=g, dtype=torch.float64).to(DEV, dtype)
    opac = sc.opacities.double()[None].repeat(C, 1).to(DEV, dtype)
    tex = sc.textures.to(DEV, dtype) if textured else None
    assert_prepass_exact(_fwd_pair(st, colors, opac, tex, monkeypatch, tile_size))


@pytest.mark.parametrize("tile_size", [4, 12, 24, 32])
def test_raster_any_tile_size(tile_size):
    """Tile sizes other than 8 and 16 (the reference launches tile_size^2 threads per
    tile, so any size up to 32): forward and every gradient of the whole frame against
    the oracle in float64 (1e-9 / 1e-8), through the autograd path."""
    from gpu_util import all_tiles, raster_tile_parity
    sc = scenes.config1(n=1500, width=70, height=53, tex_res=10, seed=8, sh_degree=None)
    st = gpu_stages(sc, torch.float64, tile_size=tile_size)
    C, N = st["radii"].shape
    g = torch.Generator().manual_seed(9)
    colors = torch.rand(C, N, 3, generator=g, dtype=torch.float64).to(DEV)
    opac = sc.opacities.double()[None].repeat(C, 1).to(DEV)
    s = st["scene"]
    raster_tile_parity(st, colors, opac, s.textures, s.backgrounds, torch.float64, all_tiles(st), uv="exact")


@pytest.mark.parametrize("uv", ["none", "exact"])
def test_saturated_frame_f32(uv):
    """A frame where most pixels composite down to the 1e-4 transmittance stop (large,
    nearly opaque splats with opaque textures): the float32 backward, which walks T back
    by division from the forward's final T, against the oracle on every tile.  Tolerances
    are the ones stated in gpu_util."""
    from gpu_util import all_tiles, raster_tile_parity
    sc = scenes.config1(n=6000, width=128, height=128, tex_res=12, seed=8, sh_degree=None, logscale_mu=-2.4)
    sc.opacities = 0.75 + 0.24 * sc.opacities
    sc.textures[..., 3] = 0.7 + 0.3 * sc.textures[..., 3]
    st = gpu_stages(sc, torch.float32)
    C, N = st["radii"].shape
    g = torch.Generator().manual_seed(2)
    colors = torch.rand(C, N, 3, generator=g).to(DEV)
    opac = st["scene"].opacities[None].contiguous()
    s = st["scene"]
    # most pixels end at the stop: final T below 1e-3
    fo = _ops.rasterize_to_pixels_fwd_textured_gaussians_full(
        st["means2d"], st["rt"], colors, opac, s.textures, 3.0, 3.0, st["normals"], s.backgrounds, None, s.width,
        s.height, 16, st["offsets"], st["flatten_ids"], 0.0, 1.0)
    assert float((fo[1] > 0.999).float().mean()) > 0.6
    fo = None
    stats = raster_tile_parity(st, colors, opac, s.textures, s.backgrounds, torch.float32, all_tiles(st), uv=uv)
    print("saturated", uv, stats)


def test_texel_marks_cover_the_gradient_and_change_nothing():
    """The measurement hook that counts distinct texels read (tgs_set_texel_marks): the
    forward's outputs are bit-identical with it set, every texel with a nonzero gradient
    is marked, and only rows in the tile lists are."""
    from textured_gaussians_amd.cuda import _wrapper
    sc = scenes.config1(n=3000, width=96, height=80, tex_res=12, seed=4, sh_degree=None)
    st = gpu_stages(sc, torch.float32)
    C, N = st["radii"].shape
    g = torch.Generator().manual_seed(1)
    colors = torch.rand(C, N, 3, generator=g).to(DEV)
    opac = st["scene"].opacities[None].contiguous()
    tex = st["scene"].textures
    s = st["scene"]

    def render(marks):
        _ops.set_texel_marks(marks)
        try:
            t = tex.detach().clone().requires_grad_(True)
            out = _wrapper.rasterize_to_pixels_textured_gaussians(
                st["means2d"], st["rt"], colors, opac, t, 3.0, 3.0, st["normals"], torch.zeros_like(st["means2d"]),
                s.width, s.height, 16, st["offsets"], st["flatten_ids"], backgrounds=s.backgrounds)
        finally:
            _ops.set_texel_marks(None)
        (out[0] * torch.rand(out[0].shape, generator=torch.Generator(device=DEV).manual_seed(2), device=DEV)).sum() \
            .backward()
        return [o.detach() for o in out[:5]], t.grad
    marks = torch.zeros(tex.shape[:3], dtype=torch.uint8, device=DEV)
    a, ga = render(None)
    b, gb = render(marks)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    touched = (gb != 0).any(-1)
    assert bool(touched.any()) and not bool((touched & (marks == 0)).any())
    listed = _ops.mark_rows(st["flatten_ids"], N).bool()
    assert not bool(marks[~listed].any())
    assert int(marks.count_nonzero()) >= int(touched.count_nonzero())


@pytest.mark.parametrize("tex_dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("tile_size,width,height", [(16, 80, 64), (8, 75, 61)])
def test_pair_samples_replace_the_texel_reads(tex_dtype, tile_size, width, height):
    """With the forward's pair samples (its texture sample of every composited in-range
    pair) the reference-rule backward reads no texels and takes the forward's
    composited-lane masks for its keep decisions: every gradient equals the backward
    that re-reads the texels (float32; the samples are the same values, so only the
    order of the float atomics differs), through the records path and through float
    atomics, and the forward's outputs do not change."""
    dtype = torch.float32
    sc = scenes.config1(n=2500, width=width, height=height, tex_res=12, seed=9, sh_degree=None)
    st = gpu_stages(sc, dtype, tile_size=tile_size)
    s = st["scene"]
    C, N = st["radii"].shape
    g = torch.Generator().manual_seed(10)
    colors = torch.rand(C, N, 3, generator=g, dtype=torch.float64).to(DEV, dtype)
    opac = sc.opacities.double()[None].repeat(C, 1).to(DEV, dtype)
    tex = sc.textures.to(DEV, dtype).to(tex_dtype)
    flat = st["flatten_ids"]
    fargs = (st["means2d"], st["rt"], colors, opac, tex, 3.0, 3.0, st["normals"], None, None, s.width, s.height,
             tile_size, st["offsets"], flat, 0.0, 1.0)
    live = _ops.live_entries_buffer(flat, tile_size)
    live0 = _ops.live_entries_buffer(flat, tile_size)
    smp = _ops.pair_samples_buffer(flat, tile_size)
    cnt, cnt0 = _ops.tex_counts_buffer(tex), _ops.tex_counts_buffer(tex)
    f1 = _ops.rasterize_to_pixels_fwd_textured_gaussians_full(*fargs, live_entries=live, tex_counts=cnt,
                                                             pair_samples=smp)
    f0 = _ops.rasterize_to_pixels_fwd_textured_gaussians_full(*fargs, live_entries=live0, tex_counts=cnt0)
    for i, (a, b) in enumerate(zip(f0, f1)):
        if i == 7:  # contribution sums: float atomics in any order
            assert rel_err(b, a) < 1e-6
        else:
            assert torch.equal(a, b), i
    assert torch.equal(cnt, cnt0)
    H, W = s.height, s.width
    gen = torch.Generator().manual_seed(11)
    vc = torch.randn(C, H, W, 3, generator=gen).to(DEV)
    va = torch.randn(C, H, W, 1, generator=gen).to(DEV)
    vn = torch.randn(C, H, W, 3, generator=gen).to(DEV)
    vd = torch.randn(C, H, W, 1, generator=gen).to(DEV)
    vm = torch.randn(C, H, W, 1, generator=gen).to(DEV)
    densify = torch.zeros_like(st["means2d"])
    bargs = [st["means2d"], st["rt"], colors, opac, tex, 3.0, 3.0, st["normals"], densify, None, None, W, H,
             tile_size, st["offsets"], flat, 1.0, f1[0], f1[1], f1[5], f1[6], vc, va, vn, vd, vm, False]
    for records in (False, True):
        opts = dict(live_entries=live, render_trans=f1[10])
        if records:
            opts["tex_counts"] = cnt
        ref = _ops.rasterize_to_pixels_bwd_textured_gaussians(*bargs, **opts)
        if records:
            opts["tex_counts"] = cnt.clone()
        got = _ops.rasterize_to_pixels_bwd_textured_gaussians(*bargs, pair_samples=smp, **opts)
        for k, (a, b) in enumerate(zip(ref[1:], got[1:])):
            assert torch.isfinite(b).all(), (records, k)
            assert rel_err(b, a) < 1e-5, (records, k, rel_err(b, a))
        assert float(got[5].abs().sum()) > 0


def test_gather_rows_with_many_records():
    """Large magnified splats put thousands of records on a texture row, many on one texel
    quad: the row gather then adds chunks past the ones its adders prefetch (4 x 2 x 64
    records) and its float64 LDS atomics meet on equal addresses.  The product path (records
    and the gather) matches the oracle over the whole frame in float32, and equals the
    float-atomic path's gradients."""
    from textured_gaussians_amd.cuda import _wrapper
    from gpu_util import all_tiles, raster_tile_parity
    dtype = torch.float32
    sc = scenes.config1(n=400, width=128, height=96, tex_res=50, logscale_mu=-2.2, logscale_sigma=0.6, seed=21)
    st = gpu_stages(sc, dtype)
    s = st["scene"]
    colors = torch.clamp_min(0.28209479177387814 * s.colors[None, :, 0, :] + 0.5, 0.0).contiguous()
    opac = s.opacities[None].contiguous()
    tex = s.textures
    cnt = _ops.tex_counts_buffer(tex)
    _ops.rasterize_to_pixels_fwd_textured_gaussians_full(
        st["means2d"], st["rt"], colors, opac, tex, 3.0, 3.0, st["normals"], None, None, s.width, s.height, 16,
        st["offsets"], st["flatten_ids"], 0.0, 1.0, tex_counts=cnt)
    assert int(cnt.max()) > 4 * 4 * 2 * 64, int(cnt.max())
    assert bool(_wrapper._use_records(tex, dtype))
    stats = raster_tile_parity(st, colors, opac, tex, s.backgrounds, dtype, all_tiles(st))
    print("many records per row", int(cnt.max()), stats)
    # the same gradients through float atomics
    grads = {}
    saved = _wrapper.TEXTURE_GRAD_RECORDS
    try:
        for mode in (True, False):
            _wrapper.TEXTURE_GRAD_RECORDS = mode
            t = tex.clone().requires_grad_(True)
            m2 = st["means2d"].clone().requires_grad_(True)
            out = _wrapper.rasterize_to_pixels_textured_gaussians(
                m2, st["rt"], colors, opac, t, 3.0, 3.0, st["normals"], torch.zeros_like(st["means2d"]), s.width,
                s.height, 16, st["offsets"], st["flatten_ids"], backgrounds=s.backgrounds)
            (out[0].square().sum() + out[1].sum()).backward()
            grads[mode] = (t.grad.clone(), m2.grad.clone())
    finally:
        _wrapper.TEXTURE_GRAD_RECORDS = saved
    for a, b in zip(grads[False], grads[True]):
        assert rel_err(b, a) < 1e-5, rel_err(b, a)
