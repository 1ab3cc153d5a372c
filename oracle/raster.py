"""Textured / plain 2DGS rasterization to pixels: CPU restatement.

Test infrastructure only.  Forward semantics follow
``textured_gaussians/cuda/csrc/rasterize_to_pixels_textured_gaussians_2dgs_fwd.cu:105-445``
with the bilinear lookup of ``filters/bilinear.cuh:15-64`` (textured mode) and
``rasterize_to_pixels_2dgs_fwd.cu`` (plain mode: ``alpha = min(.999, o e^-sigma)``,
no texture).  The per-pixel sequential walk is written as a vectorised scan per
tile: an entry is *valid* if ``ray_cross.z != 0``, ``sigma >= 0`` and
``alpha >= 1/255``; it is *included* if valid and no earlier-or-same valid
entry drove ``T * (1 - alpha)`` to ``<= 1e-4`` (the stopping entry itself is
excluded, ``:349-354``).

The function is differentiable with torch autograd.  ``uv_grad`` selects how
the bilinear sample position depends on the splat coordinate ``s``:

* ``"exact"``  - the true derivative (``du/ds = (res-1) / (2 range)``);
* ``"none"``   - sample weights held constant, the rule of the reference's
  ``bilinear`` backward (``rasterize_to_pixels_textured_gaussians_2dgs_bwd.cu``,
  which has no d(texture)/ds term);
* ``"double"`` - twice the true derivative, the rule of the reference's
  ``bilinear_bwd2`` backward (``..._bwd2.cu:626-628`` uses
  ``du/ds = (res-1) / range``).

Textures are indexed with the Gaussian id ``flatten_id mod N`` and the
per-Gaussian contribution counters likewise (the reference indexes both with
the flatten id, which is out of bounds for cameras >= 1).
"""
from typing import Dict, Optional

import torch
from torch import Tensor


def _bilinear(textures: Tensor, n: Tensor, s: Tensor, rx: float, ry: float, uv_grad: str):
    """Bilinear RGBA sample of ``textures[n]`` at splat coords ``s`` [P,L,2].

    Returns (sample [P,L,4], in_range [P,L]).  ``filters/bilinear.cuh:15-64``:
    out of range if ``s.x < -rx or s.x > rx or s.y < -ry or s.y > ry``;
    ``u = (s.x + rx) / (2 rx) * (tw - 1)`` clamped to [0, tw-1]; texels at
    floor/ceil of u, v."""
    th, tw = textures.shape[1], textures.shape[2]
    in_range = (s[..., 0] >= -rx) & (s[..., 0] <= rx) & (s[..., 1] >= -ry) & (s[..., 1] <= ry)
    if uv_grad == "none":
        s_use = s.detach()
    else:
        s_use = s
    u = (s_use[..., 0] + rx) / (rx * 2) * (tw - 1)
    v = (s_use[..., 1] + ry) / (ry * 2) * (th - 1)
    if uv_grad == "double":
        u = u.detach() + 2.0 * (u - u.detach())
        v = v.detach() + 2.0 * (v - v.detach())
    # out-of-range pairs are masked afterwards; keep their indices in bounds
    u = torch.where(in_range, u, torch.zeros_like(u))
    v = torch.where(in_range, v, torch.zeros_like(v))
    u = torch.clamp(u, 0, tw - 1)
    v = torch.clamp(v, 0, th - 1)
    u0 = torch.floor(u.detach()).long()
    v0 = torch.floor(v.detach()).long()
    u1 = torch.ceil(u.detach()).long()
    v1 = torch.ceil(v.detach()).long()
    wu = u - u0.to(u.dtype)
    wv = v - v0.to(v.dtype)
    nn = n[None, :].expand_as(u0)
    t00 = textures[nn, v0, u0]
    t10 = textures[nn, v0, u1]
    t01 = textures[nn, v1, u0]
    t11 = textures[nn, v1, u1]
    wu_ = wu[..., None]
    wv_ = wv[..., None]
    samp = ((1 - wu_) * (1 - wv_) * t00 + wu_ * (1 - wv_) * t10
            + (1 - wu_) * wv_ * t01 + wu_ * wv_ * t11)
    return samp, in_range


def mip_base(log_res: int, level: int) -> int:
    """First texel of mip level ``level`` (0 = finest, resolution 2^(log_res - level))
    in the flat pyramid, stored coarsest first: (4^(log_res - level) - 1) / 3
    (``filters/mip2.cuh`` t_base_index)."""
    return ((1 << (2 * (log_res - level))) - 1) // 3


def generate_mipmap(textures: Tensor) -> Tensor:
    """[N, R, R, C] (R = 2^L) -> flat pyramid [N, (4^(L+1) - 1) / 3, C]: level l+1 is
    the 2x2 box average of level l, summed as ((t00 + t01) + t10) + t11 then / 4
    (``generate_mipmap_fwd.cu:40-58``).  Differentiable."""
    N, R, R2, C = textures.shape
    assert R == R2 and R & (R - 1) == 0, textures.shape
    L = R.bit_length() - 1
    levels = [textures]
    cur = textures
    for _ in range(L):
        a, b = cur[:, 0::2, 0::2], cur[:, 0::2, 1::2]
        c, d = cur[:, 1::2, 0::2], cur[:, 1::2, 1::2]
        cur = (a + b + c + d) / 4
        levels.append(cur)
    return torch.cat([lv.reshape(N, -1, C) for lv in reversed(levels)], 1)


def _mip2(mip: Tensor, n: Tensor, s: Tensor, dsdx: Tensor, dsdy: Tensor, log_res: int, rx: float, ry: float):
    """Trilinear sample of the pyramid ``mip[n]`` at splat coords ``s`` [P,L,2] with
    the level of detail from the splat-coordinate change over one pixel
    (``filters/mip2.cuh`` precompute): duv/dx uses the x range for both components and
    duv/dy the y range (as there); t = log2(max |duv/dx|, |duv/dy|) clamped to [0, 31];
    u = (s.x + rx) / (2 rx) (res - 1) unclamped; per level l, u_l = (u + 0.5) / 2^l - 0.5
    and texels floor / ceil; invalid when any of them leaves its level.  Sample
    positions and weights carry no gradient (the reference's rule).
    Returns (sample [P,L,4], valid [P,L], edge [P,L]); ``edge`` marks pairs within
    float32 rounding of a place where ``valid`` flips: a level coordinate at 0 or at
    its level's last texel, or t at an integer (the levels, and so their extents,
    change there)."""
    res = 1 << log_res
    s, dsdx, dsdy = s.detach(), dsdx.detach(), dsdy.detach()
    kx, ky = (res - 1) / (2 * rx), (res - 1) / (2 * ry)
    duv = torch.maximum(torch.sqrt((dsdx[..., 0] * kx) ** 2 + (dsdx[..., 1] * kx) ** 2),
                        torch.sqrt((dsdy[..., 0] * ky) ** 2 + (dsdy[..., 1] * ky) ** 2))
    t = torch.log2(duv)
    edge = (t > -1e-4) & (t < 31) & ((t - torch.round(t)).abs() < 1e-4)
    t = torch.where(t < 0, torch.zeros_like(t), torch.where(t > 31, torch.full_like(t, 31.0), t))
    u = (s[..., 0] + rx) / (2 * rx) * (res - 1)
    v = (s[..., 1] + ry) / (2 * ry) * (res - 1)
    tl = torch.floor(t)
    th = torch.ceil(t)
    valid = torch.ones_like(u, dtype=torch.bool)
    corners = []
    for lvl in (tl, th):
        sc = torch.pow(2.0, lvl)
        ul = (u + 0.5) / sc - 0.5
        vl = (v + 0.5) / sc - 0.5
        u0, u1 = torch.floor(ul), torch.ceil(ul)
        v0, v1 = torch.floor(vl), torch.ceil(vl)
        hi = torch.floor((res - 1) / sc)  # (res - 1) >> level
        valid &= (u0 >= 0) & (v0 >= 0) & (u1 <= hi) & (v1 <= hi)
        edge |= ((ul.abs() < 2e-4) | ((ul - hi).abs() < 2e-4) | (vl.abs() < 2e-4) | ((vl - hi).abs() < 2e-4))
        corners.append((lvl, ul - u0, vl - v0, u0, u1, v0, v1))
    wt = t - tl
    nn = n[None, :].expand_as(u)
    samp = torch.zeros(u.shape + (4,), dtype=mip.dtype)
    for j, (lvl, wu, wv, u0, u1, v0, v1) in enumerate(corners):
        lv = torch.where(valid, lvl, torch.zeros_like(lvl)).long().clamp(max=log_res)
        lres = torch.div(torch.full_like(lv, res), torch.pow(2, lv), rounding_mode="floor")
        base = torch.div(torch.pow(4, log_res - lv) - 1, 3, rounding_mode="floor")
        z = torch.zeros_like(u0)
        idx = lambda vv, uu: base + torch.where(valid, vv, z).long() * lres + torch.where(valid, uu, z).long()  # noqa
        wl = (1 - wt) if j == 0 else wt
        for (vv, uu, w) in ((v0, u0, (1 - wu) * (1 - wv)), (v0, u1, wu * (1 - wv)), (v1, u0, (1 - wu) * wv),
                            (v1, u1, wu * wv)):
            samp = samp + (w * wl)[..., None] * mip[nn, idx(vv, uu)]
    return samp, valid, edge


def rasterize_to_pixels(means2d: Tensor, ray_transforms: Tensor, colors: Tensor,
                        opacities: Tensor, normals: Tensor, image_width: int,
                        image_height: int, tile_size: int, isect_offsets: Tensor,
                        flatten_ids: Tensor, textures: Optional[Tensor] = None,
                        texture_range=(3.0, 3.0), backgrounds: Optional[Tensor] = None,
                        masks: Optional[Tensor] = None, g_weight: float = 1.0,
                        gs_contrib_threshold: float = 0.0, uv_grad: str = "exact",
                        distort_grad: str = "exact", pair_means: Optional[list] = None,
                        tiles: Optional[list] = None, mip_log_res: Optional[int] = None) -> Dict[str, Tensor]:
    """Render [C,H,W,*] images.  Shapes as the reference op
    ``rasterize_to_pixels_fwd_textured_gaussians`` (``bindings.h:509-538``):
    means2d [C,N,2], ray_transforms [C,N,3,3], colors [C,N,K], opacities
    [C,N], normals [C,N,3], textures [N,th,tw,4] or None (plain 2DGS).

    If ``pair_means`` is a list, the 2D means used by the 2D-kernel branch
    are per-(pixel, entry) copies appended to it as (cam, flatten ids, tensor)
    with ``retain_grad``; their gradients give the reference's ``absgrad``.

    ``distort_grad="reference"`` reproduces the reference backward's
    distortion rule: it seeds the depth accumulator with
    ``render_colors[..., K-1]`` (``..._bwd.cu:189-191``), which exceeds the
    forward's ``sum_i w_i d_i`` by ``delta = sum_i w_i tex_i[K-1] + T bg[K-1]``;
    that backward is the gradient of ``distort - 2 delta (1 - T)`` with delta
    held constant, which is what is added here (value unchanged).

    The output ``fragile`` [C,H,W] marks pixels where some entry sits within
    a float32 rounding margin of a discrete decision (alpha = 1/255, the
    texture-range edge, the 1e-4 stop, the 0.5 median test), where a float32
    kernel may legitimately decide the other way.  ``cond_max`` [C,N] is,
    per Gaussian, the largest condition number |hu||hv|/|cross.z| of the splat
    coordinate over its kept pairs (large for splats seen edge-on);
    ``texel_edge`` [C,N] marks Gaussians with a kept pair whose texture
    coordinate is within 1e-4 texel of a texel edge.  ``fragile_rows`` [C,N]
    marks Gaussians with a fragile pair or a valid pair in a fragile pixel: the
    rows whose gradient a flipped float32 decision can move; ``fragile_own`` [C,N]
    the subset with a fragile pair of their own (a flip there adds or removes that
    pair's whole term from the Gaussian's gradient).

    ``mip_log_res``: ``textures`` is the flat pyramid [N, T, 4] of ``generate_mipmap``
    over 2^mip_log_res textures and sampling is the reference's ``mipmapped2`` filter
    (``rasterize_to_pixels_mip2_textured_gaussians_2dgs_fwd.cu:294-395``): the pair
    also needs non-zero z for the crosses at px +- 1/2 and py +- 1/2, from which
    ds/dx and ds/dy come; ``uv_grad`` must be "none".

    ``tiles`` (optional list of (camera, tile_y, tile_x) or (camera, tile_y,
    tile_x, start, end)) restricts the work to those tiles, optionally with
    explicit ranges into ``flatten_ids``; other pixels stay zero.
    """
    C, N = means2d.shape[:2]
    K = colors.shape[-1]
    dt = means2d.dtype
    th_t, tw_t = isect_offsets.shape[1:]
    I = flatten_ids.numel()
    H, W, ts = image_height, image_width, tile_size
    rx, ry = float(texture_range[0]), float(texture_range[1])
    textured = textures is not None
    out_c = torch.zeros(C, H, W, K, dtype=dt)
    out_a = torch.zeros(C, H, W, 1, dtype=dt)
    out_n = torch.zeros(C, H, W, 3, dtype=dt)
    out_d = torch.zeros(C, H, W, 1, dtype=dt)
    out_m = torch.zeros(C, H, W, 1, dtype=dt)
    last_ids = torch.zeros(C, H, W, dtype=torch.int32)
    median_ids = torch.zeros(C, H, W, dtype=torch.int32)
    csum = torch.zeros(N, dtype=dt)
    ccnt = torch.zeros(N, dtype=dt)
    fragile = torch.zeros(C, H, W, dtype=torch.bool)
    cond_max = torch.zeros(C * N, dtype=torch.float64)
    texel_edge = torch.zeros(C * N, dtype=torch.bool)
    fragile_rows = torch.zeros(C * N, dtype=torch.bool)
    fragile_own = torch.zeros(C * N, dtype=torch.bool)
    offs = isect_offsets.reshape(-1).tolist()
    fl = flatten_ids.long()
    mflat = means2d.reshape(C * N, 2)
    Tflat = ray_transforms.reshape(C * N, 3, 3)
    cflat = colors.reshape(C * N, K)
    oflat = opacities.reshape(C * N)
    nflat = normals.reshape(C * N, 3)
    if tiles is None:
        tiles = [(c, ty, tx) for c in range(C) for ty in range(th_t) for tx in range(tw_t)]
    for tile in tiles:
        c, ty, tx = tile[:3]
        bg = backgrounds[c] if backgrounds is not None else torch.zeros(K, dtype=dt)
        tid = ty * tw_t + tx
        y0, x0 = ty * ts, tx * ts
        y1, x1 = min(y0 + ts, H), min(x0 + ts, W)
        if y0 >= H or x0 >= W:
            continue
        ys = torch.arange(y0, y1)
        xs = torch.arange(x0, x1)
        yy, xx = torch.meshgrid(ys, xs, indexing="ij")
        yy, xx = yy.reshape(-1), xx.reshape(-1)
        if masks is not None and not bool(masks[c, ty, tx]):
            out_c[c, yy, xx] = bg.expand(yy.numel(), K)
            continue
        k = c * th_t * tw_t + tid
        start = offs[k]
        end = I if k == C * th_t * tw_t - 1 else offs[k + 1]
        if len(tile) == 5:  # explicit [start, end) into flatten_ids
            start, end = int(tile[3]), int(tile[4])
        px = xx.to(dt) + 0.5
        py = yy.to(dt) + 0.5
        P = px.numel()
        if end <= start:
            out_c[c, yy, xx] = bg.expand(P, K)
            continue
        g = fl[start:end]
        L = g.numel()
        n = g % N
        Tm = Tflat[g]
        T0, T1, T2 = Tm[:, 0], Tm[:, 1], Tm[:, 2]
        hu = px[:, None, None] * T2[None] - T0[None]
        hv = py[:, None, None] * T2[None] - T1[None]
        cr = torch.cross(hu, hv, dim=-1)
        cz = cr[..., 2]
        okz = cz != 0
        czs = torch.where(okz, cz, torch.ones_like(cz))
        s = cr[..., :2] / czs[..., None]
        if mip_log_res is not None:
            # splat coordinate at the pixel's edges: crosses with h_u -/+ w/2, h_v -/+ w/2
            half = 0.5 * T2[None]
            sh = []
            for a_, b_ in ((hu - half, hv), (hu + half, hv), (hu, hv - half), (hu, hv + half)):
                q = torch.cross(a_, b_, dim=-1).detach()
                okz = okz & (q[..., 2] != 0)
                sh.append(q[..., :2] / torch.where(q[..., 2] != 0, q[..., 2], torch.ones_like(q[..., 2]))[..., None])
            dsdx, dsdy = sh[1] - sh[0], sh[3] - sh[2]
        w3 = s[..., 0] * s[..., 0] + s[..., 1] * s[..., 1]
        mp = mflat[g][None].expand(P, L, 2)
        if pair_means is not None:
            mp = mp.clone()
            mp.retain_grad()
            pair_means.append((c, g, mp))
        dx = mp[..., 0] - px[:, None]
        dy = mp[..., 1] - py[:, None]
        w2 = 2.0 * (dx * dx + dy * dy)
        gw = torch.where(w3 <= w2, w3, w2)
        sigma = 0.5 * gw
        ker = torch.exp(-sigma)
        op = oflat[g][None]
        if textured:
            vis = (0.998 - g_weight) + g_weight * ker
            if mip_log_res is not None:
                assert uv_grad == "none", "the mip2 filter has the reference's rule only"
                samp, inr, mip_edge = _mip2(textures, n, s, dsdx, dsdy, mip_log_res, rx, ry)
            else:
                samp, inr = _bilinear(textures, n, s, rx, ry, uv_grad)
            a_s = torch.where(inr, samp[..., 3], torch.ones_like(samp[..., 3]))
            tex = torch.where(inr[..., None], samp[..., :K], torch.zeros_like(samp[..., :K]))
            araw = op * vis * a_s
        else:
            araw = op * ker
            tex = None
        alpha = torch.clamp(araw, max=0.999)
        valid = okz & ~(sigma < 0) & ~(alpha < 1.0 / 255.0)
        a_eff = torch.where(valid, alpha, torch.zeros_like(alpha))
        one_m = 1 - a_eff
        Tincl = torch.cumprod(one_m, dim=1)
        stop = valid & (Tincl <= 1e-4)
        stopped = torch.cumsum(stop.to(torch.int32), dim=1) > 0
        inc = valid & ~stopped
        Texcl = torch.cat([torch.ones(P, 1, dtype=dt), Tincl[:, :-1]], 1)
        w = torch.where(inc, a_eff * Texcl, torch.zeros_like(a_eff))
        Tfin = torch.prod(torch.where(inc, one_m, torch.ones_like(one_m)), dim=1)
        col = cflat[g][None].expand(P, L, K)
        full = col + tex if textured else col
        out_c[c, yy, xx] = (w[..., None] * full).sum(1) + Tfin[:, None] * bg[None]
        out_a[c, yy, xx] = (1 - Tfin)[:, None]
        out_n[c, yy, xx] = (w[..., None] * nflat[g][None]).sum(1)
        dep = cflat[g][:, K - 1][None]
        wd = w * dep
        A_ex = torch.cumsum(wd, 1) - wd
        dist = (2 * (wd * (1 - Texcl) - w * A_ex)).sum(1)
        if distort_grad == "reference":
            delta = ((w * tex[..., K - 1]).sum(1) if textured else 0.0) + Tfin * bg[K - 1]
            extra = -2.0 * delta.detach() * (1 - Tfin)
            dist = dist + extra - extra.detach()
        out_d[c, yy, xx] = dist[:, None]
        with torch.no_grad():
            # margins: a few times the float32 error of each quantity (alpha to ~1e-6
            # relative, T to ~1e-6 relative per entry, s to ~1e-6 for well-conditioned splats)
            fr = okz & ((alpha - 1.0 / 255.0).abs() < 2e-7)
            fr |= (Tincl / 1e-4 - 1).abs() < 2e-5
            fr |= (Texcl - 0.5).abs() < 2e-6
            if textured and mip_log_res is None:
                fr |= ((s[..., 0].abs() - rx).abs() < 2e-5) | ((s[..., 1].abs() - ry).abs() < 2e-5)
            elif textured:
                # a validity flip moves the pair's alpha between op*vis*a_s and op*vis: it
                # matters where op*vis reaches 1/255 and no earlier entry stopped the pixel
                stop_i = stop.to(torch.int32)
                before = (torch.cumsum(stop_i, 1) - stop_i) > 0
                fr |= okz & mip_edge & (op * vis >= (1.0 - 1e-6) / 255.0) & ~before
            fr_pix = fr.any(1)
            fragile[c, yy, xx] = fr_pix
            # Gaussians whose gradient a flipped decision can move: the fragile pair's own
            # Gaussian, and every Gaussian with a valid pair in a fragile pixel (a flip there
            # changes that pixel's transmittance and accumulated colour behind them)
            fr_g = (fr | (valid & fr_pix[:, None])).any(0)
            fragile_rows.index_put_((g,), fr_g, accumulate=True)
            fragile_own.index_put_((g,), fr.any(0), accumulate=True)
            # conditioning of s = cross.xy / cross.z for kept pairs: |hu| |hv| / |cross.z|
            cnd = (hu.norm(dim=-1) * hv.norm(dim=-1) / cz.abs().clamp_min(1e-300)).to(torch.float64)
            cnd = torch.where(inc, cnd, torch.zeros_like(cnd)).amax(0)
            cond_max.scatter_reduce_(0, g, cnd, reduce="amax")
            if textured and mip_log_res is None:
                # kept in-range pairs whose texture coordinate lies within rounding
                # distance of a texel edge (the bilinear derivative jumps there)
                uu = (s[..., 0] + rx) / (2 * rx) * (textures.shape[2] - 1)
                vv = (s[..., 1] + ry) / (2 * ry) * (textures.shape[1] - 1)
                near = ((uu - torch.round(uu)).abs() < 1e-4) | ((vv - torch.round(vv)).abs() < 1e-4)
                inr = (s[..., 0].abs() <= rx) & (s[..., 1].abs() <= ry)
                edge = (inc & inr & near).any(0)
                texel_edge.index_put_((g,), edge, accumulate=True)
            pos = torch.arange(L)[None].expand(P, L)
            has = inc.any(1)
            lastp = torch.where(inc, pos, torch.full_like(pos, -1)).max(1).values
            last_ids[c, yy, xx] = torch.where(has, start + lastp, 0).to(torch.int32)
            med = inc & (Texcl > 0.5)
            hasm = med.any(1)
            medp = torch.where(med, pos, torch.full_like(pos, -1)).max(1).values
            median_ids[c, yy, xx] = torch.where(hasm, start + medp, 0).to(torch.int32)
            cm = inc & (alpha > gs_contrib_threshold)
            csum.index_add_(0, n, torch.where(cm, alpha, torch.zeros_like(alpha)).sum(0).detach())
            ccnt.index_add_(0, n, cm.to(dt).sum(0))
        medp_c = medp.clamp(min=0)
        mdep = cflat[g][medp_c, K - 1]
        out_m[c, yy, xx] = torch.where(hasm, mdep, torch.zeros_like(mdep))[:, None]
    return {"render_colors": out_c, "render_alphas": out_a, "render_normals": out_n,
            "render_distort": out_d, "render_median": out_m, "last_ids": last_ids,
            "median_ids": median_ids, "contrib_sum": csum, "contrib_count": ccnt, "fragile": fragile,
            "cond_max": cond_max.reshape(C, N), "texel_edge": texel_edge.reshape(C, N),
            "fragile_rows": fragile_rows.reshape(C, N), "fragile_own": fragile_own.reshape(C, N)}


def rasterize_pixel_loop(means2d, ray_transforms, colors, opacities, normals, textures,
                         texture_range, backgrounds, W, H, tile_size, isect_offsets,
                         flatten_ids, g_weight=1.0, c=0, pix=None):
    """Literal per-pixel sequential walk (python floats, float64), for a few
    pixels of tiny scenes; pins the vectorised scan above.  Returns a list of
    (color list, alpha, last_id) per requested pixel (i, j)."""
    import math
    C, N = means2d.shape[:2]
    K = colors.shape[-1]
    th_t, tw_t = isect_offsets.shape[1:]
    I = flatten_ids.numel()
    rx, ry = texture_range
    res = []
    for (i, j) in pix:
        ty, tx = i // tile_size, j // tile_size
        k = c * th_t * tw_t + ty * tw_t + tx
        start = int(isect_offsets.reshape(-1)[k])
        end = I if k == C * th_t * tw_t - 1 else int(isect_offsets.reshape(-1)[k + 1])
        px, py = j + 0.5, i + 0.5
        T = 1.0
        acc = [0.0] * K
        cur = 0
        for idx in range(start, end):
            g = int(flatten_ids[idx])
            n = g % N
            M = ray_transforms.reshape(-1, 3, 3)[g].tolist()
            hu = [px * M[2][q] - M[0][q] for q in range(3)]
            hv = [py * M[2][q] - M[1][q] for q in range(3)]
            cx = hu[1] * hv[2] - hu[2] * hv[1]
            cy = hu[2] * hv[0] - hu[0] * hv[2]
            czz = hu[0] * hv[1] - hu[1] * hv[0]
            if czz == 0:
                continue
            sx, sy = cx / czz, cy / czz
            mx, my = means2d.reshape(-1, 2)[g].tolist()
            sig = 0.5 * min(sx * sx + sy * sy, 2.0 * ((mx - px) ** 2 + (my - py) ** 2))
            if textures is not None:
                th, tw = textures.shape[1], textures.shape[2]
                if -rx <= sx <= rx and -ry <= sy <= ry:
                    u = min(max((sx + rx) / (2 * rx) * (tw - 1), 0.0), tw - 1.0)
                    v = min(max((sy + ry) / (2 * ry) * (th - 1), 0.0), th - 1.0)
                    u0, v0, u1, v1 = math.floor(u), math.floor(v), math.ceil(u), math.ceil(v)
                    a, b = u - u0, v - v0
                    tx_ = textures[n]
                    smp = [((1 - a) * (1 - b) * float(tx_[v0, u0, q]) + a * (1 - b) * float(tx_[v0, u1, q])
                            + (1 - a) * b * float(tx_[v1, u0, q]) + a * b * float(tx_[v1, u1, q]))
                           for q in range(4)]
                    a_s, tex = smp[3], smp[:K]
                else:
                    a_s, tex = 1.0, [0.0] * K
                alpha = min(0.999, float(opacities.reshape(-1)[g]) *
                            (0.998 - g_weight + g_weight * math.exp(-sig)) * a_s)
            else:
                tex = [0.0] * K
                alpha = min(0.999, float(opacities.reshape(-1)[g]) * math.exp(-sig))
            if sig < 0 or alpha < 1.0 / 255.0:
                continue
            nT = T * (1 - alpha)
            if nT <= 1e-4:
                break
            for q in range(K):
                acc[q] += (float(colors.reshape(-1, K)[g, q]) + tex[q]) * alpha * T
            cur = idx
            T = nT
        bgc = backgrounds[c].tolist() if backgrounds is not None else [0.0] * K
        res.append(([acc[q] + T * bgc[q] for q in range(K)], 1 - T, cur))
    return res
