"""Shared helpers for the GPU parity tests: run the GPU stages and hand their
outputs to the CPU oracle as float64 CPU tensors."""
import json
import math
import os

import numpy as np
import torch

from textured_gaussians_amd.cuda import _ops

DEV = torch.device("cuda:0")


def cpu64(t):
    return None if t is None else t.detach().cpu().double()


def gpu_stages(sc, dtype=torch.float32, tile_size=16):
    """Projection + tile lists on the GPU for scene ``sc`` (CPU tensors)."""
    s = sc.to(DEV, dtype)
    radii, means2d, depths, rt, normals = _ops.fully_fused_projection_fwd_2dgs(
        s.means, s.quats, s.scales, s.viewmats, s.Ks, s.width, s.height, 0.3, 0.01, 1e10, 0.0)
    C = s.viewmats.shape[0]
    tw = math.ceil(s.width / tile_size)
    th = math.ceil(s.height / tile_size)
    tpg, ids, flat = _ops.isect_tiles(means2d, radii, depths, None, None, C, tile_size, tw, th, True, True)
    offs = _ops.isect_offset_encode(ids, C, tw, th)
    return dict(scene=s, radii=radii, means2d=means2d, depths=depths, rt=rt, normals=normals, tpg=tpg,
                isect_ids=ids, flatten_ids=flat, offsets=offs, tw=tw, th=th, C=C, tile_size=tile_size)


def gpu_stages_packed(sc, dtype=torch.float32, tile_size=16):
    """As gpu_stages, in the packed layout: per-entry tensors [nnz, ...] for the visible
    (camera, Gaussian) pairs, with their camera_ids and gaussian_ids."""
    s = sc.to(DEV, dtype)
    _, cid, gid, radii, means2d, depths, rt, normals = _ops.fully_fused_projection_packed_fwd_2dgs(
        s.means, s.quats, s.scales, s.viewmats, s.Ks, s.width, s.height, 0.01, 1e10, 0.0)
    C = s.viewmats.shape[0]
    tw = math.ceil(s.width / tile_size)
    th = math.ceil(s.height / tile_size)
    tpg, ids, flat = _ops.isect_tiles(means2d, radii, depths, cid, gid, C, tile_size, tw, th, True, True)
    offs = _ops.isect_offset_encode(ids, C, tw, th)
    return dict(scene=s, radii=radii, means2d=means2d, depths=depths, rt=rt, normals=normals, tpg=tpg,
                isect_ids=ids, flatten_ids=flat, offsets=offs, tw=tw, th=th, C=C, tile_size=tile_size,
                camera_ids=cid, gaussian_ids=gid)


def rel_err(a, b):
    """max |a-b| / max |b| (0 if both vanish)."""
    a = a.detach().cpu().double()
    b = b.detach().cpu().double()
    den = b.abs().max().item()
    num = (a - b).abs().max().item()
    return num / den if den > 0 else num


# ------------------------------------------------------------------ tile parity
# Float32 kernels vs the float64 oracle, element by element:
#   forward, pixels the oracle does not mark fragile: |gpu - ref| <= 1e-4 (max-abs);
#   forward, fragile pixels (an entry within float32 rounding of 1/255, the texture-range
#   edge, the 1e-4 stop or the 0.5 median test): |gpu - ref| <= 2e-2, the most one flipped
#   pair can move a pixel (alpha ~ 1/255 times a colour below ~5);
#   gradients, per element e of Gaussian row j, rows without a fragile pair:
#     |gpu_e - ref_e| <= 1e-3 |ref_e| + 1e-4 max_j |ref| + 3 |ref32_e - ref_e|,
#   i.e. 1e-3 relative with an absolute floor of 1e-4 of the row's largest element, plus
#   three times the float32 oracle's own error on that element (what float32 arithmetic of
#   the same formula cannot avoid);
#   gradients, rows with a valid pair in a fragile pixel (a flip there moves that pixel's
#   transmittance, the colour accumulated behind them, or which entry is the median):
#     |gpu_e - ref_e| <= 5e-2 max_j |ref| + 1e-5 max |ref| + 3 |ref32_e - ref_e|
#   (the second term: the tensor's largest element, for rows whose own values are tiny);
#   rows with a fragile pair of their own (a flip adds or removes that pair's whole term)
#   and, for the exact and doubled uv rules, rows with an ill-conditioned splat coordinate
#   (condition > 1e3) or a kept pair on a texel edge (the bilinear derivative jumps
#   there): compared only inside the tensor's L2 norm, and they must stay under 10% of
#   the rows;
#   every tensor in the L2 norm: within max(1e-3, 3x the float32 oracle's own L2 error)
#   (ray-transform gradients carry pixel-coordinate-sized terms that cancel, so at
#   1297 px float32 arithmetic of the same formula is itself ~1e-3 off).
# Float64 kernels: every pixel within 1e-9, every gradient tensor within 1e-8 of its
# largest element (max-abs relative).
FWD_TOL32, FWD_FRAGILE_TOL32 = 1e-4, 2e-2
GRAD_RTOL32, GRAD_ATOL_ROW32, GRAD_LOOSE32, GRAD_FLOOR32 = 1e-3, 1e-4, 5e-2, 1e-5


def sample_tiles(st, k, min_len=50, extra=()):
    """``k`` tiles with more than ``min_len`` list entries, spread evenly over the tile
    grid of all cameras, plus the tiles ``extra`` given as (camera, ty, tx).  Returns
    (camera, ty, tx, start, end) per tile."""
    offs = st["offsets"].reshape(-1).cpu().numpy().astype(np.int64)
    I = st["flatten_ids"].numel()
    ends = np.append(offs[1:], I)
    per = ends - offs
    tpc = st["tw"] * st["th"]
    cand = [t for t in range(offs.size) if per[t] > min_len]
    pick = {cand[i] for i in np.linspace(0, len(cand) - 1, k).astype(int)} if cand else set()
    for c, ty, tx in extra:
        pick.add(c * tpc + ty * st["tw"] + tx)
    out = []
    for t in sorted(pick):
        c, r = divmod(t, tpc)
        out.append((c, r // st["tw"], r % st["tw"], int(offs[t]), int(ends[t])))
    return out


def all_tiles(st):
    return sample_tiles(st, 0, extra=[(c, ty, tx) for c in range(st["C"]) for ty in range(st["th"])
                                      for tx in range(st["tw"])])


# a JSON-lines file that parity failures are appended to, for reading offline (off unless set)
REPORT = os.environ.get("TGS_PARITY_REPORT")


def _report(rec):
    """Append a failure record (one JSON line) to REPORT, when set; best effort."""
    if not REPORT:
        return
    try:
        os.makedirs(os.path.dirname(os.path.abspath(REPORT)), exist_ok=True)
        with open(REPORT, "a") as f:
            f.write(json.dumps(rec, default=float) + "\n")
    except OSError:
        pass


def check_grad(name, got, r64, r32, loose_rows, dtype, row_info=None, failures=None, norm_rows=None):
    """Per-element gradient check (rules above); rows on dim 0.  Every row is held to
    the strict bound; only ``loose_rows`` may fall back to the loose one, and
    ``norm_rows`` are checked in the tensor's L2 norm only.  Violations are appended to
    ``failures`` (or raised at once when it is None).  Returns stats."""
    G = r64.shape[0]
    g = cpu64(got).reshape(G, -1)
    r = r64.detach().reshape(G, -1)
    if dtype == torch.float64:
        e = rel_err(g, r)
        if not e < 1e-8:
            (failures.append if failures is not None else _raise)({"tensor": name, "max_rel": e})
        return {"max_rel": e}
    q = r32.detach().double().reshape(G, -1)
    err = (g - r).abs()
    f32e = (q - r).abs()
    rowmax = r.abs().amax(1, keepdim=True)
    strict = GRAD_RTOL32 * r.abs() + GRAD_ATOL_ROW32 * rowmax + 3 * f32e
    loose = GRAD_LOOSE32 * rowmax + GRAD_FLOOR32 * r.abs().max() + 3 * f32e
    if norm_rows is None:
        norm_rows = torch.zeros_like(loose_rows)
    over_s = (err > strict).any(1) & ~norm_rows
    over_l = (err > loose).any(1) & ~norm_rows
    bad = (over_s & ~loose_rows) | over_l
    l2 = float((g - r).norm() / r.norm().clamp_min(1e-300))
    l2_32 = float((q - r).norm() / r.norm().clamp_min(1e-300))
    if not (l2 < max(GRAD_RTOL32, 3 * l2_32) and float(norm_rows.float().mean()) < 0.10):
        (failures.append if failures is not None else _raise)({"tensor": name, "l2_rel": l2, "l2_rel_f32": l2_32,
                                                               "norm_rows": int(norm_rows.sum())})
    stats = {"rows": G, "loose_allowed": int(loose_rows.sum()), "loose_used": int(over_s.sum()),
             "norm_only": int(norm_rows.sum()), "l2_rel_f32": l2_32,
             "worst_strict": float((err / strict.clamp_min(1e-300)).max()) if err.numel() else 0.0,
             "l2_rel": float((g - r).norm() / r.norm().clamp_min(1e-300))}
    for j in bad.nonzero().flatten()[:6].tolist():
        rec = {"tensor": name, "row": j, "stats": stats, "got": g[j].tolist()[:12], "ref64": r[j].tolist()[:12],
               "ref32": q[j].tolist()[:12], "loose_allowed": bool(loose_rows[j]),
               "ratio_strict": float((err[j] / strict[j].clamp_min(1e-300)).max()),
               "ratio_loose": float((err[j] / loose[j].clamp_min(1e-300)).max())}
        if row_info is not None:
            rec.update({k: (v[j].item() if torch.is_tensor(v) else v[j]) for k, v in row_info.items()})
        (failures.append if failures is not None else _raise)(rec)
    return stats


def _raise(rec):
    raise AssertionError(rec)


def raster_tile_parity(st, colors, opac, textures, bg, dtype, tiles, uv="none", seed=5, backward=True,
                       gaussian_ids=None, n_gaussians=None, filtering=None, zero_fragile_cotangents=False):
    """Rasterizer forward over the whole frame on the GPU, then its backward with upstream
    gradients (all five rendered outputs) nonzero only inside ``tiles``, so every
    per-Gaussian and per-texel gradient comes from those tiles and must equal the
    oracle's autograd on those tiles alone (a compact problem holding only the
    (camera, Gaussian) entries their lists reference).  Inputs are GPU tensors of the
    kernel dtype (textures may be another float dtype, bfloat16 included; the oracle
    reads them widened to float64/float32).  ``backward=False``: forward parity only.
    ``gaussian_ids`` [nnz]: the packed layout (per-entry inputs [nnz, ...], entry i
    samples texture row gaussian_ids[i]; ``n_gaussians`` rows when untextured).
    ``filtering="mipmapped2"``: the textures are the [N, R, R, 4] base level; both sides
    build the pyramid and differentiate through it.  The background gradient is
    checked too when ``bg`` is given.  ``zero_fragile_cotangents``: the upstream gradient
    is zero at the pixels the float64 oracle marks fragile (for filters whose flips move a
    pair's alpha by a large step, which can move a pixel's median entry to another
    Gaussian: a whole term that no bound on rounding covers).
    Returns a dict of statistics; asserts the tolerances above."""
    from oracle.raster import generate_mipmap as oracle_mipmap
    from oracle.raster import rasterize_to_pixels
    s = st["scene"]
    packed = gaussian_ids is not None
    C = st["offsets"].shape[0]
    if packed:
        E = st["means2d"].shape[0]
        N = textures.shape[0] if textures is not None else int(n_gaussians)
        lead = 1
    else:
        N = st["means2d"].shape[1]
        E = C * N
        lead = 2
    K = colors.shape[-1]
    W, H, ts = s.width, s.height, st["tile_size"]
    textured = textures is not None
    exact = uv == "exact"
    mip = filtering == "mipmapped2"
    L_mip = int(textures.shape[1]).bit_length() - 1 if mip else None
    assert not mip or uv == "none"
    # the op-level forward for the list positions (last_ids, median_ids)
    tex_fwd = _ops.generate_mipmap_fwd(textures, L_mip) if mip else textures
    fo = _ops._raster_fwd(textured, st["means2d"], st["rt"], colors, opac, tex_fwd, 3.0, 3.0, st["normals"], bg, None,
                          W, H, ts, st["offsets"], st["flatten_ids"], 0.0, 1.0, gaussian_ids=gaussian_ids,
                          n_gaussians=N if packed else None, mip_log_res=L_mip)
    ids = (fo[5], fo[6])
    fo = tex_fwd = None

    # the compact problem: the (camera, Gaussian) entries the sampled tiles reference
    flat = st["flatten_ids"].cpu().numpy()
    sel_ids = np.unique(np.concatenate([flat[a:b] for *_, a, b in tiles]))
    u = torch.from_numpy(sel_ids.astype(np.int64))
    rank = np.full(E, -1, np.int64)
    rank[sel_ids] = np.arange(sel_ids.size)
    new_flat = torch.from_numpy(np.concatenate([rank[flat[a:b]] for *_, a, b in tiles]).astype(np.int32))
    ltiles, o = [], 0
    for c, ty, tx, a, b in tiles:
        ltiles.append((0, ty, tx, o, o + (b - a)))
        o += b - a
    Gn = u.numel()
    tex_rows = None
    # texture row (Gaussian) of each compact entry
    row_of = gaussian_ids.cpu()[u] if packed else u % N
    if textured:
        tex_rows = textures[row_of.to(textures.device)].cpu()

    def oracle(dt, grad=backward):
        def take(t):
            tail = t.shape[lead:]
            return t.detach().cpu().to(dt).reshape(E, *tail)[u].reshape(1, Gn, *tail).clone().requires_grad_(True)
        L = dict(m2=take(st["means2d"]), rt=take(st["rt"]), col=take(colors),
                 op=opac.detach().cpu().to(dt).reshape(-1)[u].reshape(1, Gn).clone().requires_grad_(True),
                 nrm=take(st["normals"]))
        if textured:
            L["tex"] = tex_rows.to(dt).requires_grad_(True)
        if bg is not None:
            L["bg"] = bg.detach().cpu().to(dt).clone().requires_grad_(True)
        tex_in = None
        if textured:
            tex_in = oracle_mipmap(L["tex"]) if mip else L["tex"]
        offs_dummy = torch.zeros(1, st["th"], st["tw"], dtype=torch.int32)
        loss, outs = 0.0, []
        for (c, ty, tx, a, b), lt in zip(tiles, ltiles):  # one tile at a time: each has its own camera
            r = rasterize_to_pixels(L["m2"], L["rt"], L["col"], L["op"], L["nrm"], W, H, ts, offs_dummy, new_flat,
                                    textures=tex_in, backgrounds=None if bg is None else L["bg"][c:c + 1],
                                    uv_grad=uv, distort_grad="exact" if exact else "reference", tiles=[lt],
                                    mip_log_res=L_mip)
            ys, xs = slice(ty * ts, (ty + 1) * ts), slice(tx * ts, (tx + 1) * ts)
            outs.append((c, ys, xs, r))
            if grad:
                for k, v in cot.items():
                    loss = loss + (r[k][0, ys, xs] * v[c, ys, xs].to(dt)).sum()
        if grad:
            loss.backward()
        return L, outs

    sel = torch.zeros(C, H, W, 1, dtype=torch.float64)
    for c, ty, tx, _, _ in tiles:
        sel[c, ty * ts:(ty + 1) * ts, tx * ts:(tx + 1) * ts] = 1.0
    n_zeroed = 0
    if zero_fragile_cotangents:
        # no upstream gradient at pixels with a pair within float32 rounding of a flip (their
        # forward is still checked at the fragile bound)
        with torch.no_grad():
            _, pre = oracle(torch.float64, grad=False)
        for (c, ys, xs, r) in pre:
            fr = r["fragile"][0, ys, xs]
            n_zeroed += int(fr.sum())
            sel[c, ys, xs, 0] *= (~fr).double()
        pre = None
    g = torch.Generator().manual_seed(seed)
    cot = {"render_colors": torch.randn(C, H, W, K, generator=g, dtype=torch.float64) * sel,
           "render_alphas": torch.randn(C, H, W, 1, generator=g, dtype=torch.float64) * sel,
           "render_normals": torch.randn(C, H, W, 3, generator=g, dtype=torch.float64) * sel,
           "render_distort": torch.randn(C, H, W, 1, generator=g, dtype=torch.float64) * sel,
           "render_median": torch.randn(C, H, W, 1, generator=g, dtype=torch.float64) * sel}
    # the product path: the rasterizer's autograd function (records, gather, fused row fill,
    # the forward's transmittance handed to the backward), as the rendering API calls it
    from textured_gaussians_amd.cuda import _wrapper as Wr
    leaves = {"m2": st["means2d"].detach().clone().requires_grad_(backward),
              "rt": st["rt"].detach().clone().requires_grad_(backward),
              "col": colors.detach().clone().requires_grad_(backward),
              "op": opac.detach().clone().requires_grad_(backward),
              "nrm": st["normals"].detach().clone().requires_grad_(backward)}
    dens = torch.zeros_like(st["means2d"])
    bg_leaf = None if bg is None else bg.detach().clone().requires_grad_(backward)
    if textured:
        tex_leaf = textures.detach().requires_grad_(backward and textures.is_floating_point()
                                                    and textures.dtype != torch.bfloat16)
        filt = "mipmapped2" if mip else {"none": "bilinear", "double": "bilinear_bwd2", "exact": "bilinear_exact"}[uv]
        outs_gpu = Wr.rasterize_to_pixels_textured_gaussians(
            leaves["m2"], leaves["rt"], leaves["col"], leaves["op"], tex_leaf, 3.0, 3.0, leaves["nrm"], dens, W, H,
            ts, st["offsets"], st["flatten_ids"], backgrounds=bg_leaf, filtering=filt, packed=packed,
            gaussian_ids=gaussian_ids)
    else:
        outs_gpu = Wr.rasterize_to_pixels_2dgs(
            leaves["m2"], leaves["rt"], leaves["col"], leaves["op"], leaves["nrm"], dens, W, H, ts, st["offsets"],
            st["flatten_ids"], backgrounds=bg_leaf, grad_mode="exact" if exact else "reference", packed=packed,
            gaussian_ids=gaussian_ids, n_gaussians=N if packed else None)
    fo = [o.detach() for o in outs_gpu[:5]] + list(ids)
    assert float(fo[1].min()) >= 0.0 and float(fo[1].max()) <= 1.0
    v_m2 = v_rt = v_col = v_op = v_tex = v_nrm = v_bg = None
    if backward:
        names5 = ["render_colors", "render_alphas", "render_normals", "render_distort", "render_median"]
        torch.autograd.backward(list(outs_gpu[:5]), [cot[k].to(DEV, dtype) for k in names5])
        v_m2, v_rt, v_col, v_op, v_nrm = (leaves[k].grad for k in ("m2", "rt", "col", "op", "nrm"))
        v_bg = None if bg_leaf is None else bg_leaf.grad
        if textured:
            v_tex = tex_leaf.grad
            tex_leaf.grad = None
    outs_gpu = leaves = None
    if textured:
        textures = textures.detach()

    with torch.set_grad_enabled(backward):
        L64, outs64 = oracle(torch.float64)
        L32, outs32 = oracle(torch.float32) if (dtype == torch.float32 and backward) else (None, None)

    # forward
    stats = {"tiles": len(tiles), "gaussians": Gn, "cotangent_zeroed_pixels": n_zeroed}
    failures = []
    names = ["render_colors", "render_alphas", "render_normals", "render_distort", "render_median"]
    n_frag, n_pix, worst, frag_over = 0, 0, {}, 0
    for (c, ys, xs, r) in outs64:
        frag = r["fragile"][0, ys, xs]
        n_frag += int(frag.sum())
        n_pix += frag.numel()
        over_any = torch.zeros_like(frag)
        for nm, o in zip(names, fo[:5]):
            d = (cpu64(o[c, ys, xs]) - r[nm][0, ys, xs].detach()).abs()
            d = d.amax(-1) if d.dim() == 3 else d
            worst[nm] = max(worst.get(nm, 0.0), float(d.max()))
            tol = 1e-9 if dtype == torch.float64 else FWD_TOL32
            over = d > tol
            over_any |= over
            # every pixel is held to the tolerance; only fragile pixels may use the looser bound
            bad = (over & ~frag) | (d > FWD_FRAGILE_TOL32) if dtype == torch.float32 else over
            for (i, j) in bad.nonzero()[:4].tolist():
                failures.append({"output": nm, "camera": c, "pixel": [ys.start + i, xs.start + j],
                                 "d": float(d[i, j]), "fragile": bool(frag[i, j])})
        frag_over += int(over_any.sum())
        # list positions (bit-exact away from fragile pixels): the compact problem's positions
        # mapped back to the full lists; pixels with nothing composited hold 0 in both
        keep = ~frag if dtype == torch.float32 else torch.ones_like(frag)
        hit = r["render_alphas"][0, ys, xs, 0] > 0
        for oi, nm in ((5, "last_ids"), (6, "median_ids")):
            exp = torch.where(hit, _global_ids(r[nm][0, ys, xs], tiles, ltiles, c, ys, xs, ts),
                              torch.zeros_like(r[nm][0, ys, xs]))
            got_ids = fo[oi][c, ys, xs].cpu()
            if not torch.equal(got_ids[keep], exp[keep]):
                failures.append({"output": nm, "camera": c, "tile": [ys.start, xs.start],
                                 "mismatches": int((got_ids != exp)[keep].sum())})
    stats["fragile_pixels"] = n_frag / max(n_pix, 1)
    stats["pixels_over_strict"] = frag_over
    stats["fwd_worst"] = worst

    if backward:
        # backward: gradients of the referenced entries
        fr_rows = outs64[0][3]["fragile_rows"][0].clone()
        fr_own = outs64[0][3]["fragile_own"][0].clone()
        cond = outs64[0][3]["cond_max"][0].clone()
        edge = outs64[0][3]["texel_edge"][0].clone()
        for (_, _, _, r) in outs64[1:]:
            fr_rows |= r["fragile_rows"][0]
            fr_own |= r["fragile_own"][0]
            cond = torch.maximum(cond, r["cond_max"][0])
            edge |= r["texel_edge"][0]
        norm_rows = fr_own | (((cond > 1e3) | edge) if uv != "none" else torch.zeros_like(fr_rows))
        info = {"entry": u, "fragile_row": fr_rows, "fragile_own": fr_own, "cond_max": cond, "texel_edge": edge}
        stats["fragile_rows"] = float(fr_rows.float().mean())
        g32 = (lambda k: L32[k].grad[0]) if L32 is not None else (lambda k: None)
        ud = u.to(DEV)
        stats["norm_only_rows"] = float(norm_rows.float().mean())
        for nm, got, key in (("colors", v_col.reshape(E, K), "col"), ("opacities", v_op.reshape(E, 1), "op"),
                             ("normals", v_nrm.reshape(E, 3), "nrm"), ("means2d", v_m2.reshape(E, 2), "m2"),
                             ("ray_transforms", v_rt.reshape(E, 9), "rt")):
            ref = L64[key].grad[0].reshape(Gn, -1)
            r32 = None if L32 is None else L32[key].grad[0].reshape(Gn, -1)
            stats[nm] = check_grad(nm, got[ud], ref, r32, fr_rows, dtype, info, failures, norm_rows)
        if textured:
            # a Gaussian seen by several sampled cameras has one texture row on the GPU but one
            # copy per (camera, Gaussian) in the compact problem: sum the copies
            rows, inv = torch.unique(row_of, return_inverse=True)

            def tex_sum(t):
                return torch.zeros(rows.numel(), *t.shape[1:], dtype=t.dtype).index_add_(0, inv, t)
            fr_tex = torch.zeros(rows.numel(), dtype=torch.bool).index_put_((inv,), fr_rows, accumulate=True)
            nr_tex = torch.zeros(rows.numel(), dtype=torch.bool).index_put_((inv,), norm_rows, accumulate=True)
            stats["textures"] = check_grad("textures", v_tex[rows.to(v_tex.device)], tex_sum(L64["tex"].grad),
                                           None if L32 is None else tex_sum(L32["tex"].grad.double()), fr_tex,
                                           dtype, {"row": rows, "fragile_row": fr_tex}, failures, nr_tex)
            # texture rows outside the sampled tiles' lists: exactly zero
            mask = torch.ones(N, dtype=torch.bool, device=v_tex.device)
            mask[rows.to(v_tex.device)] = False
            z = 0.0
            for a0 in range(0, N, 1 << 16):
                blk = v_tex[a0:a0 + (1 << 16)]
                m = mask[a0:a0 + (1 << 16)]
                if m.any():
                    z = max(z, float(blk[m].abs().max()))
            if z != 0.0:
                failures.append({"tensor": "textures", "unlisted_rows_max": z})
        if bg is not None:
            # background: one row per camera, the sum over the sampled tiles' pixels of
            # T_final * v_colors; cameras with a fragile pixel may use the loose bound
            cam_frag = torch.zeros(C, dtype=torch.bool)
            for (c, ys, xs, r) in outs64:
                cam_frag[c] |= bool(r["fragile"][0, ys, xs].any())
            stats["backgrounds"] = check_grad("backgrounds", v_bg, L64["bg"].grad,
                                              None if L32 is None else L32["bg"].grad, cam_frag, dtype, None, failures)
        # per-entry gradients of unreferenced entries: exactly zero
        others = torch.ones(E, dtype=torch.bool)
        others[u] = False
        od = others.to(DEV)
        for nm, t in (("colors", v_col), ("opacities", v_op), ("means2d", v_m2), ("ray_transforms", v_rt)):
            tt = t.reshape(E, -1)[od]
            if tt.numel() and float(tt.abs().max()) != 0.0:
                failures.append({"tensor": nm, "unreferenced_max": float(tt.abs().max())})
    # release the GPU outputs before a failure's traceback can keep them (tens of GB at size)
    fo = v_m2 = v_rt = v_col = v_op = v_tex = v_nrm = v_bg = tex_leaf = bg_leaf = None  # noqa: F841
    textures = colors = opac = tex_rows = L64 = L32 = outs64 = outs32 = st = None  # noqa: F841
    if failures:
        _report({"case": f"C={C} N={N} E={E} {W}x{H} K={K} tex={textured} {dtype} uv={uv} packed={packed} "
                         f"filter={filtering}", "stats": stats,
                 "failures": failures[:40]})
    assert not failures, (stats, failures[:6])
    return stats


def _global_ids(local, tiles, ltiles, c, ys, xs, ts):
    """Map list positions of the compact problem back to positions in the full lists."""
    for (cc, ty, tx, a, b), (_, _, _, la, lb) in zip(tiles, ltiles):
        if cc == c and ys.start == ty * ts and xs.start == tx * ts:
            return (local.to(torch.int64) - la + a).to(torch.int32)
    raise KeyError((c, ys, xs))
