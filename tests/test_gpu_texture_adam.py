"""Bfloat16 textures and their row-wise Adam step on the GPU.

- The rasterizer with bfloat16 textures computes exactly what it computes with
  float32 textures holding the same values (the texels widen exactly).
- The records-only backward plus tgs_texture_adam_rows equals the dense float32
  texture gradient followed by the reference's selective Adam in float32
  (adam.cu:16-47 arithmetic: no bias correction): the float32 master to float32
  rounding, the moments to their bfloat16 rounding; invisible rows keep every bit.
- The moments' stochastic rounding is unbiased where round-to-nearest is not.
- The config-5 training step (bf16 textures, L1 + D-SSIM, fused Adam) runs."""
import pytest
import torch

import textured_gaussians_amd as tg
from textured_gaussians_amd import scenes
from textured_gaussians_amd.cuda import _ops
from textured_gaussians_amd.texture_adam import BF16TextureAdam, join_master, split_master

from gpu_util import DEV, gpu_stages, rel_err

pytestmark = pytest.mark.gpu


def _setup(tex_res=8, seed=6, n=2500):
    sc = scenes.config1(n=n, width=80, height=64, tex_res=tex_res, seed=seed, sh_degree=None)
    st = gpu_stages(sc, torch.float32)
    s = st["scene"]
    C, N = st["radii"].shape
    g = torch.Generator().manual_seed(seed + 2)
    colors = torch.rand(C, N, 3, generator=g).to(DEV)
    opac = sc.opacities.float()[None].repeat(C, 1).to(DEV)
    master = (sc.textures.float() + 0.05 * torch.randn(sc.textures.shape, generator=g)).to(DEV)
    vc = torch.randn(C, s.height, s.width, 3, generator=g).to(DEV)
    va = torch.randn(C, s.height, s.width, 1, generator=g).to(DEV)
    return st, colors, opac, master, vc, va


def _fwd(st, colors, opac, tex, **kw):
    s = st["scene"]
    return _ops.rasterize_to_pixels_fwd_textured_gaussians_full(
        st["means2d"], st["rt"], colors, opac, tex, 3.0, 3.0, st["normals"], None, None, s.width, s.height, 16,
        st["offsets"], st["flatten_ids"], 0.0, 1.0, **kw)


def _bwd(st, colors, opac, tex, fwd, vc, va, **kw):
    s = st["scene"]
    C, H, W = vc.shape[:3]
    z3 = torch.zeros(C, H, W, 3, device=DEV)
    z1 = torch.zeros(C, H, W, 1, device=DEV)
    return _ops.rasterize_to_pixels_bwd_textured_gaussians(
        st["means2d"], st["rt"], colors, opac, tex, 3.0, 3.0, st["normals"], torch.zeros_like(st["means2d"]), None,
        None, W, H, 16, st["offsets"], st["flatten_ids"], 1.0, fwd[0], fwd[1], fwd[5], fwd[6], vc, va, z3, z1, z1,
        False, **kw)


def test_bf16_textures_render_and_backprop_as_their_float32_values():
    st, colors, opac, master, vc, va = _setup()
    hi = master.to(torch.bfloat16)
    f_b = _fwd(st, colors, opac, hi)
    f_f = _fwd(st, colors, opac, hi.float())
    for i, (a, b) in enumerate(zip(f_b, f_f)):
        if i in (7, 8):  # per-Gaussian contribution sums: float atomics, order varies
            assert rel_err(a, b) < 1e-6
        else:  # per-pixel outputs: identical arithmetic
            assert torch.equal(a, b), i
    b_b = _bwd(st, colors, opac, hi, f_b, vc, va)
    b_f = _bwd(st, colors, opac, hi.float(), f_f, vc, va)
    assert b_b[5].dtype == torch.float32  # the texture gradient stays float32
    for a, b in zip(b_b[1:], b_f[1:]):
        assert rel_err(a, b) < 1e-5  # float atomics in another order
    with pytest.raises(RuntimeError):
        _fwd(st, colors.double(), opac.double(), hi)  # bf16 textures need float32 compute


def _adam_state(master, seed=9):
    g = torch.Generator(device=DEV).manual_seed(seed)
    hi, lo = split_master(master)
    m = (torch.randn(master.shape, generator=g, device=DEV) * 1e-3).to(torch.bfloat16)
    v = (torch.rand(master.shape, generator=g, device=DEV) * 9e-6 + 1e-6).to(torch.bfloat16)
    return hi.contiguous(), lo.contiguous(), m, v


def _bf16_ulp(x):
    e = (x.abs().view(torch.int32) & 0x7F800000).view(torch.float32)
    return torch.clamp(e, min=2.0 ** -126) * 2.0 ** -7


def test_texture_adam_rows_matches_dense_gradient_and_float32_adam():
    st, colors, opac, master, vc, va = _setup()
    N = master.shape[0]
    hi, lo, m, v = _adam_state(master)
    cnt = _ops.tex_counts_buffer(hi)
    fwd = _fwd(st, colors, opac, hi, tex_counts=cnt)
    recs = _bwd(st, colors, opac, hi, fwd, vc, va, v_textures_out=False, tex_counts=cnt)[5]
    assert torch.equal(recs[2], recs[0])  # the backward wrote exactly the pairs the forward counted
    g = _bwd(st, colors, opac, hi.float(), _fwd(st, colors, opac, hi.float()), vc, va)[5]
    vis = _ops.mark_rows(st["flatten_ids"], N)
    extra = torch.arange(N, device=DEV) % 7 == 0  # visible rows without records: zero gradient
    vis = (vis.bool() | extra).to(torch.uint8).contiguous()
    assert 0 < int(vis.sum()) < N
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    # reference: float32 selective Adam on the float32 master and widened moments
    p = join_master(hi, lo)
    mf, vf = m.float(), v.float()
    m_ref = b1 * mf + (1 - b1) * g
    v_ref = b2 * vf + (1 - b2) * g * g
    p_ref = p - lr * m_ref / (torch.sqrt(v_ref) + eps)
    rows = vis.bool()
    h2, l2, m2, v2 = hi.clone(), lo.clone(), m.clone(), v.clone()
    _ops.texture_adam_rows(recs, vis, h2.view(torch.int16), l2, m2.view(torch.int16), v2.view(torch.int16), lr, b1,
                           b2, eps, 1234)
    torch.cuda.synchronize()
    # invisible rows: every plane bit-identical
    for a, b in ((h2, hi), (l2, lo), (m2, m), (v2, v)):
        assert torch.equal(a[~rows].view(torch.int16), b[~rows].view(torch.int16))
    # visible rows: master to float32 rounding of the same update
    p2 = join_master(h2, l2)
    assert float((p2[rows] - p_ref[rows]).abs().max()) < 2e-6
    assert float((p2[rows] - p[rows]).abs().max()) > 1e-3  # it moved
    # the render copy is the master rounded to nearest
    hh, ll = split_master(p2)
    assert torch.equal(hh.view(torch.int16), h2.view(torch.int16)) and torch.equal(ll, l2)
    # moments: one of the two bfloat16 neighbours of the kernel's float32 value, which is
    # the reference's to float32 rounding (2 ulps where that crosses a binade boundary)
    # plus what the gradient's summation order moves them (|dg| <= 1e-5 max|g|)
    dg = 1e-5 * float(g.abs().max())
    g_r = g[rows].flatten().abs()
    for name, got, ref, slack in (("m", m2, m_ref, (1 - b1) * dg), ("v", v2, v_ref, (1 - b2) * 2 * g_r * dg)):
        gr, rr = got.float()[rows].flatten(), ref[rows].flatten()
        d = (gr - rr).abs()
        u = _bf16_ulp(rr)
        k = int(((d - slack) / u).argmax())
        assert bool((d <= 2 * u + slack).all()), (name, float(gr[k]), float(rr[k]), float(d[k] / u[k]))
        assert float((d <= 1.01 * u).float().mean()) > 0.999, name


def test_texture_adam_moment_rounding_is_unbiased():
    """Visible rows with no records: m <- b1 m, v <- b2 v exactly, then rounded.  b2 v
    lies 0.1% below v, inside half a bfloat16 ulp: round-to-nearest would return v
    unchanged every step; the stochastic rounding's mean is b2 v."""
    N, res = 4096, 8
    c_m, c_v = 0.0625, 0.015625  # exact in bfloat16
    hi = torch.zeros(N, res, res, 4, dtype=torch.bfloat16, device=DEV)
    lo = torch.zeros(N, res, res, 4, dtype=torch.int16, device=DEV)
    m = torch.full_like(hi, c_m)
    v = torch.full_like(hi, c_v)
    recs = (torch.zeros(N, dtype=torch.int32, device=DEV), torch.zeros(N, dtype=torch.int64, device=DEV),
            torch.zeros(N, dtype=torch.int32, device=DEV), torch.zeros(32, dtype=torch.uint8, device=DEV))
    vis = torch.ones(N, dtype=torch.uint8, device=DEV)
    b1, b2 = 0.9, 0.999
    _ops.texture_adam_rows(recs, vis, hi.view(torch.int16), lo, m.view(torch.int16), v.view(torch.int16), 1e-3, b1,
                           b2, 1e-8, 77)
    n = m.numel()
    ulp_v = c_v * 2.0 ** -7
    mean_v = float(v.double().mean())
    assert abs(mean_v - b2 * c_v) < 0.02 * ulp_v, (mean_v, b2 * c_v)
    assert float((v.float() < c_v).double().mean()) > 0.05  # some elements moved down
    ulp_m = c_m * 2.0 ** -8  # b1 m lies in the binade below
    assert abs(float(m.double().mean()) - b1 * c_m) < 0.02 * ulp_m
    # the parameter moved by -lr * m' / (sqrt(v') + eps) with the unrounded moments
    want = -1e-3 * (b1 * c_m) / ((b2 * c_v) ** 0.5 + 1e-8)
    p = join_master(hi, lo)
    assert float((p - want).abs().max()) < 1e-7 and n == N * res * res * 4


def test_texture_adam_through_the_rendering_api():
    """texture_adam= steps exactly the rows in the tile lists; the other gradients are
    those of the same call without it."""
    sc = scenes.config1(n=2000, width=96, height=64, tex_res=8, seed=3, sh_degree=None).to(DEV)
    opt = BF16TextureAdam.from_textures(sc.textures.float(), lr=1e-2)
    before = [t.clone() for t in (opt.textures, opt.lo, opt.exp_avg, opt.exp_avg_sq)]
    grads = []
    for use in (False, True):
        means = sc.means.clone().requires_grad_(True)
        colors = sc.colors.clone().requires_grad_(True)
        out = tg.rasterization_textured_gaussians(means, sc.quats, sc.scales, sc.opacities, colors, opt.textures,
                                                  sc.viewmats, sc.Ks, sc.width, sc.height,
                                                  backgrounds=sc.backgrounds, texture_adam=opt if use else None)
        (out[0] - sc.target).abs().mean().backward()
        grads.append((means.grad, colors.grad, out[-1]["flatten_ids"]))
    assert rel_err(grads[1][0], grads[0][0]) < 1e-5 and rel_err(grads[1][1], grads[0][1]) < 1e-5
    rows = _ops.mark_rows(grads[1][2], sc.means.shape[0]).bool()
    assert opt.steps == 1 and 0 < int(rows.sum()) < rows.numel()
    for a, b in zip((opt.textures, opt.lo, opt.exp_avg, opt.exp_avg_sq), before):
        assert torch.equal(a[~rows].view(torch.int16), b[~rows].view(torch.int16))
    assert bool((opt.exp_avg_sq[rows].float().abs().sum(dim=(1, 2, 3)) > 0).any())


def test_training_step_config5_shape():
    """The full iteration (bf16 64x64 textures, SH3, L1 + 0.2 D-SSIM, fused Adam on every
    parameter) on a small scene of the config-5 kind: finite loss, textures and
    geometry stepped."""
    from textured_gaussians_amd import train
    sc = scenes.config2(n=20_000, width=192, height=108, tex_res=None).to(DEV)
    texopt = train.scene_texture_adam(sc.means.shape[0], 64, DEV)
    state = train.init_state(sc.means, sc.quats, sc.scales, sc.opacities, sc.colors, texopt)
    m0 = state.params["means"].detach().clone()
    t0 = texopt.master().clone()
    losses = [float(train.train_step(state, sc.viewmats, sc.Ks, sc.width, sc.height, sc.target, sc.backgrounds))
              for _ in range(3)]
    assert all(torch.isfinite(torch.tensor(losses)))
    assert texopt.steps == 3
    assert float((state.params["means"].detach() - m0).abs().max()) > 0
    assert float((texopt.master() - t0).abs().max()) > 0


def test_texture_adam_in_the_packed_layout_matches_the_grid():
    """texture_adam= with packed=True over three cameras: each entry's records go to its
    Gaussian's row, so one step moves the textures, their moments and the geometry
    gradients as the grid layout does (summation order aside)."""
    cams = [(3.5, 0.3, 0.4), (-3.0, 1.5, 0.8), (0.5, -3.6, -0.6)]
    sc = scenes.make_scene(1500, 72, 56, 50.0, cams, 8, logscale_mu=-3.0, logscale_sigma=0.4, sh_degree=1,
                           seed=8).to(DEV)
    res = []
    for packed in (False, True):
        opt = BF16TextureAdam.from_textures(sc.textures.float(), lr=1e-2)
        means = sc.means.clone().requires_grad_(True)
        out = tg.rasterization_textured_gaussians(means, sc.quats, sc.scales, sc.opacities, sc.colors, opt.textures,
                                                  sc.viewmats, sc.Ks, sc.width, sc.height, sh_degree=1,
                                                  backgrounds=sc.backgrounds, texture_adam=opt, packed=packed)
        (out[0] - sc.target).abs().mean().backward()
        assert opt.steps == 1
        res.append((opt.master(), opt.exp_avg.float(), opt.exp_avg_sq.float(), means.grad))
    (p0, m0, v0, g0), (p1, m1, v1, g1) = res
    assert float((p0 - sc.textures.float()).abs().max()) > 0  # the step moved texels
    # master values agree to float32 summation order; moments to their bfloat16 rounding
    assert rel_err(p1, p0) < 1e-6
    assert rel_err(m1, m0) < 1e-2 and rel_err(v1, v0) < 1e-2
    assert rel_err(g1, g0) < 1e-5
