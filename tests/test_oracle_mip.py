"""The oracle's mipmap pyramid and trilinear (``mipmapped2``) sampler, pinned by
closed forms: each pyramid level is the block mean of the finest texture; at an
exact integer level of detail the sampler is the bilinear sampler (itself pinned
against the reference's golden renders) on that level with texel-centre-aligned
coordinates; between levels it interpolates linearly in the level; the pyramid's
backward is the adjoint (autograd)."""
import math

import pytest
import torch

from oracle.raster import _bilinear, _mip2, generate_mipmap, mip_base


def _tex(n=3, L=4, seed=0):
    g = torch.Generator().manual_seed(seed)
    return torch.rand(n, 1 << L, 1 << L, 4, generator=g, dtype=torch.float64)


def test_pyramid_levels_are_block_means():
    t = _tex()
    mip = generate_mipmap(t)
    L = 4
    assert mip.shape == (3, (4 ** (L + 1) - 1) // 3, 4)
    for lvl in range(L + 1):
        r = (1 << L) >> lvl
        b = 1 << lvl
        want = t.reshape(3, r, b, r, b, 4).mean(dim=(2, 4)).reshape(3, r * r, 4)
        got = mip[:, mip_base(L, lvl):mip_base(L, lvl) + r * r]
        assert torch.allclose(got, want, atol=1e-14, rtol=0), lvl
    assert mip_base(L, L) == 0 and mip_base(L, 0) == (4 ** L - 1) // 3


def test_pyramid_backward_is_the_adjoint():
    t = _tex(n=2, L=3, seed=1).requires_grad_(True)
    v = torch.randn(2, (4 ** 4 - 1) // 3, 4, generator=torch.Generator().manual_seed(2), dtype=torch.float64)
    (generate_mipmap(t) * v).sum().backward()
    # adjoint by hand: a texel receives its own cotangent plus 4^-k of each ancestor's
    want = torch.zeros_like(t)
    for lvl in range(4):
        r = 8 >> lvl
        vl = v[:, mip_base(3, lvl):mip_base(3, lvl) + r * r].reshape(2, r, r, 4)
        want += vl.repeat_interleave(1 << lvl, 1).repeat_interleave(1 << lvl, 2) / 4 ** lvl
    assert torch.allclose(t.grad, want, atol=1e-14, rtol=0)


def _s_for_u(u, res, r):
    return u / (res - 1) * 2 * r - r


def test_mip2_at_integer_lod_is_bilinear_on_that_level():
    L, rx = 4, 3.0
    t = _tex(n=2, L=L, seed=3)
    mip = generate_mipmap(t)
    res = 1 << L
    g = torch.Generator().manual_seed(4)
    for lvl in (0, 1, 2):
        sc = 1 << lvl
        r = res >> lvl
        # texel coordinates at level lvl well inside it, and the finest-level u they come from
        ul = torch.rand(5, 2, generator=g, dtype=torch.float64) * (r - 1.2) + 0.1
        vl = torch.rand(5, 2, generator=g, dtype=torch.float64) * (r - 1.2) + 0.1
        u = (ul + 0.5) * sc - 0.5
        v = (vl + 0.5) * sc - 0.5
        s = torch.stack([_s_for_u(u, res, rx), _s_for_u(v, res, rx)], -1)
        # |duv/dx| = 2^lvl exactly: ds = 2^lvl * 2 rx / (res - 1) along x
        d = torch.full_like(s[..., 0], sc * 2 * rx / (res - 1))
        dsdx = torch.stack([d, torch.zeros_like(d)], -1)
        dsdy = torch.zeros_like(dsdx)
        n = torch.tensor([0, 1])
        got, ok, _ = _mip2(mip, n, s, dsdx, dsdy, L, rx, rx)
        assert bool(ok.all())
        lev = mip[:, mip_base(L, lvl):mip_base(L, lvl) + r * r].reshape(2, r, r, 4)
        want, inr = _bilinear(lev, n, torch.stack([_s_for_u(ul, r, rx), _s_for_u(vl, r, rx)], -1), rx, rx, "none")
        assert bool(inr.all())
        assert torch.allclose(got, want, atol=1e-12, rtol=0), lvl


def test_mip2_interpolates_linearly_between_levels():
    L, rx = 4, 3.0
    mip = generate_mipmap(_tex(n=1, L=L, seed=5))
    res = 1 << L
    s = torch.tensor([[[0.4, -0.7]]], dtype=torch.float64)
    n = torch.tensor([0])

    def at(lod):
        d = torch.full((1, 1), 2.0 ** lod * 2 * rx / (res - 1), dtype=torch.float64)
        dsdx = torch.stack([torch.zeros_like(d), d], -1)  # |dsdx| counts, whatever the axis
        return _mip2(mip, n, s, dsdx, torch.zeros_like(dsdx), L, rx, rx)[:2]

    s1, ok1 = at(1.0)
    s2, ok2 = at(2.0)
    sm, okm = at(1.0 + math.log2(1.5))  # t = 1.585: weight 0.585 on level 2
    f = math.log2(1.5)
    assert bool(ok1.all() and ok2.all() and okm.all())
    assert torch.allclose(sm, (1 - f) * s1 + f * s2, atol=1e-12, rtol=0)


def test_mip2_validity_shrinks_with_the_level():
    L, rx = 4, 3.0
    mip = generate_mipmap(_tex(n=1, L=L, seed=6))
    res = 1 << L
    n = torch.tensor([0])
    u = torch.tensor([[0.2]], dtype=torch.float64)  # near the left edge
    s = torch.stack([_s_for_u(u, res, rx), _s_for_u(u + 5, res, rx)], -1)
    small = torch.zeros(1, 1, 2, dtype=torch.float64)
    _, ok0, _ = _mip2(mip, n, s, small, small, L, rx, rx)
    d = torch.full((1, 1), 4.0 * 2 * rx / (res - 1), dtype=torch.float64)
    _, ok2, _ = _mip2(mip, n, s, torch.stack([d, d * 0], -1), small, L, rx, rx)
    # level 0: u = 0.2 has texels 0 and 1; level 2: (0.2 + 0.5) / 4 - 0.5 < 0 leaves the level
    assert bool(ok0.all()) and not bool(ok2.any())


def test_mip2_edge_marks_where_validity_flips():
    """The edge flag is set exactly where a small move of u or of the level of detail
    changes validity: u at a level's first texel, and t at an integer."""
    L, rx = 4, 3.0
    mip = generate_mipmap(_tex(n=1, L=L, seed=6))
    res = 1 << L
    n = torch.tensor([0])
    small = torch.zeros(1, 1, 2, dtype=torch.float64)

    def run(u, lod=None):
        uu = torch.tensor([[u]], dtype=torch.float64)
        s = torch.stack([_s_for_u(uu, res, rx), _s_for_u(uu + 5, res, rx)], -1)
        if lod is None:
            return _mip2(mip, n, s, small, small, L, rx, rx)
        d = torch.full((1, 1), 2.0 ** lod * 2 * rx / (res - 1), dtype=torch.float64)
        return _mip2(mip, n, s, torch.stack([d, d * 0], -1), small, L, rx, rx)
    # level 0, u just inside / outside the first texel: validity flips, both flagged
    _, ok_in, e_in = run(1e-6)
    _, ok_out, e_out = run(-1e-6)
    assert bool(ok_in.all()) and not bool(ok_out.any()) and bool(e_in.all()) and bool(e_out.all())
    # well inside: no flag
    _, ok, e = run(3.3)
    assert bool(ok.all()) and not bool(e.any())
    # u = 1.2 is valid at level 1 ((1.2 + 0.5) / 2 - 0.5 = 0.35) but not at level 2
    # (-0.075); t just below / above 1 flips it (level 2 enters the blend)
    _, ok_lo, e_lo = run(1.2, 1.0 - 1e-6)
    _, ok_hi, e_hi = run(1.2, 1.0 + 1e-6)
    assert bool(ok_lo.all()) and not bool(ok_hi.any()) and bool(e_lo.all()) and bool(e_hi.all())
    _, ok_mid, e_mid = run(1.2, 0.5)
    assert bool(ok_mid.all()) and not bool(e_mid.any())
