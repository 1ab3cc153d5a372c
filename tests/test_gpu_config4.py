"""The neural-texture form at full size: the headline scene's 1,000,000
Gaussians with textures decoded from per-Gaussian 32-d latents N(0, 1) by the
32 -> 128 -> 128 -> 1024 decoder (16x16 RGBA, bfloat16 matrix cores, float32
accumulation).

- The fused inference kernel over all 1M latents, against the oracle on a sample
  of latents (the same bound as the small-size test: every output within 2e-3,
  mean error under 1e-6, under 2% of latents beyond 1e-5; the outliers are hidden
  activations within float32 error of a bfloat16 rounding boundary).
- The training path's backward over all 1M rows with the upstream gradient
  nonzero only on sampled rows: the latent gradients of those rows and the weight
  and bias gradients (sums over the sampled rows only) against the oracle, 2e-3
  in the L2 norm as the small-size test.
- Decode -> render: the decoded textures through the rasterizer forward and
  backward on sampled tiles against the oracle (tolerances of gpu_util).
"""
import numpy as np
import pytest
import torch

from oracle.neural import mlp_backward, mlp_forward
from textured_gaussians_amd import neural, scenes
from textured_gaussians_amd.cuda import _ops, _wrapper

from gpu_util import DEV, gpu_stages, raster_tile_parity, sample_tiles

pytestmark = pytest.mark.gpu

N = 1_000_000


@pytest.fixture(scope="module")
def model():
    torch.cuda.empty_cache()
    z = torch.randn(N, 32, generator=torch.Generator().manual_seed(2)).to(DEV)
    dec = neural.NeuralTextureDecoder(device=DEV, seed=3)
    yield z, dec
    torch.cuda.empty_cache()


def _params(dec):
    return [t.detach().double().cpu() for t in (dec.W1, dec.b1, dec.W2, dec.b2, dec.W3, dec.b3)]


def test_config4_fused_decode_sampled_latents(model):
    z, dec = model
    packed = _ops.decoder_pack(dec.W1.detach(), dec.W2.detach(), dec.W3.detach())
    out = _ops.decode_textures(z, packed, dec.b1.detach(), dec.b2.detach(), dec.b3.detach())
    assert out.shape == (N, 16, 16, 4)
    assert bool(torch.isfinite(out).all()) and float(out.min()) > 0.0 and float(out.max()) < 1.0
    rows = torch.cat([torch.randperm(N, generator=torch.Generator().manual_seed(4))[:4096],
                      torch.arange(N - 70, N)])  # the tail: the last, partial 64-latent block
    ref, _, _ = mlp_forward(z[rows.to(DEV)].double().cpu(), *_params(dec))
    d = (out[rows.to(DEV)].reshape(rows.numel(), -1).double().cpu() - ref).abs()
    assert float(d.max()) < 2e-3 and float(d.mean()) < 1e-6
    assert float((d.amax(1) > 1e-5).double().mean()) < 0.02


def test_config4_decoder_backward_sampled_rows(model):
    z, dec = model
    zz = z.clone().requires_grad_(True)
    for p in dec.parameters():
        p.grad = None
    tex = neural.decode_mlp(zz, dec.W1, dec.b1, dec.W2, dec.b2, dec.W3, dec.b3)  # training path, 1M rows
    g = torch.Generator().manual_seed(5)
    rows = torch.randperm(N, generator=g)[:2048]
    gout = torch.zeros(N, 1024, device=DEV)
    gsel = torch.randn(rows.numel(), 1024, generator=g)
    gout[rows.to(DEV)] = gsel.to(DEV)
    tex.backward(gout)
    P = _params(dec)
    zr = z[rows.to(DEV)].double().cpu()
    ref_out, h1, h2 = mlp_forward(zr, *P)
    ref = mlp_backward(gsel.double(), zr, P[0], P[2], P[4], h1, h2, ref_out)
    mine = [zz.grad[rows.to(DEV)], dec.W1.grad, dec.b1.grad, dec.W2.grad, dec.b2.grad, dec.W3.grad, dec.b3.grad]
    for name, m, r in zip(["dz", "dW1", "db1", "dW2", "db2", "dW3", "db3"], mine, ref):
        e = float((m.double().cpu() - r).norm() / r.norm())
        assert e < 2e-3, (name, e)
    # rows without upstream gradient get none
    other = torch.ones(N, dtype=torch.bool)
    other[rows] = False
    assert float(zz.grad[other.to(DEV)].abs().max()) == 0.0


def test_config4_decode_then_render_sampled_tiles(model):
    z, dec = model
    with torch.no_grad():
        tex = dec(z).contiguous()  # the fused kernel
    sc = scenes.config2(tex_res=None)
    st = gpu_stages(sc)
    s = st["scene"]
    with torch.no_grad():
        camc = torch.linalg.inv(s.viewmats)[:, None, :3, 3]
        col = _wrapper.spherical_harmonics_views(3, s.means[None] - camc, s.colors, masks=st["radii"] > 0)
        colors = torch.clamp_min(col + 0.5, 0.0).contiguous()
    tiles = sample_tiles(st, 10, extra=[(0, st["th"] - 1, st["tw"] - 1)])
    stats = raster_tile_parity(st, colors, s.opacities[None].contiguous(), tex, s.backgrounds, torch.float32, tiles)
    print("config4 render", stats)
