"""The full training iteration at size: N = 2,000,000 Gaussians, 64x64 RGBA
textures stored bfloat16 with a float32 master and bfloat16 moments (262 GB of
texture state), one 1920x1080 camera, SH degree 3, 0.8 L1 + 0.2 D-SSIM, selective
Adam on every parameter (the textures stepped inside the backward).

- The bfloat16-texture rasterizer forward on sampled tiles against the oracle fed
  the same texels widened to float64 (tolerances of gpu_util).  A dense texture
  gradient (131 GB in float32) does not fit beside the state, so the backward at
  this size is checked through the step's effects, and against the dense path at
  small size (tests/test_gpu_texture_adam.py).
- One iteration: the loss is finite; every texture row outside the step's tile
  lists keeps every bit of its four planes (checksums of all rows, and an exact
  copy of a sample of rows); rows in the lists are stepped; geometry moves.
"""
import pytest
import torch

from textured_gaussians_amd import scenes, train
from textured_gaussians_amd.cuda import _ops, _wrapper

from gpu_util import DEV, gpu_stages, raster_tile_parity, sample_tiles

pytestmark = pytest.mark.gpu

N = 2_000_000


def _row_checksums(plane, chunk=8192):
    """Per-row weighted sum of the 16-bit words (int64), chunk by chunk."""
    out = torch.empty(plane.shape[0], dtype=torch.int64, device=plane.device)
    n = plane[0].numel() // 2
    w = (torch.arange(n, device=plane.device, dtype=torch.int64) * 2654435761 + 12345) % (1 << 31)
    for a in range(0, plane.shape[0], chunk):
        x = plane[a:a + chunk].reshape(-1, n * 2).view(torch.int32).to(torch.int64)
        out[a:a + chunk] = (x * w).sum(1)
    return out


def test_config5_bf16_render_and_train_step():
    torch.cuda.empty_cache()
    sc = scenes.config5().to(DEV)
    texopt = train.scene_texture_adam(N, 64, DEV)
    assert texopt.textures.dtype == torch.bfloat16 and texopt.textures.shape == (N, 64, 64, 4)
    st = gpu_stages(sc)
    s = st["scene"]
    with torch.no_grad():
        camc = torch.linalg.inv(s.viewmats)[:, None, :3, 3]
        col = _wrapper.spherical_harmonics_views(3, s.means[None] - camc, s.colors, masks=st["radii"] > 0)
        colors = torch.clamp_min(col + 0.5, 0.0).contiguous()
    tiles = sample_tiles(st, 10, extra=[(0, st["th"] - 1, st["tw"] - 1)])
    stats = raster_tile_parity(st, colors, s.opacities[None].contiguous(), texopt.textures, s.backgrounds,
                               torch.float32, tiles, backward=False)
    print("config5 bf16 forward", stats)

    listed = _ops.mark_rows(st["flatten_ids"], N).bool()
    assert 0 < int(listed.sum()) < N
    del st, colors
    torch.cuda.empty_cache()
    planes = (texopt.textures, texopt.lo, texopt.exp_avg, texopt.exp_avg_sq)
    before = [_row_checksums(p) for p in planes]
    unl = torch.nonzero(~listed).flatten()
    samp = unl[torch.randperm(unl.numel(), generator=torch.Generator().manual_seed(1))[:2000].to(DEV)]
    copies = [p[samp].clone() for p in planes]
    state = train.init_state(sc.means, sc.quats, sc.scales, sc.opacities, sc.colors, texopt)
    m0 = state.params["means"].detach().clone()
    loss = float(train.train_step(state, sc.viewmats, sc.Ks, sc.width, sc.height, sc.target, sc.backgrounds))
    assert loss == loss and abs(loss) < float("inf")
    assert texopt.steps == 1
    after = [_row_checksums(p) for p in planes]
    for b, a in zip(before, after):
        assert torch.equal(a[~listed], b[~listed])
    for p, c in zip(planes, copies):
        assert torch.equal(p[samp].view(torch.int16), c.view(torch.int16))
    # the listed rows took the step: their second moments left zero, their masters moved
    moved = (after[1] != before[1]) | (after[0] != before[0])
    assert float(moved[listed].float().mean()) > 0.5
    assert bool((texopt.exp_avg_sq[listed.nonzero()[:64].flatten()].float().abs().sum() > 0))
    assert float((state.params["means"].detach() - m0).abs().max()) > 0
