"""tgs_rasterize_fwd_noalloc (the forward whose workspace is the library's static
per-GPU scratch) against tgs_rasterize_fwd on the same inputs: every pixel output,
the live-entry bytes and the texture pair counts bit-identical, the per-Gaussian
contribution sums (float atomics, order-dependent) within 1e-6 relative.  Two calls
in a row, the second on a smaller scene, so the second reads scratch the first left
behind; then a request larger than the scratch is refused before anything runs."""
import ctypes

import pytest
import torch

from textured_gaussians_amd import scenes
from textured_gaussians_amd.cuda import _ops
from textured_gaussians_amd.cuda._backend import lib

from gpu_util import DEV, gpu_stages

SCRATCH = 64 << 20  # TGS_RASTERIZE_NOALLOC_SCRATCH_BYTES


def _fwd(st, colors, opac, tex, bg, noalloc):
    s = st["scene"]
    # bytes past the last entry a wave reached are left as they are: start both calls from one pattern
    live = _ops.live_entries_buffer(st["flatten_ids"], st["tile_size"]).fill_(0xA5)
    cnt = _ops.tex_counts_buffer(tex)
    out = _ops._raster_fwd(True, st["means2d"], st["rt"], colors, opac, tex, 3.0, 3.0, st["normals"], bg, None,
                           s.width, s.height, st["tile_size"], st["offsets"], st["flatten_ids"], 0.0, 1.0,
                           live_entries=live, tex_counts=cnt, noalloc=noalloc)
    return list(out) + [live, cnt]


def _case(n, w, h, tex_res, dtype, seed):
    sc = scenes.config1(n=n, width=w, height=h, tex_res=tex_res, seed=seed, sh_degree=None)
    st = gpu_stages(sc, dtype)
    C, N = st["radii"].shape
    g = torch.Generator().manual_seed(seed + 3)
    colors = torch.rand(C, N, 3, generator=g, dtype=dtype).to(DEV)
    opac = sc.opacities.to(DEV, dtype)[None].repeat(C, 1).contiguous()
    bg = torch.rand(C, 3, generator=g, dtype=dtype).to(DEV)
    return st, colors, opac, sc.textures.to(DEV, dtype).contiguous(), bg


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_forward_matches_workspace_forward(dtype):
    names = ["colors", "alphas", "normals", "distort", "median", "last_ids", "median_ids", "contrib_sum",
             "contrib_count", "accum_depth", "trans", "live_entries", "tex_counts"]
    # the first call's entries fill most of the scratch (58 MB of 64 MiB); the second is smaller
    for n, w, h, tex_res, seed in [(500_000 if dtype == torch.float32 else 250_000, 320, 240, 16, 1),
                                   (3_000, 80, 64, 12, 2)]:
        st, colors, opac, tex, bg = _case(n, w, h, tex_res, dtype, seed)
        ref = _fwd(st, colors, opac, tex, bg, noalloc=False)
        got = _fwd(st, colors, opac, tex, bg, noalloc=True)
        torch.cuda.synchronize()
        assert st["flatten_ids"].numel() > 0
        for name, a, b in zip(names, got, ref):
            if name.startswith("contrib"):
                assert torch.allclose(a, b, rtol=1e-6, atol=0), name
            else:
                assert torch.equal(a, b), (n, name)


def test_noalloc_refuses_more_than_the_scratch():
    """Host-side check: the entry point returns TGS_ERR_ARG before it touches a pointer
    or the GPU (null buffers are never read)."""
    L = lib()
    null = None
    for dtype, per in ((0, 29 * 4), (1, 29 * 8)):  # TGS_F32, TGS_F64: 29 values per entry
        too_many = SCRATCH // per + 1
        rc = L.tgs_rasterize_fwd_noalloc(dtype, 1, too_many, 3, *([null] * 6), 4, 4, 1.0, 1.0, null, null, null, 0,
                                         16, 16, 16, 1, 1, null, null, 0, 0.0, 1.0, *([null] * 16), 0, 0, null)
        assert rc == 1, rc
        assert b"static scratch" in L.tgs_last_error()
