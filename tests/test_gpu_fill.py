"""The native zero-fill used for the texture gradient: exact extent, any size
that is a multiple of 4 bytes, neighbours untouched."""
import pytest
import torch

from textured_gaussians_amd.cuda import _ops

from gpu_util import DEV

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 3, 4, 5, 1023, 250_001, 3_000_003])
def test_zero_fill_exact_extent(n):
    buf = torch.rand(n + 16, device=DEV) + 1.0  # all entries >= 1
    view = buf[4:4 + n]  # 16-byte aligned start
    _ops.zero_fill_(view)
    torch.cuda.synchronize()
    assert torch.all(view == 0)
    assert torch.all(buf[:4] >= 1) and torch.all(buf[4 + n:] >= 1)


def test_zero_fill_rejects_misaligned():
    buf = torch.ones(64, device=DEV)
    with pytest.raises(RuntimeError, match="aligned"):
        _ops.zero_fill_(buf[1:9])


def test_mark_rows_and_row_selective_fill():
    N = 10
    flat = torch.tensor([0, 5, 5, 7 + N, 2 * N + 9], dtype=torch.int32, device=DEV)  # ids of cameras 0..2
    mask = _ops.mark_rows(flat, N)
    assert mask.cpu().tolist() == [1, 0, 0, 0, 0, 1, 0, 1, 0, 1]
    buf = torch.full((N, 3, 2, 4), 7.0, device=DEV)
    _ops.zero_fill_rows_(buf, mask, 1)
    torch.cuda.synchronize()
    rows = buf.reshape(N, -1)
    sel = mask.bool().cpu()
    assert torch.all(rows[sel] == 0) and torch.all(rows[~sel] == 7)
    _ops.zero_fill_rows_(buf, mask, 0)
    assert torch.all(buf == 0)


def test_fused_fill_with_no_intersections():
    # no tile lists: the forward and backward kernels have nothing to fuse the fill into
    N, C, W, H = 6, 1, 32, 32
    f32 = torch.float32
    m2 = torch.zeros(C, N, 2, device=DEV)
    rt = torch.zeros(C, N, 3, 3, device=DEV)
    cols = torch.rand(C, N, 3, device=DEV)
    op = torch.rand(C, N, device=DEV)
    nrm = torch.zeros(C, N, 3, device=DEV)
    tex = torch.rand(N, 5, 5, 4, device=DEV)
    offs = torch.zeros(C, 2, 2, dtype=torch.int32, device=DEV)
    flat = torch.zeros(0, dtype=torch.int32, device=DEV)
    v_tex = torch.full_like(tex, float("nan"))
    mask = _ops.mark_rows(flat, N)
    out = _ops.rasterize_to_pixels_fwd_textured_gaussians_full(
        m2, rt, cols, op, tex, 3.0, 3.0, nrm, None, None, W, H, 16, offs, flat, 0.0, 1.0, zero_rows=(v_tex, mask))
    g = [torch.ones_like(out[0]), torch.ones_like(out[1]), torch.zeros_like(out[2]), torch.zeros_like(out[3]),
         torch.zeros_like(out[4])]
    res = _ops.rasterize_to_pixels_bwd_textured_gaussians(
        m2, rt, cols, op, tex, 3.0, 3.0, nrm, torch.zeros_like(m2), None, None, W, H, 16, offs, flat, 1.0,
        out[0], out[1], out[5], out[6], *g, False, v_textures_out=v_tex, zero_rows=(v_tex, mask))
    torch.cuda.synchronize()
    assert torch.all(res[5] == 0)
    assert out[0].dtype == f32 and torch.all(out[1] == 0)
