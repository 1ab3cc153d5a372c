"""The CPU-runnable workload, whole frame: N = 10,000 Gaussians, 16x16 RGBA
textures, SH degree 0, one 256x256 camera (fov 40) at (4, 0, 0), 16x16 tiles,
white background.  Every tile of the frame goes through the oracle, in float64
and float32, for the forward and for the backward under the three texture
gradient rules (the reference's ``bilinear`` and ``bilinear_bwd2``, and the
exact one), with random upstream gradients on all five rendered outputs.
Tolerances are the ones stated in gpu_util."""
import pytest
import torch

from textured_gaussians_amd import scenes

from gpu_util import DEV, all_tiles, gpu_stages, raster_tile_parity

pytestmark = pytest.mark.gpu


def _stages(dtype):
    sc = scenes.config1()
    st = gpu_stages(sc, dtype)
    s = st["scene"]
    # SH degree 0: colour = max(0.2820948 c + 0.5, 0), one view
    st["colors"] = torch.clamp_min(0.28209479177387814 * s.colors[None, :, 0, :] + 0.5, 0.0).contiguous()
    st["opac"] = s.opacities[None].contiguous()
    return st


@pytest.mark.parametrize("dtype,uv,path", [(torch.float32, "none", "default"), (torch.float32, "double", "default"),
                                           (torch.float32, "exact", "default"), (torch.float64, "none", "default"),
                                           (torch.float64, "exact", "default"), (torch.float32, "none", "records"),
                                           (torch.float64, "exact", "records"), (torch.float32, "none", "atomics"),
                                           (torch.float32, "exact", "hybrid"), (torch.float64, "none", "hybrid"),
                                           (torch.float32, "none", "dense_fill"),
                                           (torch.float32, "none", "records_dense_fill"),
                                           (torch.float32, "none", "pair_samples")])
def test_config1_full_frame(dtype, uv, path, monkeypatch):
    """Texture-gradient paths: "records" - every texture pair through per-pair records and
    the row gather, the backward reading the texels, with the unlisted-row fill split
    between forward and backward; "records_dense_fill" - records, the unlisted rows
    zero-filled by one launch before the forward; "pair_samples" - float atomics with the
    backward reading the forward's texture samples instead of the texels; "atomics" - float
    atomics only; "hybrid" - rows with at least 16 pairs through records (many rows at this
    size), the others atomics; "dense_fill" - float atomics with the gradient zero-filled
    by one launch before the forward instead of along the rasterizer walks; "default" - as
    shipped (float32: records with the streaming gather, the backward reading the forward's
    pair samples under the reference rule; float64: float atomics)."""
    from textured_gaussians_amd.cuda import _wrapper
    if path == "records":
        monkeypatch.setattr(_wrapper, "TEXTURE_GRAD_RECORDS", True)
        monkeypatch.setattr(_wrapper, "PAIR_SAMPLES", False)
        monkeypatch.setattr(_wrapper, "FWD_FILL_RESIDUES", 0x0F)
    elif path == "pair_samples":
        monkeypatch.setattr(_wrapper, "TEXTURE_GRAD_RECORDS", False)
        monkeypatch.setattr(_wrapper, "HEAVY_ROW_PAIRS", 0)
        monkeypatch.setattr(_wrapper, "PAIR_SAMPLES", True)
    elif path == "records_dense_fill":
        monkeypatch.setattr(_wrapper, "TEXTURE_GRAD_RECORDS", True)
        monkeypatch.setattr(_wrapper, "FUSED_FILL", False)
    elif path == "atomics":
        monkeypatch.setattr(_wrapper, "TEXTURE_GRAD_RECORDS", False)
        monkeypatch.setattr(_wrapper, "HEAVY_ROW_PAIRS", 0)
    elif path == "hybrid":
        monkeypatch.setattr(_wrapper, "TEXTURE_GRAD_RECORDS", False)
        monkeypatch.setattr(_wrapper, "HEAVY_ROW_PAIRS", 16)
    elif path == "dense_fill":
        monkeypatch.setattr(_wrapper, "TEXTURE_GRAD_RECORDS", False)
        monkeypatch.setattr(_wrapper, "FUSED_FILL", False)
    st = _stages(dtype)
    assert st["flatten_ids"].numel() > 50_000
    s = st["scene"]
    tiles = all_tiles(st)
    stats = raster_tile_parity(st, st["colors"], st["opac"], s.textures, s.backgrounds, dtype, tiles, uv=uv)
    print("config1", dtype, uv, stats)
    assert stats["tiles"] == 256
