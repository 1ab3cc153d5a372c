"""Tile intersection, radix sort and offsets on the GPU vs the oracle: bit-exact."""
import math

import numpy as np
import pytest
import torch

from oracle import isect as O
from textured_gaussians_amd import scenes
from textured_gaussians_amd.cuda import _ops

from gpu_util import DEV, gpu_stages

pytestmark = pytest.mark.gpu


def _check(st, offsets=True):
    m2 = st["means2d"].float().cpu().numpy()
    r = st["radii"].cpu().numpy()
    d = st["depths"].float().cpu().numpy()
    tpg, ids, flat = O.isect_tiles(m2, r, d, st["tile_size"], st["tw"], st["th"])
    assert np.array_equal(st["tpg"].cpu().numpy(), tpg)
    assert np.array_equal(st["isect_ids"].cpu().numpy(), ids)
    assert np.array_equal(st["flatten_ids"].cpu().numpy(), flat)
    if offsets:
        offs = O.isect_offset_encode(ids, st["C"], st["tw"], st["th"])
        assert np.array_equal(st["offsets"].cpu().numpy(), offs)
    return ids.size


@pytest.mark.parametrize("n,w,h", [(3000, 96, 80), (10_000, 256, 256), (20_000, 250, 130)])
def test_isect_config1_like(n, w, h):
    sc = scenes.config1(n=n, width=w, height=h, tex_res=None)
    assert _check(gpu_stages(sc)) > 0


def test_isect_multicamera():
    cams = [(4 * math.cos(a), 4 * math.sin(a), 0.5) for a in (0.0, 2.0, 4.0)]
    sc = scenes.make_scene(5000, 120, 90, 50.0, cams, None, sh_degree=0)
    assert _check(gpu_stages(sc)) > 0


def test_isect_large_sort():
    # ~1.3M intersections: many radix-sort blocks and a multi-level scan
    sc = scenes.make_scene(300_000, 640, 480, 60.0, [(5.0, 0.0, 0.0)], None, mean_dist="normal",
                           mean_scale=2.0, logscale_mu=-4.0, logscale_sigma=0.6, sh_degree=0)
    assert _check(gpu_stages(sc)) > 500_000


def test_isect_empty():
    # every Gaussian behind the camera: no intersections, offsets all zero
    sc = scenes.config1(n=100, width=64, height=64, tex_res=None, cam_positions=[(-4.0, 0.0, 0.0)])
    sc.means[:, 0] = -10.0
    st = gpu_stages(sc)
    assert st["isect_ids"].numel() == 0
    assert int(st["offsets"].abs().sum()) == 0
    _check(st)


def test_isect_tile8_and_ragged_edges():
    sc = scenes.config1(n=4000, width=77, height=53, tex_res=None)
    _check(gpu_stages(sc, tile_size=8))


def test_sort_pairs_random_keys():
    # stable sort of random 48-bit keys with many duplicates, values = positions
    g = torch.Generator().manual_seed(3)
    n = 200_003
    keys = torch.randint(0, 1 << 20, (n,), generator=g, dtype=torch.int64) << 28
    keys |= torch.randint(0, 4, (n,), generator=g, dtype=torch.int64)
    vals = torch.arange(n, dtype=torch.int32)
    kd, vd = keys.to(DEV), vals.to(DEV)
    import ctypes
    from textured_gaussians_amd.cuda._backend import lib, check
    L = lib()
    ka, va = torch.empty_like(kd), torch.empty_like(vd)
    tb = L.tgs_sort_pairs_temp_bytes(n)
    temp = torch.empty(tb, dtype=torch.uint8, device=DEV)
    sel = ctypes.c_int32(0)
    check(L.tgs_sort_pairs(n, 48, kd.data_ptr(), vd.data_ptr(), ka.data_ptr(), va.data_ptr(), temp.data_ptr(), tb,
                           ctypes.byref(sel), torch.cuda.current_stream().cuda_stream), "sort")
    ko, vo = (ka, va) if sel.value else (kd, vd)
    ek, ev = O.sort_pairs(keys.numpy(), vals.numpy(), 48)
    assert np.array_equal(ko.cpu().numpy(), ek)
    assert np.array_equal(vo.cpu().numpy(), ev)


def test_isect_ties_negative_depths_culled():
    """Synthetic projections with many equal depths (ties go in entry order), negative
    and negative-zero depths (their sign-extended bits fill the camera and tile bits of
    the key, so those entries sort last, by their unsigned depth bits), culled entries
    and two cameras: the oracle's tile lists bit for bit.  (The per-tile
    offsets are not compared here: the reference's offset encoding indexes its output with
    the camera and tile bits of each key, which such keys fill with ones - out of range,
    so it has no defined result to match.  The projection culls depths below the near
    plane, so the product never makes these keys.)"""
    g = torch.Generator().manual_seed(5)
    C, N, ts, tw, th = 2, 6000, 16, 19, 13
    means2d = torch.rand(C, N, 2, generator=g) * torch.tensor([tw * ts * 1.1, th * ts * 1.1]) - 8.0
    radii = torch.randint(0, 40, (C, N), generator=g, dtype=torch.int32)
    depths = torch.randint(-3, 7, (C, N), generator=g).float() * 0.5
    depths[0, :50] = -0.0
    st = dict(means2d=means2d.to(DEV), radii=radii.to(DEV), depths=depths.to(DEV), tile_size=ts, tw=tw, th=th, C=C)
    tpg, ids, flat = _ops.isect_tiles(st["means2d"], st["radii"], st["depths"], None, None, C, ts, tw, th, True, True)
    st.update(tpg=tpg, isect_ids=ids, flatten_ids=flat)
    assert (depths[radii > 0] < 0).any() and (depths[radii > 0] == 0.5).sum() > 100
    assert _check(st, offsets=False) > 10_000
