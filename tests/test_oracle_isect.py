"""Pins of the isect / sort / offset oracle: brute force in the reference's
loop order, stable-sort semantics, closed-form offsets, invariants."""
import numpy as np
import pytest

from oracle import isect as O


def _random_case(seed, C=2, N=60, tw=7, th=5, ts=16):
    rng = np.random.default_rng(seed)
    m2 = rng.uniform(-40, 140, size=(C, N, 2)).astype(np.float32)
    radii = rng.integers(0, 30, size=(C, N)).astype(np.int32)
    radii[rng.random((C, N)) < 0.2] = 0
    depths = rng.uniform(0.5, 5.0, size=(C, N)).astype(np.float32)
    depths[:, ::7] = 2.0  # equal depths exercise the stable tie order
    return m2, radii, depths, ts, tw, th


@pytest.mark.parametrize("seed", range(6))
def test_vectorised_matches_bruteforce(seed):
    m2, r, d, ts, tw, th = _random_case(seed)
    a = O.isect_tiles(m2, r, d, ts, tw, th)
    b = O.isect_tiles_bruteforce(m2, r, d, ts, tw, th)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_known_answer_single_gaussian():
    # centre (24, 40) px, radius 8, 16-px tiles on a 4x4 grid:
    # x: floor(1.5-0.5)=1 .. ceil(2.0)=2 -> [1,2); y: floor(2.5-0.5)=2 .. ceil(3.0)=3 -> [2,3)
    m2 = np.array([[[24.0, 40.0]]], np.float32)
    tpg, ids, flat = O.isect_tiles(m2, np.array([[8]], np.int32), np.array([[1.5]], np.float32), 16, 4, 4)
    assert tpg.tolist() == [[1]]
    tile = 2 * 4 + 1
    dbits = np.array([1.5], np.float32).view(np.int32)[0]
    assert ids.tolist() == [(tile << 32) | int(dbits)]
    assert flat.tolist() == [0]
    offs = O.isect_offset_encode(ids, 1, 4, 4).reshape(-1)
    assert offs.tolist() == [0] * (tile + 1) + [1] * (16 - tile - 1)


def test_clamp_at_image_border():
    # a Gaussian straddling the top-left corner keeps only in-image tiles
    m2 = np.array([[[-5.0, -5.0]]], np.float32)
    tpg, ids, _ = O.isect_tiles(m2, np.array([[20]], np.int32), np.array([[1.0]], np.float32), 16, 3, 3)
    assert tpg.tolist() == [[1]]  # x in [0,1), y in [0,1)


@pytest.mark.parametrize("seed", range(4))
def test_invariants(seed):
    m2, r, d, ts, tw, th = _random_case(seed + 10, C=3, N=200)
    tpg, ids, flat = O.isect_tiles(m2, r, d, ts, tw, th)
    assert tpg.sum() == ids.size == flat.size
    tb, cb = O.bit_length(tw * th), O.bit_length(3)
    k = ids.view(np.uint64) & np.uint64((1 << (32 + tb + cb)) - 1)
    assert np.all(np.diff(k.astype(np.float64)) >= 0) and np.all(k[1:] >= k[:-1])
    # ties keep ascending flatten ids (stable sort of an id-ordered input)
    same = k[1:] == k[:-1]
    assert np.all(flat[1:][same] > flat[:-1][same])
    offs = O.isect_offset_encode(ids, 3, tw, th).reshape(-1)
    assert np.all(np.diff(offs) >= 0) and offs[0] >= 0 and offs[-1] <= ids.size
    # closed form: offset j = number of intersections whose (camera, tile) < j
    hi = ids >> 32
    bins = (hi >> tb) * (tw * th) + (hi & ((1 << tb) - 1))
    for j in range(0, 3 * tw * th, 5):
        assert offs[j] == int((bins < j).sum())


def test_empty():
    tpg, ids, flat = O.isect_tiles(np.zeros((1, 4, 2), np.float32), np.zeros((1, 4), np.int32),
                                   np.ones((1, 4), np.float32), 16, 3, 2)
    assert ids.size == 0 and tpg.sum() == 0
    assert O.isect_offset_encode(ids, 1, 3, 2).sum() == 0


def test_sort_pairs_stable():
    rng = np.random.default_rng(0)
    keys = rng.integers(0, 8, size=1000).astype(np.int64) | (rng.integers(0, 3, 1000).astype(np.int64) << 50)
    vals = np.arange(1000, dtype=np.int32)
    k, v = O.sort_pairs(keys, vals, 48)  # bits >= 48 are ignored
    low = k & ((1 << 48) - 1)
    assert np.all(np.diff(low) >= 0)
    for key in range(8):
        assert np.all(np.diff(v[low == key]) > 0)
