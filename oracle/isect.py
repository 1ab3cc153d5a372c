"""Tile intersection, stable key sort and offset encode: CPU restatement.

Test infrastructure only.  Follows ``textured_gaussians/cuda/csrc/isect_tiles.cu``:
tile bounds ``:60-77`` (float32 arithmetic, the uint32 cast saturating
negatives to 0, i.e. ``clamp(floor(x), 0, tile_width)``), key layout
``:90-101``, bit budget ``:156-160``, the stable radix sort over the low
``32 + tile_n_bits + cam_n_bits`` bits ``:262-299`` (CUB ``DeviceRadixSort``
is an LSD sort, hence stable) and the offset encode ``:309-355``.
All integer results are bit-defined; inputs are numpy arrays.
"""
from typing import Tuple

import numpy as np


def bit_length(x: int) -> int:
    """``floor(log2(x)) + 1`` for x >= 1, as ``isect_tiles.cu:156-157``."""
    return int(x).bit_length()


def tile_bounds(means2d: np.ndarray, radii: np.ndarray, tile_size: int,
                tile_width: int, tile_height: int):
    """Per-(c,n) [min, max) tile rectangle, ``isect_tiles.cu:60-70``.

    means2d float32 [..., 2]; radii int32 [...].  Returns int64 arrays
    (x0, y0, x1, y1)."""
    m = means2d.astype(np.float32)
    r = radii.astype(np.float32) / np.float32(tile_size)
    tx = m[..., 0] / np.float32(tile_size)
    ty = m[..., 1] / np.float32(tile_size)
    with np.errstate(invalid="ignore", over="ignore"):
        x0 = np.floor((tx - r).astype(np.float32))
        y0 = np.floor((ty - r).astype(np.float32))
        x1 = np.ceil((tx + r).astype(np.float32))
        y1 = np.ceil((ty + r).astype(np.float32))

    def sat(v, hi):
        v = np.where(np.isnan(v), 0.0, v)
        return np.clip(v, 0, hi).astype(np.int64)

    return sat(x0, tile_width), sat(y0, tile_height), sat(x1, tile_width), sat(y1, tile_height)


def depth_bits(depths: np.ndarray) -> np.ndarray:
    """int32 bit pattern of float32 depths, sign-extended to int64 (``:92``)."""
    return depths.astype(np.float32).view(np.int32).astype(np.int64)


def isect_tiles(means2d: np.ndarray, radii: np.ndarray, depths: np.ndarray,
                tile_size: int, tile_width: int, tile_height: int, sort: bool = True
                ) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Returns (tiles_per_gauss int32 [C,N], isect_ids int64 [I], flatten_ids int32 [I])."""
    C, N = radii.shape
    n_tiles = tile_width * tile_height
    tb = bit_length(n_tiles)
    cb = bit_length(C)
    assert tb + cb <= 32
    x0, y0, x1, y1 = tile_bounds(means2d, radii, tile_size, tile_width, tile_height)
    vis = radii > 0
    tpg = np.where(vis, (y1 - y0) * (x1 - x0), 0).astype(np.int64)
    flat_tpg = tpg.reshape(-1)
    I = int(flat_tpg.sum())
    # expand every (c,n) into its tiles, row-major (y outer, x inner) in (c,n) order
    idx = np.repeat(np.arange(C * N, dtype=np.int64), flat_tpg)
    start = np.concatenate([[0], np.cumsum(flat_tpg)[:-1]]) if C * N else np.zeros(0, np.int64)
    local = np.arange(I, dtype=np.int64) - np.repeat(start, flat_tpg)
    wx = (x1 - x0).reshape(-1)[idx]
    wx = np.where(wx == 0, 1, wx)
    ty = y0.reshape(-1)[idx] + local // wx
    tx = x0.reshape(-1)[idx] + local % wx
    tile_id = ty * tile_width + tx
    cid = idx // N if N else idx
    keys = (cid << (32 + tb)) | (tile_id << 32) | depth_bits(depths).reshape(-1)[idx]
    flat = idx.astype(np.int32)
    if sort and I:
        keys, flat = sort_pairs(keys, flat, 32 + tb + cb)
    return tpg.astype(np.int32), keys.astype(np.int64), flat


def sort_pairs(keys: np.ndarray, values: np.ndarray, end_bit: int):
    """Stable sort of (key, value) pairs on the low ``end_bit`` bits of the key."""
    mask = np.uint64((1 << end_bit) - 1)
    k = keys.view(np.uint64) & mask
    order = np.argsort(k, kind="stable")
    return keys[order], values[order]


def isect_offset_encode(isect_ids: np.ndarray, C: int, tile_width: int,
                        tile_height: int) -> np.ndarray:
    """``offsets[c, ty, tx]`` = index of the first intersection of tile
    (c, ty*tw+tx) or of the next occupied tile; the tail gets I; all zeros if
    I == 0 (``isect_tiles.cu:309-355, 386-388``)."""
    n_tiles = tile_width * tile_height
    tb = bit_length(n_tiles)
    I = isect_ids.shape[0]
    if I == 0:
        return np.zeros((C, tile_height, tile_width), np.int32)
    hi = isect_ids.astype(np.int64) >> 32
    ids = (hi >> tb) * n_tiles + (hi & ((1 << tb) - 1))
    j = np.arange(C * n_tiles, dtype=np.int64)
    off = np.searchsorted(ids, j, side="left")
    return off.astype(np.int32).reshape(C, tile_height, tile_width)


def isect_tiles_bruteforce(means2d, radii, depths, tile_size, tile_width, tile_height):
    """Pure-python double loop in the reference's kernel order (tiny inputs
    only), then python's stable ``sorted`` on the masked key."""
    import struct
    C, N = radii.shape
    n_tiles = tile_width * tile_height
    tb = bit_length(n_tiles)
    cb = bit_length(C)
    f32 = np.float32
    keys, vals, tpg = [], [], np.zeros((C, N), np.int32)
    for c in range(C):
        for n in range(N):
            r = int(radii[c, n])
            if r <= 0:
                continue
            mx, my = f32(means2d[c, n, 0]), f32(means2d[c, n, 1])
            tr = f32(r) / f32(tile_size)
            txf, tyf = mx / f32(tile_size), my / f32(tile_size)
            bx0 = min(max(0, int(np.floor(f32(txf - tr)))), tile_width)
            by0 = min(max(0, int(np.floor(f32(tyf - tr)))), tile_height)
            bx1 = min(max(0, int(np.ceil(f32(txf + tr)))), tile_width)
            by1 = min(max(0, int(np.ceil(f32(tyf + tr)))), tile_height)
            tpg[c, n] = (by1 - by0) * (bx1 - bx0)
            d = struct.unpack("i", struct.pack("f", float(depths[c, n])))[0]
            for y in range(by0, by1):
                for x in range(bx0, bx1):
                    keys.append((c << (32 + tb)) | ((y * tile_width + x) << 32) | d)
                    vals.append(c * N + n)
    mask = (1 << (32 + tb + cb)) - 1
    order = sorted(range(len(keys)), key=lambda i: keys[i] & mask)
    return (tpg, np.array([keys[i] for i in order], np.int64),
            np.array([vals[i] for i in order], np.int32))
