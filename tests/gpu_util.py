"""Shared helpers for the GPU parity tests: run the GPU stages and hand their
outputs to the CPU oracle as float64 CPU tensors."""
import math

import numpy as np
import torch

from textured_gaussians_amd.cuda import _ops

DEV = torch.device("cuda:0")


def cpu64(t):
    return None if t is None else t.detach().cpu().double()


def gpu_stages(sc, dtype=torch.float32, tile_size=16):
    """Projection + tile lists on the GPU for scene ``sc`` (CPU tensors)."""
    s = sc.to(DEV, dtype)
    radii, means2d, depths, rt, normals = _ops.fully_fused_projection_fwd_2dgs(
        s.means, s.quats, s.scales, s.viewmats, s.Ks, s.width, s.height, 0.3, 0.01, 1e10, 0.0)
    C = s.viewmats.shape[0]
    tw = math.ceil(s.width / tile_size)
    th = math.ceil(s.height / tile_size)
    tpg, ids, flat = _ops.isect_tiles(means2d, radii, depths, None, None, C, tile_size, tw, th, True, True)
    offs = _ops.isect_offset_encode(ids, C, tw, th)
    return dict(scene=s, radii=radii, means2d=means2d, depths=depths, rt=rt, normals=normals, tpg=tpg,
                isect_ids=ids, flatten_ids=flat, offsets=offs, tw=tw, th=th, C=C, tile_size=tile_size)


def rel_err(a, b):
    """max |a-b| / max |b| (0 if both vanish)."""
    a = a.detach().cpu().double()
    b = b.detach().cpu().double()
    den = b.abs().max().item()
    num = (a - b).abs().max().item()
    return num / den if den > 0 else num
