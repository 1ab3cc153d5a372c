"""Selective Adam for large bfloat16 textures, stepped inside the rasterizer's
backward so the texture gradient never exists as a dense tensor.

The reference trains textures with a float32 gradient and float32 Adam state
(``selective_adam_update``, ``cuda/csrc/adam.cu:16-47``; the trainer's
``torch.optim.Adam``).  At 2M Gaussians x 64x64 RGBA that is 131 GB per tensor:
parameter, gradient and two moments (524 GB) do not fit in one GPU's HBM.  Here
the state is four 16-bit planes, 262 GB at that size:

- ``textures`` (bfloat16): the render copy, the float32 master rounded to nearest;
- ``lo`` (int16): master bits minus (render bits << 16), so render + lo hold the
  float32 master exactly;
- ``exp_avg`` / ``exp_avg_sq`` (bfloat16): the moments, stochastically rounded
  (round-to-nearest would freeze the second moment, whose per-step change
  (1 - b2) is below half a bfloat16 ulp).

The backward writes one record per (pixel, Gaussian) texture-gradient pair, and
``tgs_texture_adam_rows`` sums each visible row's records in LDS and applies the
update to that row in the same kernel.  The arithmetic is the reference's
selective Adam: no bias correction, rows not visible in the step left untouched
(visible = the Gaussian is in this step's tile lists).
"""
from typing import Optional, Tuple

import torch
from torch import Tensor

from .cuda import _ops


def split_master(master: Tensor) -> Tuple[Tensor, Tensor]:
    """float32 -> (bfloat16 rounded to nearest with ties away from zero, int16 residual
    bits) with ``join_master`` inverting it exactly."""
    assert master.dtype == torch.float32
    bits = master.contiguous().view(torch.int32)
    hi = (bits + 0x8000) >> 16  # arithmetic shift: the upper half of the rounded pattern
    lo = bits - (hi << 16)      # in [-32768, 32767]
    return hi.to(torch.int16).view(torch.bfloat16), lo.to(torch.int16)


def join_master(hi: Tensor, lo: Tensor) -> Tensor:
    """The float32 master held by (bfloat16 render copy, int16 residual)."""
    bits = (hi.view(torch.int16).to(torch.int32) << 16) + lo.to(torch.int32)
    return bits.view(torch.float32)


class BF16TextureAdam:
    """Bfloat16 textures [N, h, w, 4] with a float32 master and Adam state, updated
    by the textured rasterizer's backward.  Pass ``opt.textures`` as the textures
    and ``texture_adam=opt`` to ``rasterization_textured_gaussians``; each backward
    then applies one Adam step to the visible rows (lr, betas and eps as for
    ``torch.optim.Adam``, without bias correction, as the reference's selective Adam)."""

    def __init__(self, n: int, height: int, width: int, lr: float, betas=(0.9, 0.999), eps: float = 1e-15,
                 device=None, seed: int = 0):
        if (height * width) % 2 or 2 * height * width * 4 * 4 > 160 * 1024:
            raise ValueError("texture_adam: needs an even texel count and height*width <= 5120")
        shape = (n, height, width, 4)
        self.textures = torch.zeros(shape, dtype=torch.bfloat16, device=device)
        self.lo = torch.zeros(shape, dtype=torch.int16, device=device)
        self.exp_avg = torch.zeros(shape, dtype=torch.bfloat16, device=device)
        self.exp_avg_sq = torch.zeros(shape, dtype=torch.bfloat16, device=device)
        self.lr, self.betas, self.eps = lr, tuple(betas), eps
        self.seed = int(seed)
        self.steps = 0

    @classmethod
    def from_textures(cls, master: Tensor, lr: float, chunk_rows: int = 65536, **kw) -> "BF16TextureAdam":
        """Initialised from float32 textures (any device), ``chunk_rows`` rows at a time."""
        n, h, w, _ = master.shape
        kw.setdefault("device", master.device)
        opt = cls(n, h, w, lr, **kw)
        opt.load_master(master, chunk_rows=chunk_rows)
        return opt

    @torch.no_grad()
    def load_master(self, master: Tensor, start: int = 0, chunk_rows: int = 65536) -> None:
        """Sets rows [start, start + len(master)) of the master (and render copy)."""
        for a in range(0, master.shape[0], chunk_rows):
            blk = master[a:a + chunk_rows].to(self.textures.device, torch.float32)
            hi, lo = split_master(blk)
            self.textures[start + a:start + a + blk.shape[0]] = hi
            self.lo[start + a:start + a + blk.shape[0]] = lo

    @torch.no_grad()
    def master(self, rows: Optional[slice] = None) -> Tensor:
        """float32 master (of ``rows``, default all: 4 bytes per texel channel)."""
        rows = rows if rows is not None else slice(None)
        return join_master(self.textures[rows], self.lo[rows])

    def state_dict(self):
        return {"textures": self.textures, "lo": self.lo, "exp_avg": self.exp_avg, "exp_avg_sq": self.exp_avg_sq,
                "lr": self.lr, "betas": self.betas, "eps": self.eps, "seed": self.seed, "steps": self.steps}

    @torch.no_grad()
    def step_from_records(self, records, visibility: Tensor) -> None:
        """One Adam step from texture-gradient records (``rasterize_to_pixels_bwd_*``
        with ``v_textures_out=False``)."""
        seed = (self.seed * 0x9E3779B1 + self.steps * 0x85EBCA77 + 1) & 0xFFFFFFFF
        _ops.texture_adam_rows(records, visibility, self.textures.view(torch.int16), self.lo,
                               self.exp_avg.view(torch.int16), self.exp_avg_sq.view(torch.int16), self.lr,
                               self.betas[0], self.betas[1], self.eps, seed)
        self.steps += 1
