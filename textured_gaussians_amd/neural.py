"""Neural-texture form of the textured Gaussians: a per-Gaussian latent decoded
by a small MLP into a 16x16 RGBA texture, with the explicit texture kept as the
fallback for Gaussians whose decode error is above a threshold.

No reference counterpart. The reference's nearest relative is the coordinate-MLP
implicit texture (examples/texture_models/mlp.py:6-18), evaluated per pixel
sample. Here the network instead decodes a whole texture per Gaussian, once per
frame, and the explicit-texture rasterizer is used unchanged.

Decoder: z (32) -> Linear 128 -> ReLU -> Linear 128 -> ReLU -> Linear 1024 ->
sigmoid, read as [16, 16, 4] (RGBA innermost, the rasterizer's texture layout).

Every product runs on the matrix cores, with bfloat16 operands and float32
accumulation:
- training (autograd) through ``tgs_gemm_bf16``: bias and activation fused into
  the GEMM epilogue, the backward's ReLU masks fused into the data-gradient
  GEMMs, the weight gradients as split-K GEMMs with atomic accumulation;
- inference (no autograd) through ``tgs_decode_textures``: the whole network in
  one kernel, activations kept in registers, textures written once.

The fallback decision, the memory accounting and the training loops are below.
Training uses the fused Adam kernel (``selective_adam_update``).
"""
import math
from typing import Dict, Optional, Tuple

import torch
from torch import Tensor

from .cuda import _ops

LATENT_DIM, HIDDEN, TEX_RES = 32, 128, 16
TEX_OUT = TEX_RES * TEX_RES * 4


def _k_splits(rows: int) -> int:
    # weight gradients reduce over the batch: one workgroup slice per 4096 rows
    return max(1, math.ceil(rows / 4096))


def _fused_sizes(z, W1, W2, W3) -> bool:
    return (z.shape[1], W1.shape[1], W2.shape[1], W3.shape[1]) == _ops.DECODER_SIZES and z.shape[0] > 0


class _DecodeFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, z, W1, b1, W2, b2, W3, b3):
        E = _ops
        if _fused_sizes(z, W1, W2, W3):
            # the fused kernel writes only the output; the backward recomputes the two
            # hidden layers (K = 32 and 128: a small part of the work) instead of keeping them
            packed = E.decoder_pack(W1, W2, W3)
            out = E.decode_textures(z, packed, b1, b2, b3).reshape(z.shape[0], -1)
            ctx.save_for_backward(z, W1, b1, W2, b2, W3, out)
            ctx.recompute = True
            return out
        h1 = E.gemm_bf16(z, W1, epilogue=E.EPI_RELU, bias=b1)
        h2 = E.gemm_bf16(h1, W2, epilogue=E.EPI_RELU, bias=b2)
        out = E.gemm_bf16(h2, W3, epilogue=E.EPI_SIGMOID, bias=b3)
        ctx.save_for_backward(z, W1, h1, W2, h2, W3, out)
        ctx.recompute = False
        return out

    @staticmethod
    def backward(ctx, g_out):
        E = _ops
        if ctx.recompute:
            z, W1, b1, W2, b2, W3, out = ctx.saved_tensors
        else:
            z, W1, h1, W2, h2, W3, out = ctx.saved_tensors
        # the weights as bfloat16, the values the GEMMs would stage: half the bytes per K step
        W1b, W2b, W3b = W1.bfloat16(), W2.bfloat16(), W3.bfloat16()
        if ctx.recompute:
            h1 = E.gemm_bf16(z, W1b, epilogue=E.EPI_RELU, bias=b1)
            h2 = E.gemm_bf16(h1, W2b, epilogue=E.EPI_RELU, bias=b2)
        ks = _k_splits(z.shape[0])
        # the output-layer cotangent is stored as bfloat16: both GEMMs that read it
        # stage it as bfloat16 anyway, so their results are unchanged and it moves half the bytes
        g3, db3 = E.sigmoid_bwd_colsum(g_out.contiguous(), out, g_dtype=torch.bfloat16)
        dW3 = E.gemm_bf16(h2, g3, trans_a=True, epilogue=E.EPI_ATOMIC, out=torch.zeros_like(W3), k_splits=ks)
        g2 = E.gemm_bf16(g3, W3b, trans_b=True, epilogue=E.EPI_RELU_MASK, mask=h2)
        dW2 = E.gemm_bf16(h1, g2, trans_a=True, epilogue=E.EPI_ATOMIC, out=torch.zeros_like(W2), k_splits=ks)
        db2 = E.colsum_(g2, torch.zeros(W2.shape[1], device=g2.device))
        g1 = E.gemm_bf16(g2, W2b, trans_b=True, epilogue=E.EPI_RELU_MASK, mask=h1)
        dW1 = E.gemm_bf16(z, g1, trans_a=True, epilogue=E.EPI_ATOMIC, out=torch.zeros_like(W1), k_splits=ks)
        db1 = E.colsum_(g1, torch.zeros(W1.shape[1], device=g1.device))
        dz = E.gemm_bf16(g1, W1b, trans_b=True) if ctx.needs_input_grad[0] else None
        return dz, dW1, db1, dW2, db2, dW3, db3


def decode_mlp(z: Tensor, W1: Tensor, b1: Tensor, W2: Tensor, b2: Tensor, W3: Tensor, b3: Tensor) -> Tensor:
    """Raw decoder output [N, D_out] (sigmoid), differentiable in every input."""
    return _DecodeFn.apply(z.contiguous(), W1.contiguous(), b1.contiguous(), W2.contiguous(), b2.contiguous(),
                           W3.contiguous(), b3.contiguous())


class NeuralTextureDecoder(torch.nn.Module):
    """Shared MLP latent -> 16x16 RGBA texture (151.7 k parameters)."""

    def __init__(self, device=None, seed: int = 0, latent_dim: int = LATENT_DIM, hidden: int = HIDDEN,
                 tex_res: int = TEX_RES):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        out_dim = tex_res * tex_res * 4
        self.tex_res = tex_res

        def lin(i, o):  # uniform(-1/sqrt(i), 1/sqrt(i)), as torch.nn.Linear's default
            bound = 1.0 / math.sqrt(i)
            W = (torch.rand(i, o, generator=g) * 2 - 1) * bound
            b = (torch.rand(o, generator=g) * 2 - 1) * bound
            return torch.nn.Parameter(W.to(device)), torch.nn.Parameter(b.to(device))

        self.W1, self.b1 = lin(latent_dim, hidden)
        self.W2, self.b2 = lin(hidden, hidden)
        self.W3, self.b3 = lin(hidden, out_dim)

    def forward(self, z: Tensor) -> Tensor:
        if not torch.is_grad_enabled():
            return self.decode(z)
        out = decode_mlp(z, self.W1, self.b1, self.W2, self.b2, self.W3, self.b3)
        return out.reshape(z.shape[0], self.tex_res, self.tex_res, 4)

    @torch.no_grad()
    def decode(self, z: Tensor, out: Optional[Tensor] = None) -> Tensor:
        """Inference with the fused kernel (all layers in one launch); the weights are
        packed into matrix-core fragments on every call, so it always sees the current
        parameters."""
        if self.tex_res != TEX_RES or self.W1.shape[0] != LATENT_DIM or self.W1.shape[1] != HIDDEN:
            raise RuntimeError("the fused decode kernel is built for 32 -> 128 -> 128 -> 16x16x4")
        packed = _ops.decoder_pack(self.W1.detach(), self.W2.detach(), self.W3.detach())
        return _ops.decode_textures(z.detach().float().contiguous(), packed, self.b1.detach(), self.b2.detach(),
                                    self.b3.detach(), out=out)

    def param_bytes(self, bytes_per_param: int = 2) -> int:
        return sum(p.numel() for p in self.parameters()) * bytes_per_param


class _FusedAdam:
    """Adam over a list of tensors with the fused kernel, every row updated.  The
    kernel has no bias correction (like the reference's selective Adam); it is
    applied here through the step size, lr * sqrt(1 - b2^t) / (1 - b1^t)."""

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-15, bias_correction=True):
        self.params = list(params)
        self.lr, self.betas, self.eps = lr, betas, eps
        self.bias_correction, self.t = bias_correction, 0
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.vis = [torch.ones(p.shape[0] if p.dim() > 1 else 1, dtype=torch.bool, device=p.device)
                    for p in self.params]

    @torch.no_grad()
    def step(self):
        self.t += 1
        b1, b2 = self.betas
        lr = self.lr * (math.sqrt(1 - b2 ** self.t) / (1 - b1 ** self.t) if self.bias_correction else 1.0)
        for p, m, v, vis in zip(self.params, self.m, self.v, self.vis):
            if p.grad is None:
                continue
            N = vis.numel()
            _ops.selective_adam_update(p, p.grad.contiguous(), m, v, vis, lr, b1, b2, self.eps, N, p.numel() // N)

    def zero_grad(self):
        for p in self.params:
            p.grad = None


def fit_decoder(target: Tensor, steps: int = 3000, lr: float = 3e-3, seed: int = 0,
                decoder: Optional[NeuralTextureDecoder] = None, latents: Optional[Tensor] = None,
                weights: Optional[Tensor] = None, lr_decay: float = 0.1) -> Tuple[NeuralTextureDecoder, Tensor, list]:
    """Fits latents (one per texture, N(0,1) at start) and the shared MLP to the
    textures ``target`` [N, 16, 16, 4] by (optionally texel-weighted) mean squared
    error, full batch, with the fused Adam kernel; the learning rate decays
    geometrically to ``lr * lr_decay``.  Returns (decoder, latents, loss history)."""
    dev = target.device
    N = target.shape[0]
    T = target.reshape(N, -1).float()
    if decoder is None:
        decoder = NeuralTextureDecoder(device=dev, seed=seed)
    if latents is None:
        g = torch.Generator().manual_seed(seed + 1)
        latents = torch.randn(N, LATENT_DIM, generator=g).to(dev)
    latents = latents.detach().clone().requires_grad_(True)
    Wt = None if weights is None else weights.reshape(N, -1).float()
    norm = (Wt.sum() if Wt is not None else torch.tensor(float(T.numel()), device=dev)).clamp_min(1e-12)
    opt = _FusedAdam([latents] + list(decoder.parameters()), lr=lr)
    hist = []
    for it in range(steps):
        opt.lr = lr * (lr_decay ** (it / max(1, steps - 1)))
        out = decode_mlp(latents, decoder.W1, decoder.b1, decoder.W2, decoder.b2, decoder.W3, decoder.b3)
        d = out - T
        loss = ((d * d * Wt).sum() if Wt is not None else (d * d).sum()) / norm
        opt.zero_grad()
        loss.backward()
        opt.step()
        if it % 100 == 0 or it == steps - 1:
            hist.append((it, float(loss.detach())))
    return decoder, latents.detach(), hist


def texture_errors(decoder: NeuralTextureDecoder, latents: Tensor, explicit: Tensor,
                   weights: Optional[Tensor] = None) -> Tensor:
    """Per-Gaussian decode error [N]: RMS of (decoded - explicit) over the texels,
    optionally texel-weighted."""
    with torch.no_grad():
        d2 = (decoder(latents) - explicit).reshape(explicit.shape[0], -1) ** 2
        if weights is None:
            return d2.mean(1).sqrt()
        w = weights.reshape(explicit.shape[0], -1)
        return ((d2 * w).sum(1) / w.sum(1).clamp_min(1e-12)).sqrt()


def mixed_textures(decoder: NeuralTextureDecoder, latents: Tensor, explicit: Tensor, fallback: Tensor) -> Tensor:
    """Decoded textures, with the explicit texture for the Gaussians in ``fallback``."""
    return torch.where(fallback[:, None, None, None], explicit, decoder(latents))


def fit_decoder_render(decoder: NeuralTextureDecoder, latents: Tensor, render, ref_images: Tensor,
                       steps: int = 1000, lr: float = 3e-3, lr_decay: float = 0.1, explicit: Optional[Tensor] = None,
                       fallback: Optional[Tensor] = None) -> Tuple[Tensor, list]:
    """Fine-tunes latents and decoder on the render: mean squared error between
    ``render(textures)`` and ``ref_images`` (the explicit-texture renders), where
    the textures are decoded except for the Gaussians in ``fallback``, which keep
    their explicit texture.  Gradients run through the rasterizer's backward into
    the texture gradient and from there through the decoder's backward.  Returns
    (latents, loss history)."""
    z = latents.detach().clone().requires_grad_(True)
    opt = _FusedAdam([z] + list(decoder.parameters()), lr=lr)
    hist = []
    for it in range(steps):
        opt.lr = lr * (lr_decay ** (it / max(1, steps - 1)))
        tex = decoder(z) if fallback is None else mixed_textures(decoder, z, explicit, fallback)
        img = render(tex)
        loss = ((img - ref_images) ** 2).mean()
        opt.zero_grad()
        loss.backward()
        opt.step()
        if it % 100 == 0 or it == steps - 1:
            hist.append((it, float(loss.detach())))
    return z.detach(), hist


def neural_memory(decoder: NeuralTextureDecoder, latents: Tensor, explicit: Tensor, fallback: Tensor
                  ) -> Dict[str, float]:
    """Memory of the explicit form (float32 textures) and of the neural form
    (bfloat16 latents and MLP, plus the explicit textures of the fallback
    Gaussians), and their ratio."""
    N = explicit.shape[0]
    per_tex = explicit[0].numel() * 4
    explicit_bytes = per_tex * N
    neural_bytes = latents.shape[1] * 2 * N + decoder.param_bytes(2) + int(fallback.sum()) * per_tex
    return dict(fallback_frac=float(fallback.float().mean()), explicit_bytes=float(explicit_bytes),
                neural_bytes=float(neural_bytes), memory_ratio=float(explicit_bytes / neural_bytes))


@torch.no_grad()
def decode_with_fallback(decoder: NeuralTextureDecoder, latents: Tensor, explicit: Tensor, threshold: float,
                         weights: Optional[Tensor] = None) -> Tuple[Tensor, Tensor, Dict[str, float]]:
    """Decoded textures, with the explicit texture kept for every Gaussian whose
    decode error (``texture_errors``) exceeds ``threshold``.  Returns (textures
    [N,16,16,4], fallback mask [N], memory stats of ``neural_memory``)."""
    err = texture_errors(decoder, latents, explicit, weights)
    fb = err > threshold
    stats = neural_memory(decoder, latents, explicit, fb)
    stats.update(decode_rms_median=float(err.median()), decode_rms_max=float(err.max()))
    return mixed_textures(decoder, latents, explicit, fb), fb, stats


def refine_textures(render, textures: Tensor, targets: Tensor, steps: int = 300, lr: float = 0.01,
                    value_range: Tuple[float, float] = (0.0, 1.0)) -> Tuple[Tensor, Tensor]:
    """The plain method's refinement stage: only the explicit textures are
    optimised (fused Adam), against ``targets`` [C, H, W, 3] by mean L2, and
    projected back into ``value_range`` after every step (a few views leave
    weakly seen texels free to drift otherwise).  ``render(textures) ->
    [C, H, W, 3]`` is the differentiable textured render.  Returns (refined
    textures, per-texel observation weight: the gradient magnitude accumulated
    over the steps)."""
    tex = textures.detach().clone().requires_grad_(True)
    opt = _FusedAdam([tex], lr=lr)
    seen = torch.zeros(textures.shape[:-1], device=textures.device)
    for _ in range(steps):
        img = render(tex)
        loss = ((img - targets) ** 2).mean()
        opt.zero_grad()
        loss.backward()
        seen += tex.grad.abs().sum(-1)
        opt.step()
        with torch.no_grad():
            tex.clamp_(*value_range)
    return tex.detach(), seen


def observation_weights(seen: Tensor, floor: float = 0.05) -> Tensor:
    """Per-texel fitting weights [N, r, r, 4] from the refinement's observation
    weight: proportional to it up to its median, at least ``floor``."""
    med = seen[seen > 0].median() if (seen > 0).any() else seen.new_tensor(1.0)
    w = (seen / med.clamp_min(1e-30)).clamp(floor, 1.0)
    return w[..., None].expand(*seen.shape, 4).contiguous()


def psnr(a: Tensor, b: Tensor, peak: float = 1.0) -> float:
    mse = float(((a.double() - b.double()) ** 2).mean())
    return float("inf") if mse == 0 else 10 * math.log10(peak * peak / mse)
