"""Oracle of the neural-texture decoder MLP (test infrastructure only).

No reference counterpart: the reference's nearest relative is its coordinate
MLP implicit texture (examples/texture_models/mlp.py:6-18). The decoder is
specified by its layers: latent z (32) -> Linear 128 -> ReLU -> Linear 128 ->
ReLU -> Linear 1024 -> sigmoid, read as a 16x16x4 RGBA texture.

Mixed-precision semantics of the GPU kernels, restated in float64:
- Every GEMM operand is rounded to bfloat16 (nearest even) where the kernel
  stages it: z, h1, h2, the weights, and in the backward the pre-activation
  gradients g1, g2, g3.
- Products are exact sums; the kernel accumulates in float32.
- Bias, activations, their derivatives and the bias gradients (column sums)
  are not rounded.

``round_bf16=False`` gives the plain float64 MLP and its exact gradient. The
CPU tests pin that case against torch autograd.
"""
import torch

Tensor = torch.Tensor


def _r(x: Tensor, on: bool) -> Tensor:
    return x.to(torch.bfloat16).to(x.dtype) if on else x


def mlp_forward(z, W1, b1, W2, b2, W3, b3, round_bf16=True):
    """Returns (out [N, D_out] after the sigmoid, h1, h2). W_i are [in, out]."""
    r = lambda t: _r(t, round_bf16)  # noqa: E731
    h1 = torch.relu(r(z) @ r(W1) + b1)
    h2 = torch.relu(r(h1) @ r(W2) + b2)
    out = torch.sigmoid(r(h2) @ r(W3) + b3)
    return out, h1, h2


def mlp_backward(g_out, z, W1, W2, W3, h1, h2, out, round_bf16=True):
    """Gradients (dz, dW1, db1, dW2, db2, dW3, db3) for the upstream gradient g_out."""
    r = lambda t: _r(t, round_bf16)  # noqa: E731
    g3 = g_out * out * (1 - out)
    dW3 = r(h2).T @ r(g3)
    db3 = g3.sum(0)
    g2 = (r(g3) @ r(W3).T) * (h2 > 0)
    dW2 = r(h1).T @ r(g2)
    db2 = g2.sum(0)
    g1 = (r(g2) @ r(W2).T) * (h1 > 0)
    dW1 = r(z).T @ r(g1)
    db1 = g1.sum(0)
    dz = r(g1) @ r(W1).T
    return dz, dW1, db1, dW2, db2, dW3, db3


def gemm_reference(a, b, trans_a=False, trans_b=False, bias=None, epilogue="none", mask=None):
    """float64 product of bfloat16-rounded operands, with the kernel's epilogues."""
    A = _r(a.double(), True)
    B = _r(b.double(), True)
    A = A.T if trans_a else A
    B = B.T if trans_b else B
    c = A @ B
    if bias is not None:
        c = c + bias.double()
    if epilogue == "relu":
        c = torch.relu(c)
    elif epilogue == "sigmoid":
        c = torch.sigmoid(c)
    elif epilogue == "relu_mask":
        c = c * (mask.double() > 0)
    return c
