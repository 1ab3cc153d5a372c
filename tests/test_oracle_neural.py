"""Pins of the decoder-MLP oracle: without bfloat16 rounding it is the plain
float64 MLP, and its hand-written backward equals torch autograd; with rounding,
outputs move by no more than bfloat16 resolution allows."""
import torch

from oracle.neural import mlp_backward, mlp_forward

F64 = torch.float64


def _params(seed=0, n=50):
    g = torch.Generator().manual_seed(seed)
    z = torch.randn(n, 32, generator=g, dtype=F64)
    W1 = torch.randn(32, 128, generator=g, dtype=F64) / 32 ** 0.5
    W2 = torch.randn(128, 128, generator=g, dtype=F64) / 128 ** 0.5
    W3 = torch.randn(128, 1024, generator=g, dtype=F64) / 128 ** 0.5
    b1, b2, b3 = (torch.randn(k, generator=g, dtype=F64) * 0.1 for k in (128, 128, 1024))
    return z, W1, b1, W2, b2, W3, b3


def test_unrounded_backward_equals_autograd():
    z, W1, b1, W2, b2, W3, b3 = _params()
    leaves = [t.clone().requires_grad_(True) for t in (z, W1, b1, W2, b2, W3, b3)]
    out, h1, h2 = mlp_forward(*leaves, round_bf16=False)
    g = torch.randn(out.shape, generator=torch.Generator().manual_seed(3), dtype=F64)
    (out * g).sum().backward()
    mine = mlp_backward(g, z, W1, W2, W3, h1.detach(), h2.detach(), out.detach(), round_bf16=False)
    order = [0, 1, 2, 3, 4, 5, 6]  # dz, dW1, db1, dW2, db2, dW3, db3 vs leaves z, W1, b1, W2, b2, W3, b3
    for i, m in zip(order, mine):
        assert torch.allclose(m, leaves[i].grad, rtol=1e-12, atol=1e-12)


def test_rounded_forward_is_close_to_exact():
    p = _params(seed=1)
    a, _, _ = mlp_forward(*p, round_bf16=False)
    b, _, _ = mlp_forward(*p, round_bf16=True)
    d = (a - b).abs().max().item()
    assert 1e-5 < d < 2e-2  # rounding happens, and stays at bfloat16 scale


def test_texture_layout_is_rgba_innermost():
    p = _params(seed=2, n=3)
    out, _, _ = mlp_forward(*p)
    tex = out.reshape(3, 16, 16, 4)
    assert torch.equal(tex[1, 5, 7], out[1, (5 * 16 + 7) * 4:(5 * 16 + 7) * 4 + 4])
    assert (out > 0).all() and (out < 1).all()
