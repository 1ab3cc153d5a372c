"""Pin of the selective-Adam oracle: one step of the reference rule (no bias
correction) equals torch.optim.Adam with lr' = lr (1-b1)/sqrt(1-b2) and
eps' = eps/sqrt(1-b2) on visible rows; invisible rows are untouched."""
import torch

from oracle.adam import selective_adam_update


def test_first_step_equals_rescaled_torch_adam():
    g = torch.Generator().manual_seed(0)
    p0 = torch.randn(64, 5, generator=g, dtype=torch.float64)
    grad = torch.randn(64, 5, generator=g, dtype=torch.float64)
    vis = torch.rand(64, generator=g) > 0.4
    lr, b1, b2, eps = 1e-2, 0.9, 0.99, 1e-8
    p, m, v = selective_adam_update(p0, grad, torch.zeros_like(p0), torch.zeros_like(p0), vis, lr, b1, b2, eps)
    q = torch.nn.Parameter(p0.clone())
    opt = torch.optim.Adam([q], lr=lr * (1 - b1) / (1 - b2) ** 0.5, betas=(b1, b2), eps=eps / (1 - b2) ** 0.5)
    q.grad = grad.clone()
    opt.step()
    assert torch.allclose(p[vis], q.detach()[vis], atol=1e-12)
    assert torch.equal(p[~vis], p0[~vis])
    assert torch.allclose(m[vis], (1 - b1) * grad[vis]) and torch.allclose(v[vis], (1 - b2) * grad[vis] ** 2)
