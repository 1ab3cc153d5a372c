"""Selective Adam update: CPU restatement (test infrastructure only).

``textured_gaussians/cuda/csrc/adam.cu:36-45``: for rows ``g`` with
``visibility[g]`` set, every element of the row is updated as
``m = b1 m + (1-b1) g``, ``v = b2 v + (1-b2) g^2``,
``p -= lr m / (sqrt(v) + eps)`` - no bias correction.  Other rows are untouched.
"""
import torch


def selective_adam_update(param, grad, exp_avg, exp_avg_sq, visibility, lr, b1, b2, eps):
    N = visibility.numel()
    M = param.numel() // N
    p = param.reshape(N, M).clone()
    g = grad.reshape(N, M)
    m = exp_avg.reshape(N, M).clone()
    v = exp_avg_sq.reshape(N, M).clone()
    vis = visibility.reshape(N, 1).bool()
    m2 = b1 * m + (1.0 - b1) * g
    v2 = b2 * v + (1.0 - b2) * g * g
    p2 = p - lr * m2 / (torch.sqrt(v2) + eps)
    p = torch.where(vis, p2, p)
    m = torch.where(vis, m2, m)
    v = torch.where(vis, v2, v)
    return p.reshape(param.shape), m.reshape(param.shape), v.reshape(param.shape)
