"""Pins of the projection oracle: closed forms, the row-major layout of the
ray transform on an asymmetric case, and the reference VJP restatement."""
import math

import numpy as np
import pytest
import torch

from oracle import projection as OP


def _t(x):
    return torch.tensor(x, dtype=torch.float64)



def _cam(fx=300.0, fy=280.0, cx=64.0, cy=48.0):
    K = _t([[[fx, 0, cx], [0, fy, cy], [0, 0, 1.0]]])
    return torch.eye(4, dtype=torch.float64)[None], K


def test_on_axis_closed_form():
    V, K = _cam()
    z, s0, s1 = 5.0, 0.05, 0.02
    means = _t([[0.0, 0.0, z]])
    quats = _t([[1.0, 0, 0, 0]])
    scales = _t([[s0, s1, 1.0]])
    radii, m2, dep, T, nrm = OP.project_2dgs(means, quats, scales, V, K, 128, 96)
    assert torch.allclose(m2[0, 0], _t([64.0, 48.0]))
    assert dep[0, 0].item() == z
    exp_T = _t([[300 * s0, 0, 64 * z], [0, 280 * s1, 48 * z], [0, 0, z]])
    assert torch.allclose(T[0, 0], exp_T)
    assert radii[0, 0].item() == math.ceil(3 * max(300 * s0, 280 * s1) / z)
    # the normal faces the camera (-n . p > 0)
    assert torch.allclose(nrm[0, 0], _t([0.0, 0.0, -1.0]))


def test_rotated_off_axis_layout():
    # asymmetric case: T must equal K @ [R Rq S e0, R Rq S e1, p_c] stored row-major
    g = torch.Generator().manual_seed(1)
    ang = 0.3
    V = torch.eye(4, dtype=torch.float64)
    V[:3, :3] = _t([[math.cos(ang), 0, math.sin(ang)], [0, 1, 0], [-math.sin(ang), 0, math.cos(ang)]])
    V[:3, 3] = _t([0.2, -0.1, 0.5])
    K = _t([[310.0, 0, 70], [0, 290, 50], [0, 0, 1]])
    means = _t([[0.3, -0.2, 4.0]])
    q = torch.randn(1, 4, generator=g, dtype=torch.float64)
    scales = _t([[0.07, 0.03, 9.0]])
    radii, m2, dep, T, nrm = OP.project_2dgs(means, q, scales, V[None], K[None], 256, 256)
    # independent construction with numpy (normalised quaternion -> matrix)
    w, x, y, z = (q[0] / q[0].norm()).tolist()
    Rq = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                   [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                   [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    Rv, t = V[:3, :3].numpy(), V[:3, 3].numpy()
    pc = Rv @ means[0].numpy() + t
    M = Rv @ Rq @ np.diag([0.07, 0.03, 1.0])
    WH = np.stack([M[:, 0], M[:, 1], pc], 1)
    Tn = K.numpy() @ WH
    assert np.allclose(T[0, 0].numpy(), Tn)
    # means2d is the centre of the image of the splat's unit circle (u,v) = (cos a, sin a)
    # and the radius is 3x its largest bounding-box half extent (dense sampling of the circle)
    a = np.linspace(0, 2 * np.pi, 200001)
    P = Tn @ np.stack([np.cos(a), np.sin(a), np.ones_like(a)])
    x, y = P[0] / P[2], P[1] / P[2]
    assert np.allclose(m2[0, 0].numpy(), [(x.max() + x.min()) / 2, (y.max() + y.min()) / 2], atol=1e-6)
    half = max(x.max() - x.min(), y.max() - y.min()) / 2
    assert radii[0, 0].item() == math.ceil(3 * half)
    assert dep[0, 0].item() == pytest.approx(pc[2])


def test_culling():
    V, K = _cam()
    means = _t([[0.0, 0, -1.0], [0.0, 0, 0.001], [50.0, 0, 1.0], [0.0, 0, 3.0]])
    quats = _t([[1.0, 0, 0, 0]]).repeat(4, 1)
    scales = torch.full((4, 3), 0.01, dtype=torch.float64)
    radii, *_ = OP.project_2dgs(means, quats, scales, V, K, 128, 96)
    assert radii[0].tolist() == [0, 0, 0, radii[0, 3].item()] and radii[0, 3] > 0


def _scene(seed=0, N=40, C=2):
    g = torch.Generator().manual_seed(seed)
    means = torch.randn(N, 3, generator=g, dtype=torch.float64) * 0.5
    quats = torch.randn(N, 4, generator=g, dtype=torch.float64)
    scales = torch.exp(torch.randn(N, 3, generator=g, dtype=torch.float64) * 0.3 - 2.5)
    Vs = []
    for c in range(C):
        a = 0.4 * c
        V = torch.eye(4, dtype=torch.float64)
        V[:3, :3] = _t([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
        V[:3, 3] = _t([0.0, 0.0, 4.0])
        Vs.append(V)
    K = _t([[200.0, 0, 64], [0, 200, 48], [0, 0, 1]])
    return means, quats, scales, torch.stack(Vs), K[None].repeat(C, 1, 1)


def _cotangents(C, N, seed=3):
    g = torch.Generator().manual_seed(seed)
    f = torch.float64
    return (torch.randn(C, N, 2, generator=g, dtype=f), torch.randn(C, N, generator=g, dtype=f),
            torch.randn(C, N, 3, generator=g, dtype=f), torch.randn(C, N, 3, 3, generator=g, dtype=f))


def _autograd_vjp(means, quats, scales, V, K, cot):
    ls = [means.clone().requires_grad_(True), quats.clone().requires_grad_(True), scales.clone().requires_grad_(True)]
    radii, m2, dep, T, nrm = OP.project_2dgs(ls[0], ls[1], ls[2], V, K, 128, 96)
    vm2, vd, vn, vT = cot
    ((m2 * vm2).sum() + (dep * vd).sum() + (nrm * vn).sum() + (T * vT).sum()).backward()
    return radii, T, [l.grad for l in ls]


def test_reference_vjp_equals_autograd_on_exact_terms():
    means, quats, scales, V, K = _scene()
    C, N = V.shape[0], means.shape[0]
    vm2, vd, vn, vT = _cotangents(C, N)
    cot = (torch.zeros_like(vm2), vd, vn, vT)
    radii, T, exact = _autograd_vjp(means, quats, scales, V, K, cot)
    assert (radii > 0).sum() > 10
    ref = OP.projection_vjp_reference(means, quats, scales, V, K, radii, T.detach(), *cot)
    for a, b in zip(ref, exact):
        assert torch.allclose(a, b, atol=1e-10, rtol=1e-9)


def test_reference_means2d_term_is_summandwise_derivative():
    # The reference's d(mean2d)/d(T2) differentiates each summand T0_j T2_j / D
    # only w.r.t. its own T2_j.  Build that rule with autograd and compare.
    means, quats, scales, V, K = _scene(seed=2)
    C, N = V.shape[0], means.shape[0]
    vm2, vd, vn, vT = _cotangents(C, N, seed=4)
    zero = (vm2, torch.zeros_like(vd), torch.zeros_like(vn), torch.zeros_like(vT))
    radii, T, _ = _autograd_vjp(means, quats, scales, V, K, zero)
    T = T.detach()
    sg = _t([1.0, 1.0, -1.0])
    # per-entry partial derivatives of mean2d as the reference takes them
    vT_rule = torch.zeros_like(T)
    for j in range(3):
        T2 = T[..., 2, :].clone().requires_grad_(True)
        D = (sg * T2 * T2).sum(-1)
        gx = sg[j] * T[..., 0, j] * T2[..., j] / D
        gy = sg[j] * T[..., 1, j] * T2[..., j] / D
        (dgx,) = torch.autograd.grad(gx.sum(), T2, retain_graph=True)
        (dgy,) = torch.autograd.grad(gy.sum(), T2)
        vT_rule[..., 2, j] = vm2[..., 0] * dgx[..., j] + vm2[..., 1] * dgy[..., j]
        D0 = (sg * T[..., 2, :] ** 2).sum(-1)
        vT_rule[..., 0, j] = vm2[..., 0] * sg[j] * T[..., 2, j] / D0
        vT_rule[..., 1, j] = vm2[..., 1] * sg[j] * T[..., 2, j] / D0
    vT_rule = torch.where((radii > 0)[..., None, None], vT_rule, torch.zeros_like(vT_rule))
    via_T = OP.projection_vjp_reference(means, quats, scales, V, K, radii, T, torch.zeros_like(vm2),
                                        torch.zeros_like(vd), torch.zeros_like(vn), vT_rule)
    ref = OP.projection_vjp_reference(means, quats, scales, V, K, radii, T, *zero)
    for a, b in zip(ref, via_T):
        assert torch.allclose(a, b, atol=1e-10, rtol=1e-9)


def test_autograd_vjp_matches_central_differences():
    means, quats, scales, V, K = _scene(seed=5, N=12, C=1)
    cot = _cotangents(1, 12, seed=6)

    def f(m):
        r, m2, dep, T, nrm = OP.project_2dgs(m, quats, scales, V, K, 128, 96)
        vm2, vd, vn, vT = cot
        return ((m2 * vm2).sum() + (dep * vd).sum() + (nrm * vn).sum() + (T * vT).sum()).item()

    _, _, (gm, _, _) = _autograd_vjp(means, quats, scales, V, K, cot)
    eps = 1e-6
    for i in range(4):
        for k in range(3):
            mp, mm = means.clone(), means.clone()
            mp[i, k] += eps
            mm[i, k] -= eps
            fd = (f(mp) - f(mm)) / (2 * eps)
            assert abs(fd - gm[i, k].item()) < 1e-5 * max(1.0, abs(fd))
