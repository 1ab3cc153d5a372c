"""2DGS projection: CPU restatement (test infrastructure only).

Follows ``textured_gaussians/cuda/csrc/fully_fused_projection_2dgs_fwd.cu:72-238``
(forward) and ``utils.cuh:41-65, 659-791`` (the reference's VJPs).
Every function is dtype-generic torch code; the forward is differentiable, so
``torch.autograd`` on it gives the exact VJP.
"""
from typing import Tuple

import torch
from torch import Tensor


def quat_to_rotmat(quats: Tensor) -> Tensor:
    """Normalised (w, x, y, z) quaternion -> rotation matrix [..., 3, 3].

    Same matrix as ``utils.cuh:15-39`` (which builds it column by column)."""
    q = quats / torch.linalg.norm(quats, dim=-1, keepdim=True)
    w, x, y, z = q.unbind(-1)
    r00 = 1 - 2 * (y * y + z * z)
    r01 = 2 * (x * y - w * z)
    r02 = 2 * (x * z + w * y)
    r10 = 2 * (x * y + w * z)
    r11 = 1 - 2 * (x * x + z * z)
    r12 = 2 * (y * z - w * x)
    r20 = 2 * (x * z - w * y)
    r21 = 2 * (y * z + w * x)
    r22 = 1 - 2 * (x * x + y * y)
    return torch.stack(
        [torch.stack([r00, r01, r02], -1), torch.stack([r10, r11, r12], -1),
         torch.stack([r20, r21, r22], -1)], -2)


def _intrinsics(Ks: Tensor) -> Tensor:
    """The reference reads only fx=K[0,0], cx=K[0,2], fy=K[1,1], cy=K[1,2]
    (``fully_fused_projection_2dgs_fwd.cu:143-146``)."""
    Z = torch.zeros_like(Ks[..., 0, 0])
    O = torch.ones_like(Z)
    return torch.stack(
        [torch.stack([Ks[..., 0, 0], Z, Ks[..., 0, 2]], -1),
         torch.stack([Z, Ks[..., 1, 1], Ks[..., 1, 2]], -1),
         torch.stack([Z, Z, O], -1)], -2)


def project_2dgs(means: Tensor, quats: Tensor, scales: Tensor, viewmats: Tensor,
                 Ks: Tensor, width: int, height: int, near_plane: float = 0.01,
                 far_plane: float = 1e10, radius_clip: float = 0.0
                 ) -> Tuple[Tensor, Tensor, Tensor, Tensor, Tensor]:
    """Forward projection for C cameras x N Gaussians.

    Returns (radii int32 [C,N], means2d [C,N,2], depths [C,N],
    ray_transforms [C,N,3,3] = K @ [R Rq S e0 | R Rq S e1 | p_c] row-major,
    normals [C,N,3]).  Culled entries hold radii 0 and zeros elsewhere (the
    reference leaves them uninitialised, ``fully_fused_projection_2dgs_fwd.cu:271-276``).
    """
    R = viewmats[:, None, :3, :3]                      # [C,1,3,3]
    t = viewmats[:, None, :3, 3]                       # [C,1,3]
    p_c = (R @ means[None, :, :, None])[..., 0] + t    # [C,N,3]
    Rq = quat_to_rotmat(quats)[None]                   # [1,N,3,3]
    S = torch.stack([scales[:, 0], scales[:, 1], torch.ones_like(scales[:, 0])], -1)
    RS = R @ (Rq * S[None, :, None, :])                # columns scaled: [C,N,3,3]
    WH = torch.stack([RS[..., 0], RS[..., 1], p_c], -1)  # columns
    K = _intrinsics(Ks)[:, None]                       # [C,1,3,3]
    Tm = K @ WH                                        # [C,N,3,3], rows T0,T1,T2
    T0, T1, T2 = Tm[..., 0, :], Tm[..., 1, :], Tm[..., 2, :]
    sgn = torch.tensor([1.0, 1.0, -1.0], dtype=means.dtype)
    distance = (sgn * T2 * T2).sum(-1)
    ok_dist = distance != 0
    f = sgn / torch.where(ok_dist, distance, torch.ones_like(distance))[..., None]
    mx = (f * T0 * T2).sum(-1)
    my = (f * T1 * T2).sum(-1)
    hx = mx * mx - (f * T0 * T0).sum(-1)
    hy = my * my - (f * T1 * T1).sum(-1)
    with torch.no_grad():
        radius = torch.ceil(3 * torch.sqrt(torch.clamp(torch.maximum(hx, hy), min=1e-4)))
        depth_ok = (p_c[..., 2] >= near_plane) & (p_c[..., 2] <= far_plane)
        inside = ~((mx + radius <= 0) | (mx - radius >= width) |
                   (my + radius <= 0) | (my - radius >= height))
        valid = depth_ok & ok_dist & (radius > radius_clip) & inside
    n = RS[..., 2]
    mult = torch.where((-n * p_c).sum(-1) > 0, 1.0, -1.0).to(means.dtype)
    normal = n * mult[..., None]
    z = torch.zeros((), dtype=means.dtype)
    radii = torch.where(valid, radius, z).to(torch.int32)
    v2 = valid[..., None]
    means2d = torch.where(v2, torch.stack([mx, my], -1), z)
    depths = torch.where(valid, p_c[..., 2], z)
    ray_transforms = torch.where(valid[..., None, None], Tm, z)
    normals = torch.where(v2, normal, z)
    return radii, means2d, depths, ray_transforms, normals


def projection_vjp_reference(means: Tensor, quats: Tensor, scales: Tensor,
                             viewmats: Tensor, Ks: Tensor, radii: Tensor,
                             ray_transforms: Tensor, v_means2d: Tensor,
                             v_depths: Tensor, v_normals: Tensor,
                             v_ray_transforms: Tensor, norm_rot_grad: bool = False
                             ) -> Tuple[Tensor, Tensor, Tensor]:
    """The reference's projection VJP, restated term by term.

    ``fully_fused_projection_2dgs_bwd.cu:45-147`` + ``utils.cuh:659-724``
    (``norm_rot_grad=False``) or ``utils.cuh:726-791`` (``True``: the scale
    columns are divided by the scale instead of multiplied).  The means2d term
    keeps only the ``T0 * (f - 2 f^2 T2^2)`` part of d(mean2d)/d(T2)
    (``utils.cuh:684-703``); the exact derivative has ``N * 2 f^2 T2`` with the
    full numerator N.  Returns (v_means [N,3], v_quats [N,4], v_scales [N,3]).
    """
    C, N = radii.shape
    dt = means.dtype
    valid = radii > 0
    Tm = ray_transforms
    vT = v_ray_transforms.clone()
    vT[..., 2, 2] = vT[..., 2, 2] + v_depths
    T0, T1, T2 = Tm[..., 0, :], Tm[..., 1, :], Tm[..., 2, :]
    dist = T2[..., 0] ** 2 + T2[..., 1] ** 2 - T2[..., 2] ** 2
    f = 1.0 / torch.where(valid, dist, torch.ones_like(dist))
    vmx, vmy = v_means2d[..., 0], v_means2d[..., 1]
    sg = torch.tensor([1.0, 1.0, -1.0], dtype=dt)
    add = torch.zeros_like(vT)
    add[..., 0, :] = vmx[..., None] * f[..., None] * T2 * sg
    add[..., 1, :] = vmy[..., None] * f[..., None] * T2 * sg
    # third row: T0_j (f +- 2 f^2 T2_j^2), sign + for j<2 -> (f - 2f^2T2_j^2), j=2 -> -(f + 2f^2T2_2^2)
    ff = f[..., None]
    corr = torch.stack([ff[..., 0] - 2 * ff[..., 0] ** 2 * T2[..., 0] ** 2,
                        ff[..., 0] - 2 * ff[..., 0] ** 2 * T2[..., 1] ** 2,
                        -(ff[..., 0] + 2 * ff[..., 0] ** 2 * T2[..., 2] ** 2)], -1)
    add[..., 2, :] = vmx[..., None] * T0 * corr + vmy[..., None] * T1 * corr
    vT = vT + add
    K = _intrinsics(Ks)[:, None]
    v_WH = K.transpose(-1, -2) @ vT                    # [C,N,3,3]
    R = viewmats[:, None, :3, :3]
    Rt = R.transpose(-1, -2)
    v_RS0 = (Rt @ v_WH[..., :, 0:1])[..., 0]
    v_RS1 = (Rt @ v_WH[..., :, 1:2])[..., 0]
    v_mean = (Rt @ v_WH[..., :, 2:3])[..., 0]
    Rq = quat_to_rotmat(quats)[None].expand(C, N, 3, 3)
    t = viewmats[:, None, :3, 3]
    p_c = (R @ means[None, :, :, None])[..., 0] + t
    tn = (R @ Rq[..., :, 2:3])[..., 0]
    mult = torch.where((-tn * p_c).sum(-1) > 0, 1.0, -1.0).to(dt)
    v_tn = (Rt @ v_normals[..., None])[..., 0] * mult[..., None]
    s0 = scales[None, :, 0:1]
    s1 = scales[None, :, 1:2]
    if norm_rot_grad:
        v_Rq = torch.stack([v_RS0 / s0, v_RS1 / s1, v_tn], -1)
    else:
        v_Rq = torch.stack([v_RS0 * s0, v_RS1 * s1, v_tn], -1)
    v_quat = _quat_to_rotmat_vjp(quats[None].expand(C, N, 4), v_Rq)
    v_scale = torch.zeros(C, N, 3, dtype=dt)
    v_scale[..., 0] = (v_RS0 * Rq[..., :, 0]).sum(-1)
    v_scale[..., 1] = (v_RS1 * Rq[..., :, 1]).sum(-1)
    m = valid[..., None].to(dt)
    return ((v_mean * m).sum(0), (v_quat * m).sum(0), (v_scale * m).sum(0))


def _quat_to_rotmat_vjp(quats: Tensor, v_R: Tensor) -> Tensor:
    """VJP of ``quat_to_rotmat`` (normalisation included), via autograd on the
    closed-form rotation matrix: the reference's hand formula
    (``utils.cuh:41-65``) is the same derivative."""
    with torch.enable_grad():
        q = quats.detach().clone().requires_grad_(True)
        R = quat_to_rotmat(q)
        (g,) = torch.autograd.grad(R, q, v_R)
    return g
