"""depth_to_normal (textured_gaussians/utils.py:145-168 semantics: back-project pixel
centres, central differences of the points, normalised cross product, zero border),
pinned by closed forms: a fronto-parallel plane gives the constant normal (0, 0, -1) in
the camera frame; any plane gives minus its unit normal (oriented towards the camera);
a camera-to-world rotation rotates the result; the one-pixel border is zero."""
import math

import torch

from textured_gaussians_amd.utils import depth_to_normal, depth_to_points

F64 = torch.float64


def _K(fx=200.0, fy=180.0, cx=40.0, cy=30.0):
    return torch.tensor([[fx, 0, cx], [0, fy, cy], [0, 0, 1]], dtype=F64)


def _plane_depth(n, d, K, H, W):
    """z-depth of the plane n . p = d seen through each pixel centre."""
    y, x = torch.meshgrid(torch.arange(H, dtype=F64), torch.arange(W, dtype=F64), indexing="ij")
    u = (x - K[0, 2] + 0.5) / K[0, 0]
    v = (y - K[1, 2] + 0.5) / K[1, 1]
    return (d / (n[0] * u + n[1] * v + n[2]))[..., None]


def test_fronto_parallel_plane_gives_constant_normal():
    H, W = 24, 32
    depth = torch.full((H, W, 1), 3.7, dtype=F64)
    nrm = depth_to_normal(depth, torch.eye(4, dtype=F64), _K())
    inner = nrm[1:-1, 1:-1]
    assert torch.allclose(inner, torch.tensor([0.0, 0.0, -1.0], dtype=F64).expand_as(inner), atol=1e-12)
    # zero border
    assert float(nrm[0].abs().max()) == 0.0 and float(nrm[-1].abs().max()) == 0.0
    assert float(nrm[:, 0].abs().max()) == 0.0 and float(nrm[:, -1].abs().max()) == 0.0


def test_tilted_plane_gives_minus_its_unit_normal():
    H, W = 20, 28
    K = _K()
    n = torch.tensor([0.3, -0.5, 1.0], dtype=F64)
    depth = _plane_depth(n, 4.0, K, H, W)
    pts = depth_to_points(depth, torch.eye(4, dtype=F64), K)
    assert float((pts @ n - 4.0).abs().max()) < 1e-12  # the points lie on the plane
    nrm = depth_to_normal(depth, torch.eye(4, dtype=F64), K)[1:-1, 1:-1]
    want = -n / n.norm()
    assert float((nrm - want).abs().max()) < 1e-10


def test_camera_rotation_rotates_the_normals():
    H, W = 16, 20
    K = _K()
    n = torch.tensor([-0.2, 0.4, 1.0], dtype=F64)
    depth = _plane_depth(n, 2.5, K, H, W)
    a = 0.7
    R = torch.tensor([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]], dtype=F64)
    c2w = torch.eye(4, dtype=F64)
    c2w[:3, :3] = R
    c2w[:3, 3] = torch.tensor([1.0, -2.0, 0.5], dtype=F64)  # a translation does not change normals
    got = depth_to_normal(depth, c2w, K)[1:-1, 1:-1]
    want = R @ (-n / n.norm())
    assert float((got - want).abs().max()) < 1e-10


def test_batched_cameras():
    H, W = 12, 14
    K = _K()
    depth = torch.stack([torch.full((H, W, 1), 1.5, dtype=F64), _plane_depth(torch.tensor([0.1, 0.2, 1.0],
                                                                                          dtype=F64), 3.0, K, H, W)])
    out = depth_to_normal(depth, torch.eye(4, dtype=F64)[None].repeat(2, 1, 1), K[None].repeat(2, 1, 1))
    assert out.shape == (2, H, W, 3)
    assert torch.allclose(out[0, 1:-1, 1:-1], torch.tensor([0.0, 0.0, -1.0], dtype=F64).expand(H - 2, W - 2, 3))
