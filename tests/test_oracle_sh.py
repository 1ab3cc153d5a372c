"""Pins of the SH oracle: scipy's complex spherical harmonics converted to the
real basis used by the reference (Condon-Shortley phase, m<0 -> sqrt2 Im,
m>0 -> sqrt2 Re), the degree-0 constant, and autograd vs central differences."""
import math

import numpy as np
import pytest
import scipy.special as sp
import torch

from oracle.sh import sh_basis, spherical_harmonics


def _real_sh_scipy(l, m, dirs):
    x, y, z = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    polar = np.arccos(np.clip(z, -1, 1))
    azim = np.arctan2(y, x)
    if hasattr(sp, "sph_harm_y"):
        Y = sp.sph_harm_y(l, abs(m), polar, azim)
    else:  # older scipy: sph_harm(m, l, azimuth, polar)
        Y = sp.sph_harm(abs(m), l, azim, polar)
    if m < 0:
        return math.sqrt(2) * Y.imag
    if m > 0:
        return math.sqrt(2) * Y.real
    return Y.real


def test_basis_matches_scipy_real_sh():
    rng = np.random.default_rng(0)
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    B = sh_basis(4, torch.from_numpy(d)).numpy()
    k = 0
    for l in range(5):
        for m in range(-l, l + 1):
            assert np.allclose(B[:, k], _real_sh_scipy(l, m, d), atol=1e-12), (l, m)
            k += 1


def test_degree0_constant_and_unnormalised_dirs():
    c = torch.rand(7, 1, 3, dtype=torch.float64)
    out = spherical_harmonics(0, torch.randn(7, 3, dtype=torch.float64), c)
    assert torch.allclose(out, 0.28209479177387814 * c[:, 0])
    d = torch.randn(7, 3, dtype=torch.float64)
    c3 = torch.randn(7, 16, 3, dtype=torch.float64)
    assert torch.allclose(spherical_harmonics(3, d, c3), spherical_harmonics(3, 5.0 * d, c3))


def test_masks_zero():
    d = torch.randn(5, 3, dtype=torch.float64)
    c = torch.randn(5, 9, 3, dtype=torch.float64)
    m = torch.tensor([True, False, True, False, True])
    out = spherical_harmonics(2, d, c, m)
    assert torch.all(out[~m] == 0)


@pytest.mark.parametrize("deg", [1, 2, 3, 4])
def test_dir_gradient_central_differences(deg):
    g = torch.Generator().manual_seed(deg)
    d = torch.randn(4, 3, generator=g, dtype=torch.float64)
    c = torch.randn(4, 25, 3, generator=g, dtype=torch.float64)
    w = torch.randn(4, 3, generator=g, dtype=torch.float64)
    dd = d.clone().requires_grad_(True)
    (spherical_harmonics(deg, dd, c) * w).sum().backward()
    eps = 1e-6
    for i in range(4):
        for k in range(3):
            p, q = d.clone(), d.clone()
            p[i, k] += eps
            q[i, k] -= eps
            fd = ((spherical_harmonics(deg, p, c) - spherical_harmonics(deg, q, c)) * w).sum() / (2 * eps)
            assert abs(fd.item() - dd.grad[i, k].item()) < 1e-7
