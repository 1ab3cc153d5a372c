"""Real spherical harmonics colour evaluation: CPU restatement.

Test infrastructure only.  Same basis, sign convention and band order as the
reference (``textured_gaussians/cuda/csrc/spherical_harmonics.cuh:17-111``,
Sloan, "Efficient Spherical Harmonic Evaluation", JCGT 2013): the
orthonormal real SH with the Condon-Shortley phase, written here directly as
polynomials in the unit direction.  Differentiable (autograd gives the VJP of
``spherical_harmonics.cuh:113-387``).
"""
import math

import torch
from torch import Tensor


def sh_basis(degree: int, dirs: Tensor) -> Tensor:
    """[..., (degree+1)^2] basis values at ``dirs`` (normalised inside)."""
    out = [torch.full_like(dirs[..., 0], 0.5 / math.sqrt(math.pi))]
    if degree >= 1:
        d = dirs / torch.linalg.norm(dirs, dim=-1, keepdim=True)
        x, y, z = d.unbind(-1)
        c1 = math.sqrt(3.0 / (4 * math.pi))
        out += [-c1 * y, c1 * z, -c1 * x]
        if degree >= 2:
            c2a = 0.5 * math.sqrt(15.0 / math.pi)
            c2b = 0.25 * math.sqrt(5.0 / math.pi)
            c2c = 0.25 * math.sqrt(15.0 / math.pi)
            out += [c2a * x * y, -c2a * y * z, c2b * (3 * z * z - 1),
                    -c2a * x * z, c2c * (x * x - y * y)]
            if degree >= 3:
                a = 0.25 * math.sqrt(35.0 / (2 * math.pi))
                b = 0.5 * math.sqrt(105.0 / math.pi)
                c = 0.25 * math.sqrt(21.0 / (2 * math.pi))
                e = 0.25 * math.sqrt(7.0 / math.pi)
                g = 0.25 * math.sqrt(105.0 / math.pi)
                out += [-a * y * (3 * x * x - y * y), b * x * y * z,
                        -c * y * (5 * z * z - 1), e * z * (5 * z * z - 3),
                        -c * x * (5 * z * z - 1), g * z * (x * x - y * y),
                        -a * x * (x * x - 3 * y * y)]
                if degree >= 4:
                    a4 = 0.75 * math.sqrt(35.0 / math.pi)
                    b4 = 0.75 * math.sqrt(35.0 / (2 * math.pi))
                    c4 = 0.75 * math.sqrt(5.0 / math.pi)
                    d4 = 0.75 * math.sqrt(5.0 / (2 * math.pi))
                    e4 = 3.0 / 16.0 * math.sqrt(1.0 / math.pi)
                    f4 = 0.375 * math.sqrt(5.0 / math.pi)
                    h4 = 3.0 / 16.0 * math.sqrt(35.0 / math.pi)
                    x2, y2, z2 = x * x, y * y, z * z
                    out += [a4 * x * y * (x2 - y2), -b4 * y * z * (3 * x2 - y2),
                            c4 * x * y * (7 * z2 - 1), -d4 * y * z * (7 * z2 - 3),
                            e4 * (35 * z2 * z2 - 30 * z2 + 3),
                            -d4 * x * z * (7 * z2 - 3), f4 * (x2 - y2) * (7 * z2 - 1),
                            -b4 * x * z * (x2 - 3 * y2),
                            h4 * (x2 * (x2 - 3 * y2) - y2 * (3 * x2 - y2))]
    return torch.stack(out, -1)


def spherical_harmonics(degree: int, dirs: Tensor, coeffs: Tensor, masks: Tensor = None
                        ) -> Tensor:
    """colors[..., 3] = sum_k basis_k(dirs) * coeffs[..., k, :]; masked-out
    entries are 0 (the reference leaves them unwritten)."""
    nb = (degree + 1) ** 2
    B = sh_basis(degree, dirs)
    col = (B[..., :, None] * coeffs[..., :nb, :]).sum(-2)
    if masks is not None:
        col = torch.where(masks[..., None], col, torch.zeros_like(col))
    return col
