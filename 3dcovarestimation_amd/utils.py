"""Grid radii and the soft-edged convolution kernel of the image mask (cov3d/utils.py:258-298).

The radius grid restates ASPIRE 0.13's ``grid_2d`` / ``grid_3d`` with ``normalized=False``
(the reference imports them from aspire.utils): coordinates ceil(k) for k in [-L/2, L/2), or,
with ``shifted`` and even L, the half-pixel grid k + 1/2, so that the grid is symmetric about
the box centre.
"""

from __future__ import annotations

import math

import torch

from .projection_funcs import centered_fft2, centered_fft3


def grid_radius(L: int, dim: int, shifted: bool = False, dtype=torch.float64) -> torch.Tensor:
    """Distance of each grid point from the origin, (L,) * dim, unnormalised."""
    if shifted and L % 2 == 0:
        g = torch.arange(-L / 2 + 0.5, L / 2 + 0.5, dtype=dtype)
    else:
        g = torch.ceil(torch.arange(-L / 2, L / 2, dtype=dtype))
    axes = torch.meshgrid(*([g] * dim), indexing="ij")
    return torch.sqrt(sum(a * a for a in axes))


def soft_edged_kernel(radius: float, L: int, dim: int, radius_backoff: int = 2,
                      in_fourier: bool = False) -> torch.Tensor:
    """Normalised disc (dim 2) or ball (dim 3) with a raised-cosine edge (utils.py:258-298).

    1 inside radius - radius_backoff, (1 + cos(pi (r - r0) / (radius - r0))) / 2 on the edge,
    0 outside, on the shifted grid; divided by its sum; optionally its centred FFT.  Radii
    below 3 are raised to 3, as in the reference.
    """
    if radius < 3:
        radius = 3
    r = grid_radius(L, dim, shifted=True)
    r0 = radius - radius_backoff
    kern = torch.where(r < r0, torch.ones_like(r), torch.zeros_like(r))
    edge = (r >= r0) & (r < radius)
    kern = torch.where(edge, (1 + torch.cos(math.pi * (r - r0) / (radius - r0))) / 2, kern)
    kern = kern / kern.sum()
    if in_fourier:
        kern = centered_fft2(kern) if dim == 2 else centered_fft3(kern)
    return kern
