"""Fourier-shell statistics (cov3d/fsc_utils.py:9-367), vectorised.

Shells follow FourierShell (fsc_utils.py:9-94): radius r of the centred
integer grid ceil(arange(-L/2, L/2)); shell i holds i-0.5 < r < i+0.5, the
last shell extends to its upper edge + 1; there are ceil(max r) shells.  The
reference loops over shells in Python; here one shell-index map per (L, dim,
device) drives scatter-add averages and gathers.
"""

from __future__ import annotations

import math
from functools import lru_cache
from typing import Tuple

import torch

from .projection_funcs import centered_fft3


@lru_cache(maxsize=32)
def _shell_index(L: int, dim: int, device_str: str) -> Tuple[torch.Tensor, torch.Tensor, int]:
    ax = torch.ceil(torch.arange(-L / 2, L / 2, dtype=torch.float64))[:L]
    grids = torch.meshgrid(*([ax] * dim), indexing="ij")
    r = torch.sqrt(sum(g**2 for g in grids))
    n = int(math.ceil(float(r.max())))
    n = max(n, 1)
    idx = torch.clamp(torch.floor(r + 0.5), max=n - 1).to(torch.int64)
    size = torch.bincount(idx.reshape(-1), minlength=n).to(torch.float64)
    dev = torch.device(device_str)
    return idx.reshape(-1).to(dev), size.to(dev), n


def shell_index(L: int, dim: int, device) -> Tuple[torch.Tensor, torch.Tensor, int]:
    """(flat shell index per voxel, shell sizes, number of shells)."""
    return _shell_index(int(L), int(dim), str(torch.device(device)))


def _avg(x: torch.Tensor, dim: int) -> torch.Tensor:
    L = x.shape[-1]
    idx, size, n = shell_index(L, dim, x.device)
    flat = x.reshape(-1, L**dim)
    out = torch.zeros((flat.shape[0], n), dtype=flat.dtype, device=x.device)
    out.index_add_(1, idx, flat)
    return out / size.to(out.real.dtype if out.is_complex() else out.dtype)


def average_fourier_shell(*spectrum_signals: torch.Tensor) -> torch.Tensor:
    """Shell averages of one or more signals; returns (n_shells,) for one, (k, n_shells) for k."""
    x = torch.stack(spectrum_signals) if len(spectrum_signals) > 1 else spectrum_signals[0]
    dim = spectrum_signals[0].ndim
    if len(spectrum_signals) == 1:
        return _avg(x, dim)[0]
    return _avg(x, dim)


def rpsd(*signals: torch.Tensor) -> torch.Tensor:
    """Radial power spectral density of 3-D signals (fsc_utils.py:134-147, :212-221)."""
    x = torch.stack(signals)
    psd = torch.abs(centered_fft3(x)) ** 2
    out = _avg(psd, 3)
    return out[0] if len(signals) == 1 else out


def expand_fourier_shell(shells: torch.Tensor, L: int, dim: int) -> torch.Tensor:
    """Shell values -> full (L,)*dim grid (fsc_utils.py:164-184)."""
    idx, _, n = shell_index(L, dim, shells.device)
    s = shells.unsqueeze(0) if shells.ndim == 1 else shells
    out = s[:, idx].reshape((s.shape[0],) + (L,) * dim)
    return out[0] if shells.ndim == 1 or s.shape[0] == 1 else out


def upsample_and_expand_fourier_shell(shells: torch.Tensor, L: int, dim: int) -> torch.Tensor:
    """Linear resampling of shell values to L's shell count, then expansion (fsc_utils.py:237-256)."""
    _, _, n = shell_index(L, dim, shells.device)
    x = shells.unsqueeze(0).unsqueeze(0) if shells.ndim == 1 else shells.unsqueeze(0)
    up = torch.nn.functional.interpolate(x, size=n, mode="linear", align_corners=False).squeeze()
    return expand_fourier_shell(up, L, dim)


def sum_over_shell(shell: torch.Tensor, L: int, dim: int) -> torch.Tensor:
    _, size, _ = shell_index(L, dim, shell.device)
    s = shell.unsqueeze(0) if shell.ndim == 1 else shell
    out = torch.sum(s * size.to(s.dtype), dim=-1)
    return out[0] if shell.ndim == 1 else out


def vol_fsc(signal1: torch.Tensor, signal2: torch.Tensor) -> torch.Tensor:
    """Fourier shell correlation of two volumes (fsc_utils.py:272-291)."""
    f1, f2 = centered_fft3(signal1), centered_fft3(signal2)
    c, p1, p2 = average_fourier_shell(torch.real(f1 * torch.conj(f2)), f1.abs() ** 2, f2.abs() ** 2)
    return c / torch.sqrt(p1 * p2)


def covar_correlate(f1: torch.Tensor, f2: torch.Tensor, return_shell_size: bool = False):
    """Shell x shell correlation of two Fourier eigenvector sets (fsc_utils.py:314-343)."""
    L = f1.shape[-1]
    idx, size, n = shell_index(L, 3, f1.device)
    mat = torch.zeros((n, n), dtype=f1.dtype, device=f1.device)
    for i in range(f1.shape[0]):
        prod = (f1[i].unsqueeze(0) * torch.conj(f2)).reshape(f2.shape[0], -1)
        s = torch.zeros((f2.shape[0], n), dtype=f1.dtype, device=f1.device)
        s.index_add_(1, idx, prod)
        s = s / size.to(s.real.dtype)
        mat += s.T @ s.conj()
    return (mat, size) if return_shell_size else mat


def covar_rpsd(eigenvectors: torch.Tensor, from_fourier: bool = False) -> torch.Tensor:
    f = eigenvectors if from_fourier else centered_fft3(eigenvectors)
    c, size = covar_correlate(f, f, return_shell_size=True)
    return c.real * torch.outer(size, size).to(c.real.dtype)


def covar_fsc(eigenvecs1: torch.Tensor, eigenvecs2: torch.Tensor) -> torch.Tensor:
    """Covariance FSC matrix (fsc_utils.py:346-367)."""
    f1, f2 = centered_fft3(eigenvecs1), centered_fft3(eigenvecs2)
    corr = covar_correlate(f1, f2).real
    p1 = covar_correlate(f1, f1).real
    p2 = covar_correlate(f2, f2).real
    bottom = torch.sqrt((p1 / p1[0, 0]) * (p2 / p2[0, 0])) * torch.sqrt(p1[0, 0]) * torch.sqrt(p2[0, 0])
    return corr / bottom
