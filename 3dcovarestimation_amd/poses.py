"""Per-image in-plane translations as Fourier phase shifts (cov3d/poses.py).

The image preprocessing step (projection_funcs.py:199-252) multiplies each image's centred
spectrum by exp(i (g_x o_x + g_y o_y)); the pose module passes the negated stored offsets
(poses.py:263-264), so that a stored offset o undoes a translation of the particle by o.
"""

from __future__ import annotations

import math
from typing import Optional, Tuple

import torch


def get_phase_shift_grid(resolution: int, dtype: torch.dtype = torch.float32,
                         device: Optional[torch.device] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """Frequency grids (g_x, g_y), each (L, L) (poses.py:68-87).

    g = 2 pi ceil(k) / L for k = -L/2, -L/2 + 1, ..., L/2 - 1 (the centred DFT frequencies);
    g_x varies along the last axis and g_y along the first (``meshgrid(..., indexing="xy")``).
    """
    k = torch.ceil(torch.arange(-resolution / 2, resolution / 2, dtype=torch.float64))
    g = (k * (2 * math.pi / resolution)).to(dtype)
    gy, gx = torch.meshgrid(g, g, indexing="ij")
    dev = device if device is not None else torch.device("cpu")
    return gx.to(dev), gy.to(dev)


def offset_to_phase_shift(offsets: torch.Tensor, resolution: Optional[int] = None,
                          phase_shift_grid: Optional[Tuple[torch.Tensor, torch.Tensor]] = None) -> torch.Tensor:
    """exp(i (g_x o_x + g_y o_y)) for offsets (N, 2) -> (N, L, L) complex (poses.py:92-125).

    Raises ValueError when neither ``resolution`` nor ``phase_shift_grid`` is given, as the
    reference does.
    """
    if phase_shift_grid is None:
        if resolution is None:
            raise ValueError("Either resolution or phase_shift_grid must be provided.")
        phase_shift_grid = get_phase_shift_grid(resolution, dtype=offsets.dtype, device=offsets.device)
    gx, gy = phase_shift_grid
    phase = gx.unsqueeze(0) * offsets[:, 0].reshape(-1, 1, 1) + gy.unsqueeze(0) * offsets[:, 1].reshape(-1, 1, 1)
    return torch.polar(torch.ones_like(phase), phase)
