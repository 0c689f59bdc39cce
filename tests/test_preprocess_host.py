"""Image preprocessing helpers on the host: translations as phase shifts (poses.py:68-125) and
the soft-edged mask kernel (utils.py:258-298).  Pinned by the DFT shift theorem and by the
kernel's normalisation and symmetry; the exact ASPIRE grid convention behind the kernel's radius
(grid_2d(shifted=True)) is restated, not imported (ASPIRE is absent)."""

import math

import pytest
import torch

from tests.conftest import pkg


@pytest.mark.parametrize("L", [16, 15])
def test_phase_shift_is_an_integer_translation(L):
    """Multiplying the centred spectrum by exp(i (g_x o_x + g_y o_y)) moves the image by -o:
    out[y, x] = img[y + o_y, x + o_x] (cyclically), exactly for integer offsets."""
    P = pkg("poses")
    pf = pkg("projection_funcs")
    g = torch.Generator().manual_seed(L)
    img = torch.randn((4, L, L), generator=g, dtype=torch.float64)
    o = torch.tensor([[2.0, -3.0], [0.0, 1.0], [-5.0, 4.0], [7.0, 7.0]], dtype=torch.float64)
    out = pf.centered_ifft2(pf.centered_fft2(img) * P.offset_to_phase_shift(o, L))
    ref = torch.stack([torch.roll(img[i], shifts=(-int(o[i, 1]), -int(o[i, 0])), dims=(-2, -1)) for i in range(4)])
    assert (out.real - ref).abs().max() < 1e-12
    assert out.imag.abs().max() < 1e-12
    # the same through an explicit grid, and a unit-modulus factor that is 1 at the zero frequency
    grid = P.get_phase_shift_grid(L, dtype=torch.float64)
    ph = P.offset_to_phase_shift(o, phase_shift_grid=grid)
    assert torch.allclose(ph, P.offset_to_phase_shift(o, L))
    assert torch.allclose(ph.abs(), torch.ones_like(ph.real))
    assert torch.allclose(ph[:, L // 2, L // 2], torch.ones(4, dtype=ph.dtype))


def test_phase_shift_grid_axes_and_errors():
    P = pkg("poses")
    gx, gy = P.get_phase_shift_grid(8, dtype=torch.float64)
    k = torch.arange(-4, 4, dtype=torch.float64) * 2 * math.pi / 8
    assert torch.equal(gx[3], k) and torch.equal(gy[:, 5], k)
    with pytest.raises(ValueError):
        P.offset_to_phase_shift(torch.zeros((1, 2)))


@pytest.mark.parametrize("L,dim", [(32, 2), (31, 2), (16, 3)])
def test_soft_edged_kernel(L, dim):
    """Sums to 1 (its centred FFT is 1 at the zero frequency), is flat inside radius - 2, zero
    outside the radius, decreasing with the radius in between, and symmetric about the box
    centre (about the half-pixel point for even L)."""
    u = pkg("utils")
    k = u.soft_edged_kernel(radius=5, L=L, dim=dim)
    assert abs(float(k.sum()) - 1.0) < 1e-12
    r = u.grid_radius(L, dim, shifted=True)
    kmax = float(k.max())
    assert torch.allclose(k[r < 3], torch.full_like(k[r < 3], kmax))
    assert (k[r >= 5] == 0).all()
    edge = (r >= 3) & (r < 5)
    order = torch.argsort(r[edge])
    assert (torch.diff(k[edge][order]) <= 1e-15).all()
    # even L: the shifted grid is symmetric about the half-pixel point; odd L: about index L // 2
    for d in range(dim):
        assert torch.allclose(k, k.flip(d))
    kf = u.soft_edged_kernel(radius=5, L=L, dim=dim, in_fourier=True)
    c = (L // 2,) * dim
    assert abs(complex(kf[c]) - 1.0) < 1e-12
    # radii below 3 are raised to 3
    assert torch.equal(u.soft_edged_kernel(radius=2, L=L, dim=dim), u.soft_edged_kernel(radius=3, L=L, dim=dim))
