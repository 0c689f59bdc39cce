"""Centred transform of zero-padded real volumes on the pruned line FFTs (cov3d_fft3c_pad_real and
its adjoint), against torch's FFT of the same expression in FP64 (the reference's
centered_fft3(v / grid_correction, padding_size) of covar.py:440-446)."""

import numpy as np
import pytest
import torch

from tests.conftest import pkg, rel_l2

pytestmark = pytest.mark.gpu


def _sinc_axis(L, us):
    cv = pkg("covar")
    gc = cv.sinc_grid_correction(L, us, "bilinear")
    return gc[:, L // 2, L // 2]


@pytest.mark.parametrize("r,L,nf,weighted", [(3, 32, 64, False), (2, 33, 64, True), (4, 64, 128, True),
                                             (2, 63, 128, False), (2, 128, 256, True), (1, 100, 256, True)])
def test_forward_matches_fp64_torch(dev, r, L, nf, weighted):
    pf = pkg("projection_funcs")
    g = torch.Generator(device=dev).manual_seed(L + nf)
    v = torch.randn((r, L, L, L), device=dev, generator=g)
    w = (1.0 / _sinc_axis(L, nf // L if nf % L == 0 else 2)).to(dev) if weighted else None
    out = pf.fft3c_pad_real(v, nf, w)
    assert out.shape == (r, nf, nf, nf) and out.dtype == torch.complex64
    v64 = v.double()
    if w is not None:
        w64 = w.double()
        v64 = v64 * (w64[:, None, None] * w64[None, :, None] * w64[None, None, :])
    ref = pf.centered_fft3(v64, padding_size=(nf,) * 3)
    err = rel_l2(out.cpu().numpy(), ref.cpu().numpy())
    # the mirrored half must be the conjugate of the computed one to rounding as well
    print(f"r={r} L={L} nf={nf} weighted={weighted}: rel err {err:.2e}")
    assert err < 2e-6


@pytest.mark.parametrize("r,L,nf", [(2, 32, 64), (3, 33, 64), (2, 128, 256)])
def test_adjoint_is_the_autograd_gradient(dev, r, L, nf):
    """The native backward equals torch autograd through the FP64 expression, and the pair passes the
    dot-product test Re<A v, g> = <v, A* g>."""
    pf = pkg("projection_funcs")
    g = torch.Generator(device=dev).manual_seed(7 * L)
    v = torch.randn((r, L, L, L), device=dev, generator=g)
    w = (1.0 / _sinc_axis(L, 2)).to(dev)
    G = torch.complex(torch.randn((r, nf, nf, nf), device=dev, generator=g),
                      torch.randn((r, nf, nf, nf), device=dev, generator=g))
    va = v.clone().requires_grad_(True)
    out = pf.fft3c_pad_real(va, nf, w)
    (out * G.conj()).real.sum().backward()
    v64 = v.double().clone().requires_grad_(True)
    w64 = w.double()
    ref = pf.centered_fft3(v64 * (w64[:, None, None] * w64[None, :, None] * w64[None, None, :]),
                           padding_size=(nf,) * 3)
    (ref * G.to(torch.complex128).conj()).real.sum().backward()
    err = rel_l2(va.grad.cpu().numpy(), v64.grad.cpu().numpy())
    lhs = float((out.detach().to(torch.complex128) * G.to(torch.complex128).conj()).real.sum())
    rhs = float((v.double() * va.grad.double()).sum())
    print(f"r={r} L={L} nf={nf}: gradient rel err {err:.2e}, dot test {lhs:.6e} vs {rhs:.6e}")
    assert err < 2e-6
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)


def test_covar_and_mean_use_it_with_the_sinc_correction(dev):
    """Covar/Mean Fourier-domain volumes (bilinear sinc^2 correction, us=2) equal the reference's
    expression v / grid_correction padded and transformed, in FP64."""
    cv = pkg("covar")
    pf = pkg("projection_funcs")
    torch.manual_seed(3)
    L = 64
    covar = cv.Covar(L, 3, upsampling_factor=2).to(dev)
    covar.init_grid_correction("bilinear")
    assert covar._axis_weights() is not None and covar._axis_weights() is not False
    vf = covar.get_vectors_fourier_domain()
    V = covar.get_vectors_spatial_domain().detach().double()
    ref = pf.centered_fft3(V / covar.grid_correction.double(), padding_size=(2 * L,) * 3)
    assert rel_l2(vf.detach().cpu().numpy(), ref.cpu().numpy()) < 2e-6
    mean = cv.Mean(torch.randn((L, L, L), device=dev), L, upsampling_factor=2).to(dev)
    mean.init_grid_correction("nearest")
    mf = mean.get_volume_fourier_domain()
    M = mean.get_volume_spatial_domain().detach().double()
    ref = pf.centered_fft3(M / mean.grid_correction.double(), padding_size=(2 * L,) * 3)
    assert rel_l2(mf.detach().cpu().numpy(), ref.cpu().numpy()) < 2e-6


def test_unsupported_sizes_take_torch(dev):
    pf = pkg("projection_funcs")
    v = torch.randn((2, 16, 16, 16), device=dev)
    assert not pf.fft3c_pad_real_supported(v, 16)      # no padding room
    assert not pf.fft3c_pad_real_supported(v.double(), 32)
    out = pf.fft3c_pad_real(v, 40)                      # 40 is no pruned grid: torch's FFT
    ref = pf.centered_fft3(v.double(), padding_size=(40,) * 3)
    assert rel_l2(out.cpu().numpy(), ref.cpu().numpy()) < 1e-5
