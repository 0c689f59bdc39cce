"""Host checks of the Covar parameterisation (cov3d/covar.py:278-458): initialisation from a scalar
pixel variance or from a power spectral density, and the eigen-decomposition helpers."""

import torch

from tests.conftest import pkg


def test_white_initialisation_has_the_pixel_variance():
    cv = pkg("covar")
    torch.manual_seed(0)
    c = cv.Covar(16, 8, pixel_var_estimate=2.0)
    v = c.get_vectors_spatial_domain().detach()
    assert abs(float(v.var()) - 2.0) < 0.05


def test_psd_initialisation_follows_the_spectrum():
    """init_random_vectors_from_psd (covar.py:389-405): the radial power spectrum of the drawn
    vectors is the requested one (unit white noise has unit spectral density), for a radial and for
    an expanded full-grid PSD; a tensor pixel_var_estimate selects this initialisation."""
    cv = pkg("covar")
    fs = pkg("fsc_utils")
    torch.manual_seed(1)
    L = 16
    psd = torch.linspace(5, 0.1, 14)
    c = cv.Covar(L, 2, pixel_var_estimate=psd)
    for p in (psd, fs.expand_fourier_shell(psd, L, 3)):
        est = fs.rpsd(*c.init_random_vectors_from_psd(400, p)).mean(0)
        # shell 0 is the single DC voxel (a 7% sampling error at 400 draws); the others hold many
        assert torch.allclose(est[1:] / psd[1:], torch.ones(13), atol=0.05), est / psd
    # the constructor drew its vectors the same way: a steep spectrum leaves high shells quiet
    steep = torch.exp(-torch.arange(14.0))
    c2 = cv.Covar(L, 50, pixel_var_estimate=steep)
    r = fs.rpsd(*c2.get_vectors_spatial_domain().detach()).mean(0)
    assert float(r[6]) < 1e-2 * float(r[0])


def test_eigenvecs_are_orthonormal_and_ordered():
    cv = pkg("covar")
    torch.manual_seed(2)
    c = cv.Covar(8, 4, pixel_var_estimate=1.0)
    ev, lam = c.eigenvecs
    g = ev.reshape(4, -1) @ ev.reshape(4, -1).T
    assert torch.allclose(g, torch.eye(4), atol=1e-5)
    assert bool((lam[:-1] >= lam[1:]).all())
    V = c.get_vectors_spatial_domain().detach().reshape(4, -1)
    S = V.T @ V
    S2 = (ev.reshape(4, -1).T * lam) @ ev.reshape(4, -1)
    assert torch.allclose(S, S2, atol=1e-4 * float(S.abs().max()))
