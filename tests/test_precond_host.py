"""CPU checks of the learned-preconditioner host logic.

* ``probe_diagonal`` (the training target) equals the diagonal of the Toeplitz
  CG operator in the L^3 Fourier basis, built column by column from the
  oracle's operator (brute force on a tiny grid);
* ``preconditioner_from_log_diag`` refuses non-finite / non-positive weights
  (the plain-CG fallback), and with exact Jacobi weights PCG on the oracle
  operator needs fewer iterations than plain CG.
"""

import numpy as np
import pytest
import torch

from tests.conftest import oracle, pkg


def _lhs(nf, seed):
    rng = np.random.default_rng(seed)
    ax = np.arange(nf) - nf // 2
    r = np.sqrt(ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2)
    lhs = (1.0 + 50.0 / (1 + r)) * rng.uniform(0.8, 1.2, (nf,) * 3)
    return lhs + 0j


def test_probe_diagonal_matches_explicit_operator():
    o = oracle()
    pc = pkg("precond")
    L, nf = 4, 8
    lhs = _lhs(nf, 0)
    T = pc.probe_diagonal(torch.tensor(lhs), L).double().numpy()
    diag = np.zeros(L**3)
    for k in range(L**3):
        e = np.zeros(L**3, dtype=complex)
        e[k] = 1
        Fe = o.ifft3c(e.reshape(L, L, L))  # spatial vector whose centred FFT is e_k
        AFe = o.toeplitz_apply(Fe, lhs, L)
        diag[k] = (np.vdot(Fe.ravel(), AFe.ravel()) / np.vdot(Fe.ravel(), Fe.ravel())).real
    assert np.allclose(T, diag, rtol=1e-5)


def test_invalid_weights_fall_back_and_jacobi_helps():
    o = oracle()
    pc = pkg("precond")
    L, nf = 6, 12
    assert pc.preconditioner_from_log_diag(torch.full((L**3,), float("nan")), L) is None
    assert pc.preconditioner_from_log_diag(torch.full((L**3,), float("inf")) * -1, L) is None
    lhs = _lhs(nf, 1)
    T = pc.probe_diagonal(torch.tensor(lhs), L).double().numpy().reshape(L, L, L)
    A = lambda v: o.toeplitz_apply(v, lhs, L).ravel()
    b = np.random.default_rng(2).standard_normal(L**3) + 0j
    _, it_plain, _ = o.conjugate_gradient(A, b, rtol=1e-8)

    class _M:
        def __mul__(self, r):
            return self.__rmul__(r)

        def __rmul__(self, r):
            return o.ifft3c(o.fft3c(r.reshape(L, L, L)) / T).ravel()

    x, it_pcg, _ = o.conjugate_gradient(A, b, rtol=1e-8, M=_M())
    assert np.linalg.norm(A(x) - b) / np.linalg.norm(b) < 1e-7
    assert it_pcg < it_plain


@pytest.mark.parametrize("L", [7, 8])
def test_preconditioner_equals_centred_fft_form(L):
    """The preconditioner's two plain FFTs with rolled weights equal ifft3c(w * fft3c(r))
    (odd and even L)."""
    pc = pkg("precond")
    pf = pkg("projection_funcs")
    g = torch.Generator().manual_seed(L)
    logd = torch.randn(L**3, generator=g)
    apply = pc.preconditioner_from_log_diag(logd, L)
    r = torch.complex(torch.randn(L**3, generator=g), torch.randn(L**3, generator=g))
    w = torch.exp(-logd).reshape(L, L, L).to(torch.complex64)
    ref = pf.centered_ifft3(w * pf.centered_fft3(r.reshape(L, L, L))).reshape(-1)
    assert (apply(r) - ref).abs().max() < 1e-5 * ref.abs().max()
