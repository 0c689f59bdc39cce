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


def test_shell_average_is_the_per_shell_mean():
    """shell_average against a loop over the shells of the integer radius map (even and odd L)."""
    pc = pkg("precond")
    for L in (6, 7):
        rng = np.random.default_rng(L)
        x = rng.standard_normal(L**3)
        ax = np.ceil(np.arange(-L / 2, L / 2))[:L]
        r = np.sqrt(ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2).ravel()
        n = int(np.ceil(r.max()))
        shell = np.minimum(np.floor(r + 0.5), n - 1).astype(int)
        want = np.empty_like(x)
        for s in range(n):
            want[shell == s] = x[shell == s].mean()
        got = pc.shell_average(torch.tensor(x), L).numpy()
        assert np.allclose(got, want, atol=1e-12)
        # idempotent, and it keeps the total
        assert np.allclose(pc.shell_average(torch.tensor(got), L).numpy(), got, atol=1e-12)
        assert abs(got.sum() - x.sum()) < 1e-9


def test_feature_set_round_trips_through_save_and_load(tmp_path):
    pc = pkg("precond")
    cols = pc.FEATURE_SETS["shell"]
    m = pc.DiagonalMLP(len(cols), device="cpu", seed=3, columns=cols)
    path = tmp_path / "w.npz"
    pc.save(m, path)
    back = pc.load(path, "cpu")
    assert back.columns == cols and back.n_in == len(cols)
    assert torch.equal(back.params, m.params)
    # the committed networks: per-shell (default) and per-voxel
    assert pc.load_default("cpu").columns == pc.FEATURE_SETS["shell"]
    assert pc.load_default("cpu", "voxel").columns == pc.FEATURE_SETS["voxel"]
