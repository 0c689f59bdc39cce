"""Pins the remaining oracle functions on the CPU.

* slice extraction: torch.nn.functional.grid_sample (the op the reference
  calls, nufft_plan.py:148) on the same grid; adjoint by the dot-product test;
* ML objective: Woodbury form vs the dense d x d Gaussian likelihood
  (reference test_torch_impl.py:248-265), closed-form gradient vs central
  finite differences and vs torch autograd in FP64;
* CTF: zeros of the textbook formula at the closed-form radii;
* pts_rot: identity rotation gives the plain grid; rotations preserve radius;
* shells: the vectorised product helper vs the reference's loop definition;
* CG: converges to the dense solve; Toeplitz operator vs an explicit matrix;
* MLP: numpy forward vs a torch module built from the same parameters.
"""

import math

import numpy as np
import pytest
import torch

from tests.conftest import oracle, pkg, rel_l2


@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
@pytest.mark.parametrize("vL", [8, 9])
def test_slice_forward_matches_grid_sample(mode, vL):
    o = oracle()
    rng = np.random.default_rng(10)
    L = vL // 2 + 1
    R = np.linalg.qr(rng.standard_normal((4, 3, 3)))[0]
    pts = o.pts_rot(R, L).astype(np.float64)
    grid = o.discretized_grid(pts, vL)
    vol = rng.standard_normal((2, vL, vL, vL)) + 1j * rng.standard_normal((2, vL, vL, vL))
    ours = o.slice_forward(vol, grid, mode)
    inp = torch.tensor(np.concatenate([vol.real, vol.imag], 0)[[0, 2, 1, 3]])  # [re0, im0, re1, im1]
    g = torch.tensor(grid).reshape(1, 1, 1, -1, 3)
    ref = torch.nn.functional.grid_sample(inp.unsqueeze(0), g, mode=mode, align_corners=True)[0, :, 0, 0]
    ref = ref.numpy().reshape(2, 2, -1)
    assert rel_l2(ours, ref[:, 0] + 1j * ref[:, 1]) < 1e-12


@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
def test_slice_adjoint_dot_product(mode):
    o = oracle()
    rng = np.random.default_rng(11)
    vL, L = 8, 4
    pts = o.pts_rot(np.linalg.qr(rng.standard_normal((3, 3, 3)))[0], L).astype(np.float64)
    grid = o.discretized_grid(pts, vL)
    vol = rng.standard_normal((1, vL, vL, vL)) + 1j * rng.standard_normal((1, vL, vL, vL))
    sig = rng.standard_normal((1, grid.shape[0])) + 1j * rng.standard_normal((1, grid.shape[0]))
    lhs = np.vdot(o.slice_forward(vol, grid, mode), sig)
    rhs = np.vdot(vol, o.slice_adjoint(sig, grid, (vL, vL, vL), mode))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def _rand_ml(seed=12, B=3, r=3, d=20):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((B, r, d)), rng.standard_normal((B, d)), 0.7


def test_ml_woodbury_equals_dense_likelihood():
    o = oracle()
    P, y, s2 = _rand_ml()
    cost_i, *_ = o.ml_terms(P, y, s2)
    assert np.allclose(cost_i, o.ml_cost_naive(P, y, s2), rtol=1e-10, atol=1e-10)


def test_ml_gradient_finite_differences():
    o = oracle()
    P, y, s2 = _rand_ml(13)
    _, G = o.ml_cost_grad(P, y, s2)
    rng = np.random.default_rng(14)
    for _ in range(8):
        idx = tuple(rng.integers(0, n) for n in P.shape)
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[idx] += h
        Pm[idx] -= h
        fd = (o.ml_cost_grad(Pp, y, s2)[0] - o.ml_cost_grad(Pm, y, s2)[0]) / (2 * h)
        assert abs(fd - G[idx]) < 1e-7 * max(1.0, abs(fd))


def test_ml_gradient_autograd_fp64():
    o = oracle()
    P, y, s2 = _rand_ml(15)
    Pt = torch.tensor(P, requires_grad=True)
    yt = torch.tensor(y)
    r = P.shape[1]
    M = torch.eye(r, dtype=torch.float64) + Pt @ Pt.transpose(1, 2) / s2
    b = (Pt @ yt[..., None]) / s2
    z = torch.linalg.solve(M, b)
    f = 0.5 * torch.mean(-(b.transpose(1, 2) @ z).squeeze() + torch.logdet(M))
    f.backward()
    val, G = o.ml_cost_grad(P, y, s2)
    assert abs(val - f.item()) < 1e-12 * abs(val) + 1e-12
    assert rel_l2(G, Pt.grad.numpy()) < 1e-10


def test_ctf_zero_crossings():
    """h = sqrt(1-a^2) sin(g) - a cos(g) vanishes where g = atan(a / sqrt(1-a^2)) - pi.

    The defocus is solved in closed form so that the grid pixel (0, L/4) sits
    exactly on that zero; h(0) = -a.
    """
    o = oracle()
    L, px, kv, cs, a = 64, 1.5, 300.0, 2.7, 0.1
    lam = o.wavelength(kv)
    assert abs(lam - 0.019687) < 1e-6
    f = (0.5 / (2 * px))  # gx = 0.5 at column 3L/4
    c4 = 0.5 * np.pi * cs * 1e7 * lam**3
    phi0 = math.atan(a / math.sqrt(1 - a * a))
    df = (c4 * f**4 - (phi0 - np.pi)) / (np.pi * lam * f**2)
    h = o.ctf(np.array([[df, df, 0.0, kv, cs, a, px]]), L)[0]
    assert abs(h[L // 2, 3 * L // 4]) < 1e-9
    assert abs(h[3 * L // 4, L // 2]) < 1e-9  # radial: same zero on the other axis
    assert abs(h[L // 2, L // 2] + a) < 1e-12
    # astigmatism rotates with the angle: dU along x, dV along y
    h2 = o.ctf(np.array([[df, 2 * df, 0.0, kv, cs, a, px]]), L)[0]
    assert abs(h2[L // 2, 3 * L // 4]) < 1e-9 and abs(h2[3 * L // 4, L // 2]) > 1e-3


def test_pts_rot_identity_and_radius():
    o = oracle()
    L = 8
    p = o.pts_rot(np.eye(3)[None], L)[0]
    g = np.pi * o.grid2d_axis(L)
    gy, gx = np.meshgrid(g, g, indexing="ij")
    assert np.allclose(p[0], 0) and np.allclose(p[1], gy.ravel()) and np.allclose(p[2], gx.ravel())
    R = np.linalg.qr(np.random.default_rng(20).standard_normal((2, 3, 3)))[0]
    pr = o.pts_rot(R, L).astype(np.float64)
    rad_in = np.sqrt(gx.ravel() ** 2 + gy.ravel() ** 2)
    inner = rad_in < np.pi - 1e-3  # these frequencies can never reach the wrap
    for b in range(2):
        rad = np.linalg.norm(pr[b], axis=0)
        assert np.allclose(rad[inner], rad_in[inner], atol=1e-5)
        assert np.abs(pr[b]).max() <= np.pi


@pytest.mark.parametrize("L", [7, 8])
def test_shell_average_vectorised_matches_loop(L):
    o = oracle()
    fsc = pkg("fsc_utils")
    x = np.random.default_rng(21).standard_normal((L, L, L))
    ours = fsc.average_fourier_shell(torch.tensor(x)).numpy()
    assert np.allclose(ours, o.shell_average_loop(x), atol=1e-12)


def test_cg_and_toeplitz_small():
    o = oracle()
    L, nf = 4, 8
    rng = np.random.default_rng(22)
    lhs = rng.uniform(0.5, 2.0, (nf, nf, nf)).astype(complex)
    lhs = 0.5 * (lhs + np.conj(lhs[::-1, ::-1, ::-1]))  # keeps the operator Hermitian on even grids too
    A = np.zeros((L**3, L**3), dtype=complex)
    for k in range(L**3):
        e = np.zeros(L**3, dtype=complex)
        e[k] = 1
        A[:, k] = o.toeplitz_apply(e, lhs, L).ravel()
    assert np.allclose(A, A.conj().T, atol=1e-10)
    b = rng.standard_normal(L**3) + 0j
    x, it, hist = o.conjugate_gradient(lambda v: o.toeplitz_apply(v, lhs, L).ravel(), b, rtol=1e-10)
    assert rel_l2(x, np.linalg.solve(A, b)) < 1e-8


@pytest.mark.parametrize("L,nf", [(4, 8), (5, 10), (6, 16)])
def test_toeplitz_closed_forms(L, nf):
    """crop(ifft3c(lhs * fft3c(pad(x)))) (mean.py:118-123) pinned by what fixes it independently of
    any FFT convention: a constant lhs = c is c times the identity; the centred phase ramp
    lhs[k] = exp(-2 pi i k.s / nf) translates x by s with zero fill (the crop window never sees the
    circular wrap when nf >= 2L); and a random lhs equals the direct convolution of x with the
    kernel K[m] = nf^-3 sum_k lhs[k] exp(+2 pi i k.m / nf), the operator's Toeplitz form."""
    o = oracle()
    rng = np.random.default_rng(L * nf)
    x = rng.standard_normal((L,) * 3) + 1j * rng.standard_normal((L,) * 3)
    assert rel_l2(o.toeplitz_apply(x, np.full((nf,) * 3, 2.5 + 0j), L), 2.5 * x) < 1e-13
    k = np.arange(nf) - nf // 2
    s = (1, -2, 1)
    ramp = np.exp(-2j * np.pi * (k[:, None, None] * s[0] + k[None, :, None] * s[1] + k[None, None, :] * s[2]) / nf)
    shifted = np.zeros_like(x)
    src = [slice(max(0, -d), L - max(0, d)) for d in s]
    dst = [slice(max(0, d), L - max(0, -d)) for d in s]
    shifted[tuple(dst)] = x[tuple(src)]
    assert rel_l2(o.toeplitz_apply(x, ramp, L), shifted) < 1e-12
    lhs = rng.standard_normal((nf,) * 3) + 1j * rng.standard_normal((nf,) * 3)
    m = np.arange(-(L - 1), L)
    ph = np.exp(2j * np.pi * np.outer(k, m) / nf)  # (nf, 2L-1)
    K = np.einsum("abc,ai,bj,ck->ijk", lhs, ph, ph, ph) / nf**3  # K[m0, m1, m2], m = -(L-1)..L-1
    y = np.zeros_like(x)
    for i0 in range(L):
        for i1 in range(L):
            for i2 in range(L):
                y[i0, i1, i2] = np.sum(K[i0 - np.arange(L) + L - 1][:, i1 - np.arange(L) + L - 1][:, :, i2 - np.arange(L) + L - 1] * x)
    assert rel_l2(o.toeplitz_apply(x, lhs, L), y) < 1e-12


def test_mlp_forward_matches_torch():
    o = oracle()
    n_in, hid = 3, 64
    rng = np.random.default_rng(23)
    nparam = hid * n_in + hid + 2 * (hid * hid + hid) + hid + 1
    p = rng.standard_normal(nparam) * 0.2
    x = rng.standard_normal((10, n_in))
    W1, b1, W2, b2, W3, b3, W4, b4 = (torch.tensor(np.atleast_1d(t)) for t in o.mlp_unpack(p, n_in))
    xt = torch.tensor(x)
    g = lambda a: torch.nn.functional.gelu(a, approximate="tanh")
    ref = g(g(g(xt @ W1.T + b1) @ W2.T + b2) @ W3.T + b3) @ W4 + b4
    assert np.allclose(o.mlp_forward(p, x), ref.numpy(), atol=1e-12)
