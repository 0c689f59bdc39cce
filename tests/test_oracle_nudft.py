"""Pins the FP64 direct-NUDFT oracle (CPU only).

* brute force: explicit exp matrix on tiny inputs;
* adjointness: <type2 f, c> = <f, type1 c>;
* closed form: a single mode is a plane wave; the identity-rotation slice of
  a volume is the central plane of its centred 3-D FFT - the reference's
  known-answer test tests/test_torch_wraps.py:266-278 (pins sign, mode range
  and axis order).
"""

import numpy as np
import pytest

from tests.conftest import oracle, rel_l2


@pytest.mark.parametrize("N", [(4, 5, 6), (7, 7, 7), (8, 6, 5)])
def test_type2_matches_bruteforce(N):
    o = oracle()
    rng = np.random.default_rng(0)
    f = rng.standard_normal((2, *N)) + 1j * rng.standard_normal((2, *N))
    pts = rng.uniform(-np.pi, np.pi, (3, 50))
    assert rel_l2(o.nudft_type2(f, pts), o.nudft_type2_bruteforce(f, pts)) < 1e-12


def test_adjointness():
    o = oracle()
    rng = np.random.default_rng(1)
    N = (6, 5, 7)
    f = rng.standard_normal((3, *N)) + 1j * rng.standard_normal((3, *N))
    c = rng.standard_normal((3, 40)) + 1j * rng.standard_normal((3, 40))
    pts = rng.uniform(-np.pi, np.pi, (3, 40))
    lhs = np.vdot(o.nudft_type2(f, pts), c)
    rhs = np.vdot(f, o.nudft_type1(c, pts, N))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_single_mode_is_plane_wave():
    o = oracle()
    N = (6, 7, 8)
    f = np.zeros((1, *N), dtype=complex)
    n = (2, -3, 1)
    f[0, n[0] + N[0] // 2, n[1] + N[1] // 2, n[2] + N[2] // 2] = 1.0
    pts = np.random.default_rng(2).uniform(-np.pi, np.pi, (3, 30))
    c = o.nudft_type2(f, pts)[0]
    expect = np.exp(-1j * (n[0] * pts[0] + n[1] * pts[1] + n[2] * pts[2]))
    assert np.allclose(c, expect, atol=1e-12)


@pytest.mark.parametrize("L", [8, 9])
def test_identity_rotation_is_central_slice(L):
    """fft2c(vol_forward(V)) * L equals fft3c(V)[L//2] up to the Nyquist mask."""
    o = oracle()
    rng = np.random.default_rng(3)
    V = rng.standard_normal((1, L, L, L))
    pts = o.pts_rot(np.eye(3)[None], L)
    proj = o.vol_forward_nudft(V, pts)[0, 0]
    lhs = o.fft2c(proj) * L
    rhs = o.fft3c(V[0])[L // 2] * o.nyquist_mask(L)
    assert np.allclose(lhs, rhs, atol=1e-8 * np.abs(rhs).max())


def test_vol_forward_im_backward_adjoint():
    """<A v, y> = <v, A* y> (the reference's dot-product test, test_torch_wraps.py:374-424)."""
    o = oracle()
    rng = np.random.default_rng(4)
    L = 6
    R = np.linalg.qr(rng.standard_normal((3, 3, 3)))[0]
    pts = o.pts_rot(R, L)
    V = rng.standard_normal((1, L, L, L))
    y = rng.standard_normal((3, L, L))
    F = rng.uniform(0.2, 1.0, (3, L, L))
    F = 0.5 * (F + F[:, ::-1, ::-1])
    lhs = np.sum(o.vol_forward_nudft(V, pts, F)[:, 0] * y)
    rhs = np.sum(V[0] * o.im_backward_nudft(y, pts, F))
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


@pytest.mark.parametrize("N", [8, 9, 10])
@pytest.mark.parametrize("iflag", [-1, 1])
def test_lattice_closed_forms_match_direct_sums(N, iflag):
    """The FFT closed forms for points on a shifted integer lattice, p = 2 pi (k + delta) / N (used
    at sizes where direct sums are too slow), against the direct sums of both types and the
    explicit exp matrix; plus the mode-subset type 1 against the full one."""
    o = oracle()
    rng = np.random.default_rng(N + 10 * (iflag + 1))
    f = rng.standard_normal((2, N, N, N)) + 1j * rng.standard_normal((2, N, N, N))
    k = rng.integers(-(N // 2), N - N // 2, (3, 60))
    k[:, :5] = k[:, 5:10]  # repeated lattice sites accumulate in type 1
    delta = rng.uniform(0, 1, 3)
    P = o.lattice_points(k, delta, N)
    ref2 = o.nudft_type2_bruteforce(f, P, iflag)
    assert rel_l2(o.lattice_type2(f, k, delta, iflag), ref2) < 1e-12
    assert rel_l2(o.nudft_type2(f, P, iflag), ref2) < 1e-12
    c = rng.standard_normal((2, 60)) + 1j * rng.standard_normal((2, 60))
    ref1 = o.nudft_type1(c, P, (N, N, N), iflag)
    assert rel_l2(o.lattice_type1(c, k, delta, N, iflag), ref1) < 1e-12
    modes = rng.choice(N**3, 23, replace=False)
    assert rel_l2(o.nudft_type1_modes(c, P, (N, N, N), modes, iflag), ref1.reshape(2, -1)[:, modes]) < 1e-12
