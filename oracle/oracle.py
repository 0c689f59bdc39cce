"""CPU oracle of the COV3D hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module, and only as the checker (or the timed CPU port);
the product package never does.

Each function restates, in FP64 numpy (or the C direct NUDFT in
``nudft.c``), one piece of the reference's algorithm and cites the reference
file:line it follows.  How each piece is pinned (closed forms, torch's own
grid_sample, brute force, identities) is listed in the tests that exercise it
(tests/test_oracle_*.py).
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB = ORACLE_DIR / "liboracle.so"

_lib = None


def build(force: bool = True) -> Path:
    """Compiles nudft.c with gcc (OpenMP) into liboracle.so."""
    if LIB.exists() and not force:
        return LIB
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-fPIC", "-shared", "-fopenmp", "-std=c11", str(ORACLE_DIR / "nudft.c"), "-o", str(tmp),
           "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = ctypes.CDLL(str(LIB))
        dp = ctypes.POINTER(ctypes.c_double)
        for name in ("oracle_nudft3d_type2", "oracle_nudft3d_type1"):
            fn = getattr(lib, name)
            fn.argtypes = [ctypes.c_int] * 3 + [ctypes.c_int64, dp, dp, dp, dp, ctypes.c_int, ctypes.c_int, dp]
            fn.restype = None
        fn = lib.oracle_nudft3d_type1_modes
        fn.argtypes = [ctypes.c_int] * 3 + [ctypes.c_int64, dp, dp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_int64), dp]
        fn.restype = None
        for name in ("oracle_es_interp", "oracle_es_spread"):
            fn = getattr(lib, name)
            fn.argtypes = [ctypes.c_int] * 3 + [ctypes.c_int64, dp, dp, dp, dp, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_double, dp]
            fn.restype = None
        _lib = lib
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ---------------------------------------------------------------------------
# NUFFT (nufft_plan.py:263-392; finufft conventions: type 2 isign -1, type 1 +1,
# modes n in [-N//2, (N+1)//2 - 1] at index n + N//2, axis k <-> point coordinate k)
# ---------------------------------------------------------------------------


def wrap_points(points: np.ndarray) -> np.ndarray:
    """(remainder(p + pi, 2pi) - pi) as in nufft_plan.py:327, in the points' own dtype."""
    p = np.asarray(points)
    pi = p.dtype.type(np.pi)
    return np.remainder(p + pi, 2 * pi) - pi


def nudft_type2(f: np.ndarray, pts: np.ndarray, iflag: int = -1) -> np.ndarray:
    """f (ntrans, N0, N1, N2) complex, pts (3, M) -> c (ntrans, M) complex128."""
    f = np.ascontiguousarray(f, dtype=np.complex128)
    if f.ndim == 3:
        f = f[None]
    nt, N0, N1, N2 = f.shape
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    M = pts.shape[1]
    c = np.zeros((nt, M), dtype=np.complex128)
    x, y, z = (np.ascontiguousarray(pts[k]) for k in range(3))
    _load().oracle_nudft3d_type2(N0, N1, N2, M, _dp(x), _dp(y), _dp(z), _dp(f.view(np.float64)), nt, iflag,
                                 _dp(c.view(np.float64)))
    return c


def nudft_type1(c: np.ndarray, pts: np.ndarray, N, iflag: int = 1) -> np.ndarray:
    """c (ntrans, M) complex, pts (3, M) -> f (ntrans, N0, N1, N2) complex128."""
    c = np.ascontiguousarray(c, dtype=np.complex128)
    if c.ndim == 1:
        c = c[None]
    nt, M = c.shape
    N0, N1, N2 = N
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    f = np.zeros((nt, N0, N1, N2), dtype=np.complex128)
    x, y, z = (np.ascontiguousarray(pts[k]) for k in range(3))
    _load().oracle_nudft3d_type1(N0, N1, N2, M, _dp(x), _dp(y), _dp(z), _dp(c.view(np.float64)), nt, iflag,
                                 _dp(f.view(np.float64)))
    return f


def nudft_type1_modes(c: np.ndarray, pts: np.ndarray, N, modes: np.ndarray, iflag: int = 1) -> np.ndarray:
    """Type 1 evaluated only at the flat output indices ``modes`` (C order over N):
    c (ntrans, M), pts (3, M) -> (ntrans, len(modes)) complex128; equals
    ``nudft_type1(c, pts, N).reshape(ntrans, -1)[:, modes]``."""
    c = np.ascontiguousarray(c, dtype=np.complex128)
    if c.ndim == 1:
        c = c[None]
    nt, M = c.shape
    N0, N1, N2 = N
    modes = np.ascontiguousarray(modes, dtype=np.int64)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    out = np.zeros((nt, modes.size), dtype=np.complex128)
    x, y, z = (np.ascontiguousarray(pts[k]) for k in range(3))
    _load().oracle_nudft3d_type1_modes(N0, N1, N2, M, _dp(x), _dp(y), _dp(z), _dp(c.view(np.float64)), nt, iflag,
                                       modes.size, modes.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                       _dp(out.view(np.float64)))
    return out


def lattice_points(k: np.ndarray, delta: np.ndarray, N: int) -> np.ndarray:
    """Points 2 pi (k + delta) / N of one shifted integer lattice: k (3, M) integers, delta (3,)."""
    return 2 * np.pi * (np.asarray(k, dtype=np.float64) + np.asarray(delta, dtype=np.float64)[:, None]) / N


def _fft3(x: np.ndarray, inverse: bool) -> np.ndarray:
    """Unnormalised 3-D DFT over the last three axes, sign - (forward) or + (inverse), FP64."""
    import scipy.fft

    if inverse:
        return scipy.fft.ifftn(x, axes=(-3, -2, -1), norm="forward", workers=os.cpu_count())
    return scipy.fft.fftn(x, axes=(-3, -2, -1), workers=os.cpu_count())


def _mode_phase(delta: np.ndarray, N: int, iflag: int) -> np.ndarray:
    """exp(iflag * i * 2 pi n . delta / N) on the centred (N, N, N) mode grid."""
    n = np.arange(N) - N // 2
    ph = [np.exp(1j * iflag * 2 * np.pi * n * float(d) / N) for d in delta]
    return ph[0][:, None, None] * ph[1][None, :, None] * ph[2][None, None, :]


def lattice_type2(f: np.ndarray, k: np.ndarray, delta: np.ndarray, iflag: int = -1) -> np.ndarray:
    """Exact type 2 at the points ``lattice_points(k, delta, N)``, by FFT instead of a direct sum.

    With p = 2 pi (k + delta) / N, sum_n f[n] e^{iflag i n.p} = sum_n U[n] e^{iflag 2 pi i n.k / N},
    U = f * exp(iflag i 2 pi n.delta / N): one length-N 3-D DFT of U (modes centred, so shifted to
    index n mod N first) read at k mod N.  f (ntrans, N, N, N) -> (ntrans, M) complex128.
    """
    f = np.asarray(f, dtype=np.complex128)
    if f.ndim == 3:
        f = f[None]
    N = f.shape[-1]
    U = np.fft.ifftshift(f * _mode_phase(delta, N, iflag)[None], axes=(1, 2, 3))
    X = _fft3(U, inverse=iflag > 0)
    km = np.mod(np.asarray(k, dtype=np.int64), N)
    return X[:, km[0], km[1], km[2]]


def lattice_type1(c: np.ndarray, k: np.ndarray, delta: np.ndarray, N: int, iflag: int = 1) -> np.ndarray:
    """Exact type 1 from the points ``lattice_points(k, delta, N)`` by FFT.

    sum_j c_j e^{iflag i n.p_j} = exp(iflag i 2 pi n.delta / N) sum_m A[m] e^{iflag 2 pi i n.m / N},
    A[m] = sum of c_j over the points with k_j = m (mod N).  c (ntrans, M) -> (ntrans, N, N, N).
    """
    c = np.asarray(c, dtype=np.complex128)
    if c.ndim == 1:
        c = c[None]
    nt = c.shape[0]
    km = np.mod(np.asarray(k, dtype=np.int64), N)
    flat = (km[0] * N + km[1]) * N + km[2]
    A = np.zeros((nt, N**3), dtype=np.complex128)
    for t in range(nt):
        A[t] = np.bincount(flat, weights=c[t].real, minlength=N**3) + 1j * np.bincount(
            flat, weights=c[t].imag, minlength=N**3)
    A = A.reshape(nt, N, N, N)
    X = _fft3(A, inverse=iflag > 0)
    return np.fft.fftshift(X, axes=(1, 2, 3)) * _mode_phase(delta, N, iflag)[None]


def nudft_type2_bruteforce(f: np.ndarray, pts: np.ndarray, iflag: int = -1) -> np.ndarray:
    """Explicit-matrix type 2 (tiny sizes; pins the C sum)."""
    nt, N0, N1, N2 = f.shape
    axes = [np.arange(N) - N // 2 for N in (N0, N1, N2)]
    n0, n1, n2 = np.meshgrid(*axes, indexing="ij")
    modes = np.stack([n0.ravel(), n1.ravel(), n2.ravel()])  # (3, N^3)
    A = np.exp(1j * iflag * (pts.T.astype(np.float64) @ modes))  # (M, N^3)
    return f.reshape(nt, -1) @ A.T


# ---------------------------------------------------------------------------
# Centred FFTs and pad/crop (projection_funcs.py:11-165)
# ---------------------------------------------------------------------------


def pad(x: np.ndarray, size, axes):
    out_shape = list(x.shape)
    for a, s in zip(axes, size):
        out_shape[a] = s
    out = np.zeros(out_shape, dtype=x.dtype)
    idx = [slice(None)] * x.ndim
    for a, s in zip(axes, size):
        st = s // 2 - x.shape[a] // 2
        idx[a] = slice(st, st + x.shape[a])
    out[tuple(idx)] = x
    return out


def crop(x: np.ndarray, size, axes):
    idx = [slice(None)] * x.ndim
    for a, s in zip(axes, size):
        st = x.shape[a] // 2 - s // 2
        idx[a] = slice(st, st + s)
    return x[tuple(idx)]


def fft2c(x):
    ax = (-2, -1)
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x, axes=ax), axes=ax), axes=ax)


def ifft2c(x):
    ax = (-2, -1)
    return np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(x, axes=ax), axes=ax), axes=ax)


def fft3c(x, size=None):
    ax = (-3, -2, -1)
    if size is not None:
        x = pad(x, size, ax)
    return np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x, axes=ax), axes=ax), axes=ax)


def ifft3c(x, crop_size=None):
    ax = (-3, -2, -1)
    y = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(x, axes=ax), axes=ax), axes=ax)
    return crop(y, crop_size, ax) if crop_size is not None else y


def nyquist_mask(L: int) -> np.ndarray:
    m = np.ones((L, L))
    if L % 2 == 0:
        m[0, :] = 0
        m[:, 0] = 0
    return m


# ---------------------------------------------------------------------------
# Rotated slice frequencies (poses.py:197-269, dataset.py:146-164) and CTF
# (aspire CTFFilter formula, dataset.py:117-125)
# ---------------------------------------------------------------------------


def grid2d_axis(L: int) -> np.ndarray:
    """aspire grid_2d normalised axis: ceil(arange(-L/2, L/2)) / (L/2)."""
    return np.ceil(np.arange(-L / 2, L / 2))[:L] / (L / 2)


def pts_rot(rots: np.ndarray, L: int) -> np.ndarray:
    """(B, 3, 3) -> (B, 3, L*L): wrap(flip_zyx(R . pi*[gx, gy, 0])) in float32 like the reference."""
    g = grid2d_axis(L)
    gy, gx = np.meshgrid(g, g, indexing="ij")
    xy = np.pi * np.stack([gx.ravel(), gy.ravel(), np.zeros(L * L)])  # (3, d)
    p = np.einsum("bij,jd->bid", rots.astype(np.float64), xy)[:, ::-1, :]
    return wrap_points(p.astype(np.float32))


def wavelength(kv: float) -> float:
    """Relativistic electron wavelength in Angstrom (voltage in kV)."""
    return 12.2643247 / math.sqrt(kv * 1e3 + 0.978466 * kv * kv)


def ctf(params: np.ndarray, L: int) -> np.ndarray:
    """params (B, 7) [dU, dV, ang, kV, Cs mm, amp contrast, pixel A] -> (B, L, L)."""
    g = grid2d_axis(L)
    gy, gx = np.meshgrid(g, g, indexing="ij")
    out = np.zeros((params.shape[0], L, L))
    for i, (du, dv, ang, kv, cs, alpha, px) in enumerate(np.asarray(params, dtype=np.float64)):
        fx = np.pi * gx / (2 * np.pi * px)
        fy = np.pi * gy / (2 * np.pi * px)
        lam = wavelength(kv)
        nz = (np.abs(fx) > np.finfo(float).eps) | (np.abs(fy) > np.finfo(float).eps)
        defocus = np.zeros_like(fx)
        a = np.arctan2(fy[nz], fx[nz]) - ang
        defocus[nz] = 0.5 * (du + dv) + 0.5 * (du - dv) * np.cos(2 * a)
        r2 = fx**2 + fy**2
        gamma = -np.pi * lam * defocus * r2 + 0.5 * np.pi * (cs * 1e7) * lam**3 * r2**2
        out[i] = np.sqrt(1 - alpha**2) * np.sin(gamma) - alpha * np.cos(gamma)
    return out


# ---------------------------------------------------------------------------
# Discretised slice extraction (nufft_plan.py:94-242): grid_sample 3-D,
# align_corners=True, zero padding; adjoint = scatter with the same weights.
# ---------------------------------------------------------------------------


def discretized_grid(points: np.ndarray, vol_L: int) -> np.ndarray:
    """(n, 3, Lp^2) zyx radians -> (n*Lp^2, 3) xyz grid_sample coordinates (nufft_plan.py:94-118)."""
    g = np.ascontiguousarray(np.transpose(points, (0, 2, 1))[..., ::-1]).reshape(-1, 3).astype(np.float64)
    g = g * ((vol_L / (vol_L - 1)) / np.pi)
    if vol_L % 2 == 0:
        g = g + 1 / (vol_L - 1)
    return g


def _corners(grid: np.ndarray, D: int, H: int, W: int):
    ix = (grid[:, 0] + 1) / 2 * (W - 1)
    iy = (grid[:, 1] + 1) / 2 * (H - 1)
    iz = (grid[:, 2] + 1) / 2 * (D - 1)
    x0, y0, z0 = np.floor(ix), np.floor(iy), np.floor(iz)
    out = []
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                wx = (ix - x0) if dx else (x0 + 1 - ix)
                wy = (iy - y0) if dy else (y0 + 1 - iy)
                wz = (iz - z0) if dz else (z0 + 1 - iz)
                xs, ys, zs = (x0 + dx).astype(np.int64), (y0 + dy).astype(np.int64), (z0 + dz).astype(np.int64)
                ok = (xs >= 0) & (xs < W) & (ys >= 0) & (ys < H) & (zs >= 0) & (zs < D)
                off = np.where(ok, (zs * H + ys) * W + xs, 0)
                out.append((off, wx * wy * wz * ok))
    return out


def _nearest(grid, D, H, W):
    ix = np.rint((grid[:, 0] + 1) / 2 * (W - 1)).astype(np.int64)
    iy = np.rint((grid[:, 1] + 1) / 2 * (H - 1)).astype(np.int64)
    iz = np.rint((grid[:, 2] + 1) / 2 * (D - 1)).astype(np.int64)
    ok = (ix >= 0) & (ix < W) & (iy >= 0) & (iy < H) & (iz >= 0) & (iz < D)
    return [(np.where(ok, (iz * H + iy) * W + ix, 0), ok.astype(np.float64))]


def slice_forward(vol: np.ndarray, grid: np.ndarray, mode: str = "bilinear") -> np.ndarray:
    """vol (n, D, H, W) complex, grid (P, 3) -> (n, P) (nufft_plan.py:120-169)."""
    n, D, H, W = vol.shape
    flat = vol.reshape(n, -1)
    terms = _corners(grid, D, H, W) if mode == "bilinear" else _nearest(grid, D, H, W)
    out = np.zeros((n, grid.shape[0]), dtype=np.complex128)
    for off, w in terms:
        out += flat[:, off] * w
    return out


def slice_adjoint(sig: np.ndarray, grid: np.ndarray, shape, mode: str = "bilinear") -> np.ndarray:
    """sig (n, P) -> (n, D, H, W) scatter-add adjoint of slice_forward (nufft_plan.py:171-200)."""
    D, H, W = shape
    n = sig.shape[0]
    out = np.zeros((n, D * H * W), dtype=np.complex128)
    terms = _corners(grid, D, H, W) if mode == "bilinear" else _nearest(grid, D, H, W)
    for t in range(n):
        for off, w in terms:
            np.add.at(out[t], off, sig[t] * w)
    return out.reshape(n, D, H, W)


# ---------------------------------------------------------------------------
# Projection pair (projection_funcs.py:321-399) on the exact NUDFT
# ---------------------------------------------------------------------------


def vol_forward_nudft(vols: np.ndarray, points: np.ndarray, filters=None, fourier_domain=False) -> np.ndarray:
    """vols (r, L, L, L) real, points (B, 3, L*L) -> (B, r, L, L) as vol_forward on a NufftPlan."""
    r, L = vols.shape[0], vols.shape[-1]
    B = points.shape[0]
    pts = np.transpose(wrap_points(points), (1, 0, 2)).reshape(3, -1).astype(np.float64)
    c = nudft_type2(vols.astype(np.complex128), pts).reshape(r, B, L, L).transpose(1, 0, 2, 3)
    c = c * nyquist_mask(L)
    if filters is not None:
        c = c * filters[:, None]
    out = c if fourier_domain else ifft2c(c).real
    return out / L


def im_backward_nudft(images: np.ndarray, points: np.ndarray, filters=None, fourier_domain=False) -> np.ndarray:
    """images (B, L, L) -> (L, L, L): adjoint of vol_forward_nudft (projection_funcs.py:373-399)."""
    B, L = images.shape[0], images.shape[-1]
    im = images.astype(np.complex128) if fourier_domain else fft2c(images / L**2)
    if filters is not None:
        im = im * filters
    im = im * nyquist_mask(L)
    pts = np.transpose(wrap_points(points), (1, 0, 2)).reshape(3, -1).astype(np.float64)
    v = nudft_type1(im.reshape(1, -1), pts, (L, L, L))[0]
    return (v if fourier_domain else v.real) / L


# ---------------------------------------------------------------------------
# ML objective (covar_sgd.py:1238-1286) and latent MAP (wiener_coords.py:168-203)
# ---------------------------------------------------------------------------


def ml_terms(P: np.ndarray, y: np.ndarray, noise_var: float, lam_inv=None):
    """P (B, r, d) real, y (B, d) -> (cost_i, z, M, b) in FP64."""
    B, r, d = P.shape
    lam_inv = np.eye(r) if lam_inv is None else lam_inv
    M = np.einsum("brd,bsd->brs", P, P) / noise_var + lam_inv
    b = np.einsum("brd,bd->br", P, y) / noise_var
    z = np.linalg.solve(M, b[..., None])[..., 0]
    logdet = np.linalg.slogdet(M)[1]
    cost_i = 0.5 * (-np.einsum("br,br->b", b, z) + logdet)
    return cost_i, z, M, b


def ml_cost_grad(P: np.ndarray, y: np.ndarray, noise_var: float):
    """Batch-mean ML cost and its closed-form gradient wrt P:
    dF/dP_i = (1 / (B sigma^2)) [M_i^{-1} P_i + z_i (P_i^T z_i - y_i)^T]."""
    B = P.shape[0]
    cost_i, z, M, b = ml_terms(P, y, noise_var)
    Minv = np.linalg.inv(M)
    e = np.einsum("brd,br->bd", P, z) - y
    G = (np.einsum("brs,bsd->brd", Minv, P) + z[:, :, None] * e[:, None, :]) / (B * noise_var)
    return cost_i.mean(), G


def ml_cost_naive(P: np.ndarray, y: np.ndarray, noise_var: float) -> np.ndarray:
    """Per-image Gaussian negative log-likelihood (up to constants) with the dense d x d
    covariance C = P^T P + sigma^2 I, the form the Woodbury expression replaces
    (tests/test_torch_impl.py:248-265 in the reference)."""
    B, r, d = P.shape
    out = np.zeros(B)
    for i in range(B):
        C = P[i].T @ P[i] + noise_var * np.eye(d)
        sign, ld = np.linalg.slogdet(C)
        out[i] = 0.5 * (y[i] @ np.linalg.solve(C, y[i]) + ld - d * np.log(noise_var) - y[i] @ y[i] / noise_var)
    return out


# ---------------------------------------------------------------------------
# Mean-volume CG (mean.py:43-132)
# ---------------------------------------------------------------------------


def toeplitz_apply(x: np.ndarray, lhs: np.ndarray, L: int) -> np.ndarray:
    """crop_L(ifft3c(lhs * fft3c(pad_nf(x)))) (mean.py:118-123)."""
    nf = lhs.shape[-1]
    return ifft3c(lhs * fft3c(x.reshape(L, L, L), size=(nf,) * 3), crop_size=(L,) * 3)


def conjugate_gradient(A, b, x0=None, rtol=1e-6, maxiter=None, M=None):
    """Textbook (P)CG for Hermitian positive-definite A (mean.py:43-93); returns (x, iters, residual history)."""
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - A(x)
    z = r if M is None else M * r
    p = z.copy()
    rz = np.vdot(r, z).real
    bn = np.linalg.norm(b)
    hist = [np.linalg.norm(r) / bn]
    maxiter = 10 * b.size if maxiter is None else maxiter
    it = 0
    for it in range(1, maxiter + 1):
        Ap = A(p)
        alpha = rz / np.vdot(p, Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        hist.append(np.linalg.norm(r) / bn)
        if hist[-1] < rtol:
            break
        z = r if M is None else M * r
        rz_new = np.vdot(r, z).real
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, it, hist


# ---------------------------------------------------------------------------
# Fourier shells (fsc_utils.py:9-184)
# ---------------------------------------------------------------------------


def shell_average_loop(x: np.ndarray) -> np.ndarray:
    """Python-loop shell average exactly as FourierShell (fsc_utils.py:63-131)."""
    L, dim = x.shape[0], x.ndim
    ax = np.ceil(np.arange(-L / 2, L / 2))[:L]
    g = np.meshgrid(*([ax] * dim), indexing="ij")
    r = np.sqrt(sum(gg**2 for gg in g))
    n = int(np.ceil(r.max()))
    out = np.zeros(n, dtype=x.dtype)
    for i in range(n):
        lo, hi = i - 0.5, (i + 0.5 if i < n - 1 else i + 1)
        m = (r > lo) & (r < hi)
        out[i] = x[m].sum() / m.sum()
    return out


# ---------------------------------------------------------------------------
# Learned-preconditioner MLP (n_in -> 64 -> 64 -> 64 -> 1, tanh-GELU)
# ---------------------------------------------------------------------------


def gelu(a):
    return 0.5 * a * (1 + np.tanh(0.7978845608028654 * (a + 0.044715 * a**3)))


def mlp_unpack(params: np.ndarray, n_in: int, hid: int = 64):
    o = 0
    W1 = params[o : o + hid * n_in].reshape(hid, n_in); o += hid * n_in
    b1 = params[o : o + hid]; o += hid
    W2 = params[o : o + hid * hid].reshape(hid, hid); o += hid * hid
    b2 = params[o : o + hid]; o += hid
    W3 = params[o : o + hid * hid].reshape(hid, hid); o += hid * hid
    b3 = params[o : o + hid]; o += hid
    W4 = params[o : o + hid]; o += hid
    b4 = params[o]
    return W1, b1, W2, b2, W3, b3, W4, b4


def mlp_forward(params: np.ndarray, x: np.ndarray) -> np.ndarray:
    W1, b1, W2, b2, W3, b3, W4, b4 = mlp_unpack(params.astype(np.float64), x.shape[1])
    h = gelu(x @ W1.T + b1)
    h = gelu(h @ W2.T + b2)
    h = gelu(h @ W3.T + b3)
    return h @ W4 + b4
