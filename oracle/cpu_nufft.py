"""CPU port of one covariance-SGD step on the NUFFT path, FP32 -- TEST / BASELINE ONLY.

Timed by bench.py's ``cpu_baseline`` leg on the host cores and pinned by tests/test_cpu_port.py
against the FP64 NUDFT and the FP64 ML objective.  It is the same method as the GPU path, on
the CPU: ES-kernel type-2 / type-1 NUFFT (w = ceil(log10(1/eps)) + 1, beta = 2.30 w, upsampling 2;
cov3d/nufft_plan.py:263-392 through finufft) with the spreading and interpolation of
cpu_nufft.cpp (C++, OpenMP, FP32, all transforms per stencil weight), the fine-grid FFTs pruned
axis by axis in scipy.fft (complex64, ``workers`` threads), the deconvolution by the kernel's
Fourier transform (Gauss-Legendre quadrature), and the ML objective of
cov3d/covar_sgd.py:1238-1286 with its closed-form gradient in FP32 numpy, on one pixel of every
+-k pair of each image (the Hermitian half plane, as the GPU's chosen form).
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess
from pathlib import Path

import numpy as np
import scipy.fft

from . import oracle as O

_DIR = Path(__file__).resolve().parent
LIB = _DIR / "libcpu_nufft.so"
_lib = None


def build(force: bool = True) -> Path:
    """Compiles cpu_nufft.cpp with g++ (-O3, OpenMP) into libcpu_nufft.so."""
    if LIB.exists() and not force:
        return LIB
    tmp = LIB.with_suffix(".so.tmp")
    subprocess.run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-mavx2", "-mfma",
                    str(_DIR / "cpu_nufft.cpp"), "-o", str(tmp)], check=True)
    os.replace(tmp, LIB)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = ctypes.CDLL(str(LIB))
        vp, i64, ci, cf = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_float
        lib.cpu_es_setpts.argtypes = [ci, ci, i64, vp, vp, vp, vp, vp]
        lib.cpu_es_interp_f32.argtypes = [ci, ci, cf, i64, vp, vp, vp, vp, ci, vp]
        lib.cpu_es_spread_f32.argtypes = [ci, ci, cf, i64, vp, vp, vp, vp, vp, ci, vp, ci]
        lib.cpu_threads.restype = ci
        for fn in (lib.cpu_es_setpts, lib.cpu_es_interp_f32, lib.cpu_es_spread_f32):
            fn.restype = None
        _lib = lib
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data


def es_params(eps: float):
    """Kernel width and beta: w = ceil(log10(1/eps)) + 1 (2 <= w <= 16); beta = c w with c = 2.20,
    2.26, 2.38 for w = 2, 3, 4 and 2.30 from w = 5 on (the published FINUFFT choices at
    upsampling 2)."""
    w = min(max(int(math.ceil(math.log10(1 / eps))) + 1, 2), 16)
    b = {2: 2.20, 3: 2.26, 4: 2.38}.get(w, 2.30)
    return w, b * w


def next235even(n: int) -> int:
    n += n % 2
    while True:
        k = n
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return n
        n += 2


def deconv(N: int, nf: int, w: int, beta: float) -> np.ndarray:
    """1 / (Fourier transform of the kernel) at the N centred modes, by 128-point Gauss-Legendre."""
    x, wt = np.polynomial.legendre.leggauss(128)
    t = 0.5 * w * x
    phi = np.exp(beta * (np.sqrt(np.clip(1 - (2 * t / w) ** 2, 0, None)) - 1))
    n = np.arange(N) - N // 2
    p = (0.5 * w * wt * phi) @ np.cos(2 * np.pi * np.outer(t, n) / nf)
    return 1.0 / p


def default_workers() -> int:
    """The host threads to use: OMP_NUM_THREADS when set, else the CPUs this process may run on."""
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return n if n > 0 else len(os.sched_getaffinity(0))


class CpuNufft:
    """Type 2 / type 1 of ``ntrans`` real L^3 volumes on one point set, FP32."""

    def __init__(self, L: int, ntrans: int, eps: float = 1e-6, workers: int | None = None):
        self.L, self.ntrans = L, ntrans
        self.w, self.beta = es_params(eps)
        self.nf = next235even(max(2 * L, 2 * self.w))
        self.ip = deconv(L, self.nf, self.w, self.beta).astype(np.float32)
        self.workers = workers or default_workers()
        self.idx = (np.arange(L) - L // 2) % self.nf
        _load()

    def setpts(self, pts: np.ndarray) -> None:
        """pts (3, M) FP32 radians."""
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        self.M = pts.shape[1]
        self.base = np.empty((3, self.M), dtype=np.int32)
        self.off = np.empty((3, self.M), dtype=np.float32)
        self.order = np.empty(self.M, dtype=np.int64)
        self.starts = np.empty(self.nf + 1, dtype=np.int64)
        _load().cpu_es_setpts(self.nf, self.w, self.M, _p(pts), _p(self.base), _p(self.off), _p(self.order),
                              _p(self.starts))

    def _fft_axis(self, x, axis, inverse, keep=None, place=None):
        """Length-nf DFT along one axis (sign - or +, unnormalised); ``place`` puts the L inputs at
        their wrapped cells first, ``keep`` returns only the L wrapped outputs."""
        nf, L = self.nf, self.L
        h = L // 2  # mode i sits at cell (i - h) mod nf: modes h.. at cells 0.., modes ..h-1 at the end

        def sl(a, b):
            s = [slice(None)] * x.ndim
            s[axis] = slice(a, b)
            return tuple(s)

        if place is not None:
            shape = list(x.shape)
            shape[axis] = nf
            y = np.zeros(shape, dtype=np.complex64)
            y[sl(0, L - h)] = x[sl(h, L)]
            y[sl(nf - h, nf)] = x[sl(0, h)]
            x = y
        if inverse:
            x = scipy.fft.ifft(x, axis=axis, norm="forward", workers=self.workers, overwrite_x=True)
        else:
            x = scipy.fft.fft(x, axis=axis, workers=self.workers, overwrite_x=True)
        if keep is not None:
            x = np.concatenate([x[sl(nf - h, nf)], x[sl(0, L - h)]], axis=axis)
        return x

    def type2(self, f: np.ndarray) -> np.ndarray:
        """f (r, L, L, L) real -> (M, r) complex64 (iflag -1)."""
        r, L = f.shape[0], self.L
        ip = self.ip
        g = (np.moveaxis(f.astype(np.float32), 0, -1) * (ip[:, None, None, None] * ip[None, :, None, None] *
                                                           ip[None, None, :, None])).astype(np.complex64)
        g = self._fft_axis(g, 2, False, place=True)   # (L, L, nf, r)
        g = self._fft_axis(g, 1, False, place=True)   # (L, nf, nf, r)
        g = self._fft_axis(g, 0, False, place=True)   # (nf, nf, nf, r)
        g = np.ascontiguousarray(g, dtype=np.complex64)
        out = np.empty((self.M, r), dtype=np.complex64)
        _load().cpu_es_interp_f32(self.nf, self.w, self.beta, self.M, _p(self.base), _p(self.off), _p(self.order),
                                  _p(g), r, _p(out))
        return out

    def type1_real(self, c: np.ndarray) -> np.ndarray:
        """c (M, r) complex64 -> real part of the type 1 (iflag +1), (r, L, L, L) float32."""
        r = c.shape[1]
        nf = self.nf
        c = np.ascontiguousarray(c, dtype=np.complex64)
        g = np.empty((nf, nf, nf, r), dtype=np.complex64)
        _load().cpu_es_spread_f32(nf, self.w, self.beta, self.M, _p(self.base), _p(self.off), _p(self.order),
                                  _p(self.starts), _p(c), r, _p(g), 2)
        g = self._fft_axis(g, 0, True, keep=True)    # (L, nf, nf, r)
        g = self._fft_axis(g, 1, True, keep=True)    # (L, L, nf, r)
        g = self._fft_axis(g, 2, True, keep=True)    # (L, L, L, r)
        ip = self.ip
        f = g.real * (ip[:, None, None, None] * ip[None, :, None, None] * ip[None, None, :, None])
        return np.ascontiguousarray(np.moveaxis(f, -1, 0), dtype=np.float32)


def half_plane(L: int):
    """Row-major pixels of one of every +-k pair plus the self-paired centre, and their weights
    (2 for a pair, 1 for the centre); Nyquist row and column (even L) are excluded."""
    c = L // 2
    lo = 1 if L % 2 == 0 else 0
    pix = []
    for i in range(c, L):
        j0 = c if i == c else lo
        pix.extend(i * L + j for j in range(j0, L))
    pix = np.array(pix, dtype=np.int64)
    w = np.full(pix.size, 2.0, dtype=np.float32)
    w[pix == c * L + c] = 1.0
    return pix, w


def sgd_step_cpu(vols: np.ndarray, images: np.ndarray, pts: np.ndarray, ctf: np.ndarray, noise_var: float,
                 nufft: CpuNufft):
    """ML cost (batch mean) and its gradient with respect to the eigenvolumes, FP32.

    vols (r, L, L, L) real, images (B, L, L) real, pts (B, 3, L*L) radians, ctf (B, L, L).
    With Q = mask CTF c / L^2 and y~ = fft2c(y) / L on the half plane (weights h), the real Gram
    sum_k h_k Re(conj Q_s Q_t) equals the spatial P_s . P_t of the reference's objective; the
    gradient images fft2c(G) / L = (M^-1 Q + z (z^T Q - y~)) / (B sigma^2) go back through the
    real part of one type-1 NUFFT.
    """
    r, L = vols.shape[0], vols.shape[-1]
    B = images.shape[0]
    pix, h = half_plane(L)
    npx = pix.size
    P = np.transpose(O.wrap_points(np.asarray(pts, dtype=np.float32)), (1, 0, 2))[:, :, pix]  # (3, B, npx)
    nufft.setpts(P.reshape(3, -1))
    c = nufft.type2(vols).reshape(B, npx, r)
    mh = (ctf.reshape(B, -1)[:, pix] * O.nyquist_mask(L).reshape(-1)[pix]).astype(np.float32)  # (B, npx)
    Q = np.transpose(c, (0, 2, 1)) * (mh[:, None, :] / np.float32(L * L))  # (B, r, npx) complex64
    with scipy.fft.set_workers(nufft.workers):
        yt = (scipy.fft.fftshift(scipy.fft.fft2(scipy.fft.ifftshift(images.astype(np.float32), axes=(-2, -1))),
                                 axes=(-2, -1)).reshape(B, -1)[:, pix] / np.float32(L)).astype(np.complex64)
    sh = np.sqrt(h)
    Qr = np.concatenate([Q.real * sh, Q.imag * sh], axis=-1)  # (B, r, 2 npx): weighted real view
    yr = np.concatenate([yt.real * sh, yt.imag * sh], axis=-1)  # (B, 2 npx)
    s2 = np.float32(noise_var)
    M = np.eye(r, dtype=np.float32)[None] + np.matmul(Qr, np.transpose(Qr, (0, 2, 1))) / s2
    b = np.einsum("brk,bk->br", Qr, yr) / s2
    Lc = np.linalg.cholesky(M)
    z = np.linalg.solve(M, b[..., None])[..., 0]
    logdet = 2 * np.log(np.diagonal(Lc, axis1=1, axis2=2)).sum(axis=1)
    cost = float(np.mean(0.5 * (-np.einsum("br,br->b", b, z) + logdet)))
    Minv = np.linalg.inv(M)
    e = np.einsum("br,brk->bk", z, Q) - yt  # (B, npx) complex
    Gq = (np.matmul(Minv.astype(np.complex64), Q) + z[:, :, None] * e[:, None, :]) / np.float32(B * noise_var)
    # back to the type-1 input: d(cost)/dc = mask CTF / L^2 times the half-plane weights
    gin = Gq * (mh[:, None, :] * h[None, None, :] / np.float32(L * L))
    gv = nufft.type1_real(np.ascontiguousarray(np.transpose(gin, (0, 2, 1)).reshape(-1, r)))
    return cost, gv


def adam_step(params: np.ndarray, grad: np.ndarray, m: np.ndarray, v: np.ndarray, t: int, lr: float,
              b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> None:
    """In-place Adam update (torch.optim.Adam's defaults), FP32."""
    m *= b1
    m += (1 - b1) * grad
    v *= b2
    v += (1 - b2) * grad * grad
    mh = m / (1 - b1**t)
    vh = v / (1 - b2**t)
    params -= lr * mh / (np.sqrt(vh) + eps)
