"""CPU port of one covariance-SGD step on the NUFFT path -- TEST / BASELINE ONLY.

Used by bench.py's ``cpu_baseline`` leg (timed on the host cores) and by
tests as an independent CPU implementation of the same ES-kernel NUFFT.
Follows the same method as the gfx950 path: ES kernel (w = ceil(log10(1/eps))+1,
beta = 2.30 w, oversampling 2), deconvolution by the kernel's Fourier
transform (Gauss-Legendre quadrature), scipy.fft for the fine-grid FFT
(``workers`` threads), the C spreader/interpolator of nudft.c (OpenMP), and
the ML objective of cov3d/covar_sgd.py:1238-1286 with its closed-form
gradient in numpy.
"""

from __future__ import annotations

import math
import os

import numpy as np
import scipy.fft

from . import oracle as O


def es_params(eps: float):
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
    x, wt = np.polynomial.legendre.leggauss(128)
    t = 0.5 * w * x
    phi = np.exp(beta * (np.sqrt(np.clip(1 - (2 * t / w) ** 2, 0, None)) - 1))
    n = np.arange(N) - N // 2
    p = (0.5 * w * wt * phi) @ np.cos(2 * np.pi * np.outer(t, n) / nf)
    return 1.0 / p


class CpuNufft:
    def __init__(self, L: int, ntrans: int, eps: float = 1e-6, workers: int | None = None):
        self.L, self.ntrans = L, ntrans
        self.w, self.beta = es_params(eps)
        self.nf = next235even(max(2 * L, 2 * self.w))
        self.ip = deconv(L, self.nf, self.w, self.beta)
        self.workers = workers or os.cpu_count()

    def setpts(self, pts: np.ndarray):
        """pts (3, M) radians."""
        X = np.ascontiguousarray(pts, dtype=np.float64) * (self.nf / (2 * np.pi))
        X = np.mod(X, self.nf)
        self.X = [np.ascontiguousarray(X[k]) for k in range(3)]
        self.M = X.shape[1]

    def _modes_to_fine(self, f):
        L, nf = self.L, self.nf
        s = self.ip[:, None, None] * self.ip[None, :, None] * self.ip[None, None, :]
        g = np.zeros((f.shape[0], nf, nf, nf), dtype=np.complex128)
        idx = (np.arange(L) - L // 2) % nf
        g[np.ix_(np.arange(f.shape[0]), idx, idx, idx)] = f * s
        return g

    def type2(self, f: np.ndarray) -> np.ndarray:
        g = self._modes_to_fine(f.astype(np.complex128))
        g = scipy.fft.fftn(g, axes=(1, 2, 3), workers=self.workers)
        c = np.empty((f.shape[0], self.M), dtype=np.complex128)
        nf = self.nf
        O._load().oracle_es_interp(nf, nf, nf, self.M, O._dp(self.X[0]), O._dp(self.X[1]), O._dp(self.X[2]),
                                   O._dp(np.ascontiguousarray(g).view(np.float64)), f.shape[0], self.w, self.beta,
                                   O._dp(c.view(np.float64)))
        return c

    def type1(self, c: np.ndarray) -> np.ndarray:
        nf, L = self.nf, self.L
        c = np.ascontiguousarray(c, dtype=np.complex128)
        g = np.empty((c.shape[0], nf, nf, nf), dtype=np.complex128)
        O._load().oracle_es_spread(nf, nf, nf, self.M, O._dp(self.X[0]), O._dp(self.X[1]), O._dp(self.X[2]),
                                   O._dp(c.view(np.float64)), c.shape[0], self.w, self.beta, O._dp(g.view(np.float64)))
        g = scipy.fft.ifftn(g, axes=(1, 2, 3), workers=self.workers) * nf**3
        idx = (np.arange(L) - L // 2) % nf
        f = g[np.ix_(np.arange(c.shape[0]), idx, idx, idx)]
        s = self.ip[:, None, None] * self.ip[None, :, None] * self.ip[None, None, :]
        return f * s


def sgd_step_cpu(vols: np.ndarray, images: np.ndarray, pts: np.ndarray, ctf: np.ndarray, noise_var: float,
                 nufft: CpuNufft):
    """One ML cost + gradient evaluation on the CPU; returns (cost, grad wrt vols)."""
    r, L = vols.shape[0], vols.shape[-1]
    B = images.shape[0]
    nufft.setpts(np.transpose(O.wrap_points(pts), (1, 0, 2)).reshape(3, -1))
    c = nufft.type2(vols).reshape(r, B, L, L)
    X = np.transpose(c, (1, 0, 2, 3)) * O.nyquist_mask(L) * ctf[:, None] / L
    with scipy.fft.set_workers(nufft.workers):
        P = O.ifft2c(X).real.reshape(B, r, -1)
    cost, G = O.ml_cost_grad(P, images.reshape(B, -1), noise_var)
    # adjoint of c -> Re(ifft2c(mask * ctf * c / L)): mask * ctf * fft2c(G) / (L^2 * L)
    Gf = O.fft2c(G.reshape(B, r, L, L)) * O.nyquist_mask(L) * ctf[:, None] / (L**3)
    gv = nufft.type1(np.transpose(Gf, (1, 0, 2, 3)).reshape(r, -1)).real
    return cost, gv
