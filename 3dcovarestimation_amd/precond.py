"""Learned diagonal preconditioner for the mean-volume CG.

The CG operator of ``mean.correct_upsampling_padding`` is the Toeplitz
normal operator A = crop_L F^-1 diag(lhs) F pad_nf.  In the L^3 Fourier basis
its diagonal is exactly

    T_k = (lhs (*) |D|^2)(us*k) / nf^3,

with D the Fourier transform of the L-box on the nf grid; by the convolution
theorem this is one probe of the Toeplitz kernel ifft3c(lhs) against the
box autocorrelation (a separable triangle), i.e. two nf^3 FFTs
(``probe_diagonal``).  The Jacobi weights 1/T_k are the training targets.

``DiagonalMLP`` (in -> 64 -> 64 -> 64 -> 1, tanh-GELU, gfx950 forward and
backward kernels ``cov3d_mlp_forward`` / ``cov3d_mlp_backward``) predicts
log T_k from cheap per-voxel / per-shell features:

    [log(sampling density x CTF^2 at us*k), log(shell mean of the same),
     log(noise variance), |k| / (L/2)]

so the preconditioner needs neither the probe nor extra operator
applications.  The preconditioner M^-1 r = ifft3c(w * fft3c(r)) with w > 0
is linear and Hermitian positive-definite, so PCG stays valid; plain CG is
the fallback (``mean.complex_conjugate_gradient``) and judges the final
residual.
"""

from __future__ import annotations

import math
from typing import Callable, Optional, Sequence

import numpy as np
import torch

from . import _native
from .fsc_utils import average_fourier_shell, expand_fourier_shell
from .projection_funcs import centered_fft3, centered_ifft3

HID = 64


def n_params(n_in: int) -> int:
    return HID * n_in + HID + 2 * (HID * HID + HID) + HID + 1


def mlp_torch_reference(params: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """Differentiable torch restatement of the kernel's MLP (used as a test reference)."""
    n_in = x.shape[1]
    o = 0
    W1 = params[o : o + HID * n_in].reshape(HID, n_in); o += HID * n_in
    b1 = params[o : o + HID]; o += HID
    W2 = params[o : o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = params[o : o + HID]; o += HID
    W3 = params[o : o + HID * HID].reshape(HID, HID); o += HID * HID
    b3 = params[o : o + HID]; o += HID
    W4 = params[o : o + HID]; o += HID
    b4 = params[o]
    g = lambda a: torch.nn.functional.gelu(a, approximate="tanh")
    return g(g(g(x @ W1.T + b1) @ W2.T + b2) @ W3.T + b3) @ W4 + b4


class DiagonalMLP:
    """Parameters on the device; forward/backward run on the gfx950 kernels."""

    def __init__(self, n_in: int = 4, device="cuda", seed: int = 0, params: Optional[torch.Tensor] = None):
        self.n_in = n_in
        self.device = torch.device(device)
        if params is None:
            g = torch.Generator().manual_seed(seed)
            p = []
            for fan_in, fan_out in ((n_in, HID), (HID, HID), (HID, HID)):
                p.append(torch.randn((fan_out, fan_in), generator=g).reshape(-1) * math.sqrt(2.0 / fan_in))
                p.append(torch.zeros(fan_out))
            p.append(torch.randn(HID, generator=g) * math.sqrt(1.0 / HID))
            p.append(torch.zeros(1))
            params = torch.cat(p)
        self.params = params.to(device=self.device, dtype=torch.float32).contiguous()
        assert self.params.numel() == n_params(n_in)
        # feature normalisation (set by fit)
        self.x_mean = torch.zeros(n_in, device=self.device)
        self.x_std = torch.ones(n_in, device=self.device)
        self.y_mean = 0.0
        self.y_std = 1.0

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        _native.require_device(x, "features")
        x = x.to(torch.float32).contiguous()
        y = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        _native.check(_native.load().cov3d_mlp_forward(self.params.data_ptr(), x.data_ptr(), y.data_ptr(),
                                                        x.shape[0], self.n_in, _native.stream_ptr(x.device)))
        return y

    def backward(self, x: torch.Tensor, gy: torch.Tensor) -> torch.Tensor:
        x = x.to(torch.float32).contiguous()
        gy = gy.to(torch.float32).contiguous()
        grad = torch.zeros_like(self.params)
        _native.check(_native.load().cov3d_mlp_backward(self.params.data_ptr(), x.data_ptr(), gy.data_ptr(),
                                                         grad.data_ptr(), x.shape[0], self.n_in,
                                                         _native.stream_ptr(x.device)))
        return grad

    # --- normalised prediction of log T_k
    def predict_log_diag(self, feats: torch.Tensor) -> torch.Tensor:
        xn = (feats - self.x_mean) / self.x_std
        return self.forward(xn) * self.y_std + self.y_mean

    def fit(self, feats: torch.Tensor, targets: torch.Tensor, steps: int = 2000, lr: float = 2e-3,
            batch: int = 8192, seed: int = 0, log: Optional[list] = None) -> float:
        """Adam on the mean squared error of log T_k, gradients from the native backward."""
        self.x_mean = feats.mean(0)
        self.x_std = feats.std(0).clamp(min=1e-6)
        self.y_mean = float(targets.mean())
        self.y_std = float(targets.std().clamp(min=1e-6))
        xn = ((feats - self.x_mean) / self.x_std).contiguous()
        yn = ((targets - self.y_mean) / self.y_std).contiguous()
        m = torch.zeros_like(self.params)
        v = torch.zeros_like(self.params)
        b1, b2, eps = 0.9, 0.999, 1e-8
        g = torch.Generator(device=feats.device).manual_seed(seed)
        n = xn.shape[0]
        loss = float("nan")
        for t in range(1, steps + 1):
            idx = torch.randint(0, n, (min(batch, n),), device=feats.device, generator=g)
            xb, yb = xn[idx], yn[idx]
            pred = self.forward(xb)
            err = pred - yb
            grad = self.backward(xb, 2.0 * err / xb.shape[0])
            m.mul_(b1).add_(grad, alpha=1 - b1)
            v.mul_(b2).addcmul_(grad, grad, value=1 - b2)
            mh = m / (1 - b1**t)
            vh = v / (1 - b2**t)
            self.params.sub_(lr * mh / (vh.sqrt() + eps))
            if log is not None and (t % 100 == 0 or t == steps):
                loss = float((err**2).mean())
                log.append(loss)
        return loss

    def state_dict(self) -> dict:
        return {"n_in": self.n_in, "params": self.params.cpu(), "x_mean": self.x_mean.cpu(),
                "x_std": self.x_std.cpu(), "y_mean": self.y_mean, "y_std": self.y_std}

    @classmethod
    def from_state_dict(cls, sd: dict, device="cuda") -> "DiagonalMLP":
        m = cls(sd["n_in"], device=device, params=sd["params"])
        m.x_mean = sd["x_mean"].to(m.device)
        m.x_std = sd["x_std"].to(m.device)
        m.y_mean = float(sd["y_mean"])
        m.y_std = float(sd["y_std"])
        return m

    def make_preconditioner(self, lhs: torch.Tensor, L: int, noise_var: Optional[float]) -> Optional[Callable]:
        """Returns r -> ifft3c(w * fft3c(r)) with w = exp(-predicted log T), or None if w is invalid."""
        feats = diagonal_features(lhs, L, 1.0 if noise_var is None else noise_var)
        logd = self.predict_log_diag(feats)
        return preconditioner_from_log_diag(logd, L)


def preconditioner_from_log_diag(logd: torch.Tensor, L: int) -> Optional[Callable]:
    w = torch.exp(-logd).reshape(L, L, L)
    if not bool(torch.isfinite(w).all()) or bool((w <= 0).any()):
        return None
    w = w.to(torch.complex64)

    def apply(r: torch.Tensor) -> torch.Tensor:
        return centered_ifft3(w * centered_fft3(r.reshape(L, L, L))).reshape(-1)

    return apply


def _box_autocorr(L: int, nf: int, device) -> torch.Tensor:
    """Separable triangle (L - |n|)_+ on the centred nf grid: autocorrelation of the L-box."""
    n = torch.arange(nf, device=device, dtype=torch.float64) - nf // 2
    t = torch.clamp(L - n.abs(), min=0)
    return t[:, None, None] * t[None, :, None] * t[None, None, :]


def probe_diagonal(lhs: torch.Tensor, L: int) -> torch.Tensor:
    """Exact diagonal T_k of the Toeplitz CG operator in the L^3 Fourier basis (flattened L^3, real)."""
    nf = lhs.shape[-1]
    us = nf // L
    kern = centered_ifft3(lhs.to(torch.complex128))  # Toeplitz kernel
    conv = centered_fft3(kern * _box_autocorr(L, nf, lhs.device))
    # frequency k on the L grid sits at index us*k + nf//2 of the centred nf grid
    k = torch.arange(L, device=lhs.device) - L // 2
    idx = us * k + nf // 2
    T = conv[idx][:, idx][:, :, idx].real / (L**3)
    return T.reshape(-1).to(torch.float32)


def diagonal_features(lhs: torch.Tensor, L: int, noise_var: float) -> torch.Tensor:
    """(L^3, 4) features: log density*CTF^2 at us*k, log shell mean, log noise variance, radius."""
    nf = lhs.shape[-1]
    us = nf // L
    k = torch.arange(L, device=lhs.device) - L // 2
    idx = us * k + nf // 2
    dens = lhs.abs()[idx][:, idx][:, :, idx].float().clamp(min=1e-12)
    shell = expand_fourier_shell(average_fourier_shell(dens), L, 3).clamp(min=1e-12)
    ax = torch.ceil(torch.arange(-L / 2, L / 2, device=lhs.device, dtype=torch.float32))[:L]
    zz, yy, xx = torch.meshgrid(ax, ax, ax, indexing="ij")
    rad = torch.sqrt(zz**2 + yy**2 + xx**2) / (L / 2)
    nv = torch.full_like(dens, math.log(max(noise_var, 1e-30)))
    return torch.stack([dens.log(), shell.log(), nv, rad], dim=-1).reshape(-1, 4)
