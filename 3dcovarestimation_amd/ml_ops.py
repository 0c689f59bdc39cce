"""Fused maximum-likelihood contractions and their hand-written adjoint.

``ml_forward`` / ``ml_grad_images`` wrap the gfx950 kernels
``cov3d_ml_forward`` / ``cov3d_ml_grad_images``.  ``MLCostNufft`` is the
chosen form of the spatial ML objective of cov3d/covar_sgd.py:1238-1286 on the
NUFFT path: its forward runs the type-2 NUFFT and the per-image Gram + FP64
Cholesky; its backward forms the CTF-weighted residual-times-coefficient
gradient images and back-projects them with one type-1 NUFFT, replacing
autograd through vol_forward, matmul, solve and logdet.

Equivalence with the reference's spatial-domain cost: with
X_t = mask*CTF*c_t/L (the Fourier-domain projection) the spatial projection
is Re(ifft2c(X_t)); by Parseval its inner products equal the real inner
products of Q_t = X_t / L and y~ = fft2c(y) / L on the Hermitian part of X_t.
The anti-Hermitian part of X_t is the NUFFT's own O(eps) asymmetry for a real
volume; it enters the Gram only at O(eps^2) and is removed from the
gradient by the final real part.
"""

from __future__ import annotations

from typing import Optional, Tuple

import torch

from . import _native

PRECISION_FP32 = 0
PRECISION_BF16X3 = 1


def ml_forward(
    c: torch.Tensor,
    yhat: torch.Tensor,
    L: int,
    noise_var: float,
    ctf: Optional[torch.Tensor] = None,
    lam_inv: Optional[torch.Tensor] = None,
    mean_const: bool = False,
    precision: int = PRECISION_BF16X3,
    want_m: bool = False,
    want_minv: bool = True,
    want_b: bool = False,
):
    """Per-image ML terms from raw type-2 values.

    c: (r, B, L*L) complex64 raw NUFFT values; yhat: (B, L, L) complex64
    centred FFT of the images; ctf: (B, L, L) float32 or None.
    Returns dict with cost (B,), z (B, r) and optionally m (B, r, r),
    minv (B, r, r), b (B, r).
    """
    r, B = c.shape[0], c.shape[1]
    dev = c.device
    _native.require_device(c, "c")
    c = c.contiguous()
    yhat = yhat.to(torch.complex64).contiguous()
    ctf_t = ctf.to(torch.float32).contiguous() if ctf is not None else None
    lam_t = lam_inv.to(device=dev, dtype=torch.float32).contiguous() if lam_inv is not None else None
    out = {
        "cost": torch.empty(B, dtype=torch.float32, device=dev),
        "z": torch.empty((B, r), dtype=torch.float32, device=dev),
    }
    if want_m:
        out["m"] = torch.empty((B, r, r), dtype=torch.float32, device=dev)
    if want_minv:
        out["minv"] = torch.empty((B, r, r), dtype=torch.float32, device=dev)
    if want_b:
        out["b"] = torch.empty((B, r), dtype=torch.float32, device=dev)
    lib = _native.load()
    _native.check(
        lib.cov3d_ml_forward(
            c.data_ptr(), _native.ptr(ctf_t), yhat.data_ptr(), _native.ptr(lam_t), B, r, L,
            float(noise_var), 1 if mean_const else 0, int(precision),
            out["cost"].data_ptr(), out["z"].data_ptr(), _native.ptr(out.get("m")),
            _native.ptr(out.get("minv")), _native.ptr(out.get("b")), _native.stream_ptr(dev),
        )
    )
    return out


def ml_grad_images(
    c: torch.Tensor,
    yhat: torch.Tensor,
    z: torch.Tensor,
    minv: torch.Tensor,
    L: int,
    noise_var: float,
    grad_scale: float,
    ctf: Optional[torch.Tensor] = None,
    out: Optional[torch.Tensor] = None,
) -> torch.Tensor:
    """Gradient of the summed ML cost with respect to the raw type-2 values, times grad_scale."""
    r, B = c.shape[0], c.shape[1]
    g = out if out is not None else torch.empty_like(c)
    ctf_t = ctf.to(torch.float32).contiguous() if ctf is not None else None
    lib = _native.load()
    _native.check(
        lib.cov3d_ml_grad_images(
            c.data_ptr(), _native.ptr(ctf_t), yhat.contiguous().data_ptr(), z.contiguous().data_ptr(),
            minv.contiguous().data_ptr(), g.data_ptr(), B, r, L, float(noise_var), float(grad_scale),
            _native.stream_ptr(c.device),
        )
    )
    return g


class MLCostNufft(torch.autograd.Function):
    """Batch-mean spatial ML cost of rank-r real eigenvolumes on a NufftPlan.

    forward(vols (r, L, L, L) f32, yhat (B, L, L) c64, plan, ctf (B, L, L) or
    None, noise_var, mean_const, precision) -> scalar cost
    = 0.5 * mean_i(-b_i.z_i + logdet M_i) [+ |y_i|^2/noise_var].
    The plan's points must stay set until backward has run.
    """

    @staticmethod
    def forward(ctx, vols, yhat, plan, ctf, noise_var, mean_const, precision):
        r = vols.shape[0]
        L = vols.shape[-1]
        B = yhat.shape[0]
        c = plan.execute_forward(vols.detach())  # (batch_size, B, L, L) complex
        c = c[:r].reshape(r, B, L * L)
        res = ml_forward(c, yhat, L, noise_var, ctf=ctf, mean_const=mean_const, precision=precision)
        ctx.plan = plan
        ctx.meta = (L, B, float(noise_var))
        ctx.save_for_backward(c, yhat, res["z"], res["minv"], ctf if ctf is not None else torch.empty(0))
        ctx.has_ctf = ctf is not None
        return res["cost"].mean()

    @staticmethod
    def backward(ctx, grad_out):
        c, yhat, z, minv, ctf = ctx.saved_tensors
        L, B, noise_var = ctx.meta
        g = ml_grad_images(
            c, yhat, z, minv, L, noise_var, 1.0 / B, ctf=ctf if ctx.has_ctf else None, out=c
        )
        grad_v = ctx.plan.execute_adjoint(g.reshape(g.shape[0], -1), real_output=True)
        grad_v = grad_v[: g.shape[0]] * grad_out
        return grad_v, None, None, None, None, None, None
