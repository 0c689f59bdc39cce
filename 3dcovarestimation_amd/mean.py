"""Mean-volume reconstruction and its Toeplitz CG (cov3d/mean.py:27-291, :359-421).

``reconstruct_mean`` back-projects images and CTF^2 with the discretised
nearest-neighbour adjoint (us=2) into (us*L)^3 accumulators, divides, and
corrects the zero-padding with CG on
    A x = crop_L(ifft3c(lhs * fft3c(pad_{us*L}(x)))).
The operator runs in ``cov3d_toeplitz_apply`` (rocFFT + fused pad/multiply/
crop kernels); the CG inner products accumulate in FP64 (``cov3d_vdot_c64_f64``)
and alpha/beta are formed in FP64 on the device.

``complex_conjugate_gradient`` accepts an optional diagonal preconditioner in
the L^3 Fourier basis (see ``precond``); it stays linear and Hermitian
positive-definite so PCG remains valid.  A preconditioner that is not
finite and positive is refused and plain CG runs instead; plain CG also
restarts from the PCG iterate if PCG stagnates, and the final residual is
always recomputed from A and b.
"""

from __future__ import annotations

import logging
import math
from dataclasses import dataclass, field
from typing import Any, Callable, List, Optional, Tuple, Union

import torch

from . import _native
from .covar import sinc_grid_correction
from .fsc_utils import average_fourier_shell, expand_fourier_shell, upsample_and_expand_fourier_shell, vol_fsc
from .nufft_plan import NufftPlanDiscretized
from .projection_funcs import centered_fft3, centered_ifft3, im_backward

logger = logging.getLogger(__name__)


def get_grid_correction(L: int, upsampling_factor: int, nufft_disc: str) -> torch.Tensor:
    """mean.py:27-40."""
    return sinc_grid_correction(L, upsampling_factor, nufft_disc)


def vdot_f64(x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """sum(conj(x) * y) of complex64 vectors, accumulated in FP64; returns a complex128 device scalar."""
    _native.require_device(x, "x")
    x = _native.dense(x.reshape(-1).to(torch.complex64))
    y = _native.dense(y.reshape(-1).to(torch.complex64))
    out = torch.empty(2, dtype=torch.float64, device=x.device)
    _native.check(_native.load().cov3d_vdot_c64_f64(x.data_ptr(), y.data_ptr(), x.numel(), out.data_ptr(),
                                                     _native.stream_ptr(x.device)))
    return torch.complex(out[0], out[1])


def toeplitz_apply(x: torch.Tensor, lhs: torch.Tensor, L: int) -> torch.Tensor:
    """crop_L(ifft3c(lhs * fft3c(pad_nf(x)))) for x with L^3 complex entries (mean.py:118-123)."""
    _native.require_device(x, "x")
    nf = lhs.shape[-1]
    x = _native.dense(x.reshape(-1).to(torch.complex64))
    lhs = _native.dense(lhs.to(torch.complex64))
    y = torch.empty(L**3, dtype=torch.complex64, device=x.device)
    _native.check(_native.load().cov3d_toeplitz_apply(x.data_ptr(), lhs.data_ptr(), y.data_ptr(), L, nf,
                                                       _native.stream_ptr(x.device)))
    return y


@dataclass
class CGInfo:
    iterations: int = 0  # first iteration whose residual met the tolerance (the reference's count)
    iterations_run: int = 0  # iterations executed (up to check_every - 1 more)
    residuals: List[float] = field(default_factory=list)
    preconditioned: bool = False
    fallback: Optional[str] = None
    final_residual: float = float("nan")


def complex_conjugate_gradient(A_mv: Callable, b: torch.Tensor, x0: Optional[torch.Tensor] = None,
                               rtol: float = 1e-6, maxiter: Optional[int] = None,
                               precond: Optional[Callable] = None, return_info: bool = False,
                               stall_window: int = 1000, check_every: int = 16):
    """(Preconditioned) CG for Hermitian positive-definite A (mean.py:43-93).

    Stops when the recursively updated residual satisfies ||r|| < rtol ||b||, as the reference
    does; the true residual ||b - A x|| / ||b|| is recomputed at the end
    (``CGInfo.final_residual``).  alpha and beta are FP64 device scalars and stay on the device:
    each iteration is the operator, one FP64 dot product (pAp), ``cov3d_cg_update`` (x, r and
    ||r||^2 in one pass) and ``cov3d_cg_direction`` (plus the preconditioner and one more dot with
    ``precond``).  The residual norms of a block of ``check_every`` iterations are read back at its
    end, so the host never waits for the device inside a block and the solve may run up to
    ``check_every - 1`` iterations past the first one that meets the tolerance
    (``CGInfo.residuals`` records every iteration).  If ``precond`` stagnates (the best residual
    improves by less than 10% over ``stall_window`` iterations) plain CG resumes from the current
    iterate.
    """
    if not b.is_complex():
        raise ValueError("Input vector b must be a complex tensor.")
    _native.require_device(b, "b")
    info = CGInfo(preconditioned=precond is not None)
    dev = b.device
    lib = _native.load()
    st = _native.stream_ptr(dev)
    b = _native.dense(b.reshape(-1).to(torch.complex64))
    n = b.numel()
    x = torch.zeros_like(b) if x0 is None else _native.dense(x0.reshape(-1).to(torch.complex64)).clone()
    maxiter = n * 10 if maxiter is None else maxiter
    # FP64 device scalars: [rz0 (2) | pAp (2) | rr of the block's iterations (2 each) | rz of them (2 each)]
    S = torch.zeros(4 + 4 * check_every, dtype=torch.float64, device=dev)
    base = S.data_ptr()
    RZ0, PAP = base, base + 16
    rr_slot = lambda k: base + 32 + 16 * k
    rz_slot = lambda k: base + 32 + 16 * check_every + 16 * k

    def dot_into(u, v, ptr):
        _native.check(lib.cov3d_vdot_c64_f64(u.data_ptr(), v.data_ptr(), n, ptr, st))

    b_norm = math.sqrt(vdot_f64(b, b).real.item())
    r = _native.dense(b - A_mv(x).reshape(-1))
    z = _native.dense(precond(r).reshape(-1)) if precond is not None else r
    p = z.clone()
    dot_into(r, z, RZ0)
    rz = RZ0
    info.residuals.append(math.sqrt(vdot_f64(r, r).real.item()) / b_norm)
    best = [info.residuals[0]]  # best residual so far, per iteration
    done = info.residuals[0] < rtol
    k = 0  # iterations of the current block
    it = 0
    while not done and it < maxiter:
        Ap = _native.dense(A_mv(p).reshape(-1))
        dot_into(p, Ap, PAP)
        _native.check(lib.cov3d_cg_update(x.data_ptr(), r.data_ptr(), p.data_ptr(), Ap.data_ptr(), n, rz, PAP,
                                          rr_slot(k), st))
        if precond is not None:
            z = _native.dense(precond(r).reshape(-1))
            dot_into(r, z, rz_slot(k))
            rz_new = rz_slot(k)
        else:
            z = r
            rz_new = rr_slot(k)
        _native.check(lib.cov3d_cg_direction(p.data_ptr(), z.data_ptr(), n, rz_new, rz, st))
        rz = rz_new
        k += 1
        it += 1
        if k == check_every or it >= maxiter:
            for v in S[4:4 + 2 * k:2].cpu().tolist():
                info.residuals.append(math.sqrt(max(v, 0.0)) / b_norm)
                best.append(min(best[-1], info.residuals[-1]))
                done = done or info.residuals[-1] < rtol
            k = 0
            if not done and precond is not None and it > stall_window and best[-1] > 0.9 * best[-1 - stall_window]:
                info.fallback = "stall"
                precond = None
                z = r
                p = r.clone()
                dot_into(r, r, RZ0)
                rz = RZ0
    info.iterations_run = it
    info.iterations = next((i for i, v in enumerate(info.residuals) if v < rtol), it)
    res = b - A_mv(x).reshape(-1)
    info.final_residual = math.sqrt(vdot_f64(res, res).real.item()) / b_norm
    logger.debug(f"CG converged in {info.iterations} iterations")
    return (x, info) if return_info else x


def correct_upsampling_padding(lhs: torch.Tensor, rhs: torch.Tensor, cg_init: torch.Tensor, L: int,
                               upsampling_factor: int, precond: Optional[Callable] = None,
                               rtol: float = 1e-6, return_info: bool = False, maxiter: Optional[int] = None):
    """Solves the padded normal equations by CG from the direct estimate (mean.py:96-132)."""
    b = centered_ifft3(rhs, cropping_size=(L,) * 3).reshape(-1).to(torch.complex64)
    x0 = torch.complex(cg_init.reshape(-1).float(), torch.zeros_like(cg_init.reshape(-1).float()))
    # the operator reads lhs as dense complex64: convert it once, not once per iteration
    lhs = _native.dense(lhs.to(torch.complex64))
    x, info = complex_conjugate_gradient(lambda v: toeplitz_apply(v, lhs, L), b, x0=x0, rtol=rtol,
                                         precond=precond, return_info=True, maxiter=maxiter)
    vol = x.reshape((L, L, L)).real
    return (vol, info) if return_info else vol


def regularize_lhs(lhs: torch.Tensor) -> torch.Tensor:
    """max(|lhs|, 1e-3 * shell mean), clamped at 1e-8 (mean.py:135-142)."""
    reg = average_fourier_shell(lhs.real) * 1e-3
    reg = expand_fourier_shell(reg, lhs.shape[-1], 3)
    return torch.clamp(torch.max(lhs.abs(), reg), min=1e-8)


def _sphere_mask(L: int, device) -> torch.Tensor:
    ax = torch.ceil(torch.arange(-L / 2, L / 2, dtype=torch.float64))[:L] / (L / 2)
    zz, yy, xx = torch.meshgrid(ax, ax, ax, indexing="ij")
    return ((zz**2 + yy**2 + xx**2).sqrt() <= 1).to(device)


def backproject_lhs_rhs(dataset, idx=None, upsampling_factor: int = 2, batch_size: int = 1024):
    """rhs = sum im_backward(images), lhs = sum im_backward(CTF)/L, both (us*L)^3 (mean.py:185-216).

    Each image is first multiplied by the phase ramp of its negated stored offset
    (mean.py:200-204), which undoes the particle's in-plane translation; contrasts scale the CTF.
    """
    from .dataset import create_dataloader
    from .poses import get_phase_shift_grid, offset_to_phase_shift

    L = dataset.resolution
    plan = NufftPlanDiscretized((L,) * 3, upsample_factor=upsampling_factor, mode="nearest")
    was_spatial = dataset._in_spatial_domain
    dataset.to_fourier_domain()
    dev = dataset.device
    vL = L * upsampling_factor
    rhs = torch.zeros((vL,) * 3, dtype=torch.complex64, device=dev)
    lhs = torch.zeros((vL,) * 3, dtype=torch.complex64, device=dev)
    grid = get_phase_shift_grid(L, dtype=dataset.dtype, device=dev)
    for images, pts_rot, filters, bidx in create_dataloader(dataset, batch_size, idx):
        filters = filters * dataset.contrasts[bidx].reshape(-1, 1, 1)
        offsets = dataset.offsets[bidx].to(device=dev, dtype=dataset.dtype).reshape(-1, 2)
        images = images * offset_to_phase_shift(-offsets, phase_shift_grid=grid).reshape(images.shape)
        plan.setpts(pts_rot)
        rhs += im_backward(images, plan, filters, fourier_domain=True)[0]
        lhs += im_backward(torch.complex(filters, torch.zeros_like(filters)), plan, filters, fourier_domain=True)[0]
    lhs /= L
    if was_spatial:
        dataset.to_spatial_domain()
    return lhs, rhs


def reconstruct_mean(dataset, init_vol: Optional[torch.Tensor] = None, mask: Optional[torch.Tensor] = None,
                     upsampling_factor: int = 2, batch_size: int = 1024, idx: Optional[torch.Tensor] = None,
                     do_grid_correction: Optional[bool] = True, return_lhs_rhs: bool = False,
                     precond_model=None):
    """Mean volume from back-projections plus the CG padding correction (mean.py:145-259)."""
    L = dataset.resolution
    lhs, rhs = backproject_lhs_rhs(dataset, idx, upsampling_factor, batch_size)
    if init_vol is not None:
        from .fsc_utils import rpsd

        reg = dataset.noise_var / upsample_and_expand_fourier_shell(rpsd(init_vol.squeeze(0)), L * upsampling_factor, 3)
        reg_lhs = lhs + reg / L
    else:
        reg_lhs = regularize_lhs(lhs)
    mean_volume = centered_ifft3(rhs / reg_lhs, cropping_size=(L,) * 3).real
    if upsampling_factor > 1:
        precond = precond_model.make_preconditioner(reg_lhs, L, dataset.noise_var) if precond_model else None
        mean_volume = correct_upsampling_padding(reg_lhs, rhs, mean_volume, L, upsampling_factor, precond=precond)
    mean_volume = centered_ifft3(centered_fft3(mean_volume) * _sphere_mask(L, mean_volume.device)).real
    if do_grid_correction:
        mean_volume = mean_volume / get_grid_correction(L, upsampling_factor, "nearest").to(mean_volume.device)
    if mask is not None:
        mean_volume = mean_volume * mask.squeeze(0)
    return (mean_volume, rhs, lhs) if return_lhs_rhs else mean_volume


def regularize_mean_from_halfsets(mean_half1, rhs1, lhs1, mean_half2, rhs2, lhs2, mask=None,
                                  do_grid_correction=True, precond_model=None):
    """FSC-derived shell regulariser, then the same CG (mean.py:359-421)."""
    L = mean_half1.shape[-1]
    us = lhs1.shape[-1] // L
    gain = centered_fft3(centered_ifft3(lhs1 + lhs2, cropping_size=(L,) * 3)).abs() / 2
    avg_gain = average_fourier_shell(1 / gain)
    fsc = vol_fsc(mean_half1, mean_half2).clamp(1e-3, 1 - 1e-3)
    reg = 1 / ((fsc / (1 - fsc)) * avg_gain)
    reg = upsample_and_expand_fourier_shell(reg.unsqueeze(0), rhs1.shape[-1], 3)
    reg_lhs = regularize_lhs(lhs1 + lhs2 + reg)
    vol = centered_ifft3((rhs1 + rhs2) / reg_lhs, cropping_size=(L,) * 3).real
    if us > 1:
        precond = precond_model.make_preconditioner(reg_lhs, L, None) if precond_model else None
        vol = correct_upsampling_padding(reg_lhs, rhs1 + rhs2, vol, L, us, precond=precond)
    vol = centered_ifft3(centered_fft3(vol) * _sphere_mask(L, vol.device)).real
    if do_grid_correction:
        vol = vol / get_grid_correction(L, us, "nearest").to(vol.device)
    if mask is not None:
        vol = vol * mask.squeeze(0)
    return vol


def reconstruct_mean_from_halfsets(dataset, idx: Optional[torch.Tensor] = None, **kw):
    """Two half reconstructions then FSC regularisation (mean.py:262-291)."""
    kw["return_lhs_rhs"] = True
    n = len(dataset)
    dev = dataset.device
    if idx is None:
        idx = torch.cat([torch.arange(0, n, 2, device=dev), torch.arange(1, n, 2, device=dev)])
    m1, rhs1, lhs1 = reconstruct_mean(dataset, idx=idx[: len(idx) // 2], **kw)
    m2, rhs2, lhs2 = reconstruct_mean(dataset, idx=idx[len(idx) // 2 :], **kw)
    return regularize_mean_from_halfsets(m1, rhs1, lhs1, m2, rhs2, lhs2, mask=kw.get("mask"),
                                         do_grid_correction=kw.get("do_grid_correction", True),
                                         precond_model=kw.get("precond_model"))
