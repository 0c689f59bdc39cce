"""Projection operators of the COV3D hot path (cov3d/projection_funcs.py).

``vol_forward`` / ``im_backward`` keep the reference signatures and scaling
(projection_funcs.py:321-399): P = mask * CTF * S(V) / L in the Fourier
domain, Re(ifft2c(.)) in the spatial domain, and the exact adjoint.  The
centred FFT helpers keep the pad/crop index conventions of :11-165.
"""

from __future__ import annotations

import math
from typing import List, Optional, Tuple, Union

import torch

from .nufft_plan import BaseNufftPlan, NufftSpec, nufft_adjoint, nufft_forward


def _dims_for(tensor: torch.Tensor, n: int, dims: Optional[List[int]]) -> List[int]:
    if dims is None:
        return [(-1 - i) % tensor.ndim for i in range(n)]
    return [d % tensor.ndim for d in dims]


def pad_tensor(tensor: torch.Tensor, size: List[int], dims: Optional[List[int]] = None) -> torch.Tensor:
    """Zero-pads ``dims`` to ``size`` keeping index floor(n/2) at floor(size/2) (projection_funcs.py:11-38)."""
    dims = _dims_for(tensor, len(size), dims)
    shape = list(tensor.shape)
    for d, s in zip(dims, size):
        shape[d] = s
    out = torch.zeros(shape, dtype=tensor.dtype, device=tensor.device)
    idx = [slice(None)] * tensor.ndim
    for d, s in zip(dims, size):
        start = s // 2 - tensor.shape[d] // 2
        idx[d] = slice(start, start + tensor.shape[d])
    out[tuple(idx)] = tensor
    return out


def crop_tensor(tensor: torch.Tensor, size: List[int], dims: Optional[List[int]] = None) -> torch.Tensor:
    """Inverse of ``pad_tensor`` (projection_funcs.py:41-64)."""
    dims = _dims_for(tensor, len(size), dims)
    idx = [slice(None)] * tensor.ndim
    for d, s in zip(dims, size):
        start = tensor.shape[d] // 2 - s // 2
        idx[d] = slice(start, start + s)
    return tensor[tuple(idx)]


def crop_image(images: torch.Tensor, L_crop: int) -> torch.Tensor:
    """Centred crop of the last two axes (projection_funcs.py:67-79)."""
    L = images.shape[-1]
    idx = torch.arange(-(L_crop // 2), L_crop // 2 + L_crop % 2, device=images.device) + L // 2
    return images[..., idx, :][..., idx].reshape(*images.shape[:-2], L_crop, L_crop)


_SIGNS = {}


def _checkerboard(shape, dims, device, dtype) -> Tuple[torch.Tensor, torch.Tensor]:
    """(-1)^(sum of indices over dims), broadcastable over the other axes, and that times (-1)^(sum N/2)."""
    key = (tuple(shape), tuple(dims), str(device), dtype)
    if key not in _SIGNS:
        bshape = [1] * len(shape)
        s = torch.ones(bshape, dtype=dtype, device=device)
        half = 0
        for d in dims:
            n = shape[d]
            v = torch.ones(n, dtype=dtype, device=device)
            v[1::2] = -1
            vs = [1] * len(shape)
            vs[d] = n
            s = s * v.reshape(vs)
            half += n // 2
        _SIGNS[key] = (s, s * (-1.0 if half % 2 else 1.0))
    return _SIGNS[key]


def _centered(fn, tensor: torch.Tensor, dim: List[int], size: Optional[List[int]] = None) -> torch.Tensor:
    if size is not None:
        tensor = pad_tensor(tensor, list(size), dim)
    dims = [d % tensor.ndim for d in dim]
    if all(tensor.shape[d] % 2 == 0 for d in dims):
        # even sizes: fftshift(F(ifftshift(x)))[k] = (-1)^(k + N/2) F((-1)^m x)[k] per axis, for the forward
        # and the inverse transform alike, so the two shifts become sign flips
        real = tensor.real.dtype if tensor.is_complex() else tensor.dtype
        s_in, s_out = _checkerboard(tensor.shape, dims, tensor.device, real)
        return fn(tensor * s_in, dim=dim) * s_out
    return torch.fft.fftshift(fn(torch.fft.ifftshift(tensor, dim=dim), dim=dim), dim=dim)


def centered_fft2(image: torch.Tensor, im_dim: List[int] = [-1, -2], padding_size=None) -> torch.Tensor:
    return _centered(torch.fft.fft2, image, im_dim, padding_size)


class _CentredPadFFT3Real(torch.autograd.Function):
    """fftshift(fftn(ifftshift(pad_nf(v * w0 w1 w2)))) of real volumes v (r, L, L, L) float32 on the
    pruned line FFTs (``cov3d_fft3c_pad_real``); the backward is the native adjoint, the real part
    of the centred unnormalised inverse transform, cropped and weighted
    (``cov3d_fft3c_pad_real_adjoint``)."""

    @staticmethod
    def forward(ctx, v, nf, ip):
        from . import _native

        lib = _native.load()
        r, L = v.shape[0], v.shape[-1]
        vd = _native.dense(v.detach())
        out = torch.empty((r, nf, nf, nf), dtype=torch.complex64, device=v.device)
        work = torch.empty(int(lib.cov3d_fft3c_pad_real_work(L, nf, r)), dtype=torch.complex64, device=v.device)
        _native.check(lib.cov3d_fft3c_pad_real(vd.data_ptr(), out.data_ptr(), work.data_ptr(), ip.data_ptr(), r, L,
                                               nf, _native.stream_ptr(v.device)))
        ctx.save_for_backward(ip)
        ctx.meta = (r, L, nf)
        return out

    @staticmethod
    def backward(ctx, g):
        from . import _native

        (ip,) = ctx.saved_tensors
        r, L, nf = ctx.meta
        lib = _native.load()
        gd = _native.dense(g.to(torch.complex64))
        grad = torch.empty((r, L, L, L), dtype=torch.float32, device=g.device)
        work = torch.empty(int(lib.cov3d_fft3c_pad_real_work(L, nf, r)), dtype=torch.complex64, device=g.device)
        _native.check(lib.cov3d_fft3c_pad_real_adjoint(gd.data_ptr(), grad.data_ptr(), work.data_ptr(),
                                                       ip.data_ptr(), r, L, nf, _native.stream_ptr(g.device)))
        return grad, None, None


def fft3c_pad_real_supported(v: torch.Tensor, nf: int) -> bool:
    """Whether ``fft3c_pad_real`` runs natively for these volumes: real float32 (r, L, L, L) on a HIP
    device, nf in {64, 128, 256, 512} and 2L <= nf + 2."""
    if not (v.is_cuda and v.dtype == torch.float32 and v.ndim == 4 and v.shape[1] == v.shape[2] == v.shape[3]):
        return False
    L = v.shape[-1]
    return nf in (64, 128, 256, 512) and 2 * L <= nf + 2


def fft3c_pad_real(v: torch.Tensor, nf: int, axis_weights: Optional[torch.Tensor] = None) -> torch.Tensor:
    """centered_fft3(v * w[:, None, None] * w[None, :, None] * w[None, None, :], padding_size=(nf,) * 3)
    for real volumes v (r, L, L, L) (projection_funcs.py:115-128 with the padding of :11-38; the
    Fourier-domain eigenvolumes of covar.py:440-446, w = 1 / the per-axis grid correction).  Runs on
    the native pruned transforms when ``fft3c_pad_real_supported``, on torch's FFT otherwise."""
    L = v.shape[-1]
    if fft3c_pad_real_supported(v, nf):
        if axis_weights is None:
            ip = torch.ones(L, dtype=torch.float32, device=v.device)
        else:
            ip = axis_weights.to(device=v.device, dtype=torch.float32).contiguous()
        return _CentredPadFFT3Real.apply(v, int(nf), ip)
    if axis_weights is not None:
        w = axis_weights.to(device=v.device, dtype=v.dtype)
        v = v * (w[:, None, None] * w[None, :, None] * w[None, None, :])
    return centered_fft3(v, padding_size=(nf,) * 3)


def centered_ifft2(image: torch.Tensor, im_dim: List[int] = [-1, -2], cropping_size=None) -> torch.Tensor:
    t = _centered(torch.fft.ifft2, image, im_dim)
    return crop_tensor(t, cropping_size, im_dim) if cropping_size is not None else t


def centered_fft3(image: torch.Tensor, im_dim: List[int] = [-1, -2, -3], padding_size=None) -> torch.Tensor:
    return _centered(torch.fft.fftn, image, im_dim, padding_size)


def centered_ifft3(image: torch.Tensor, im_dim: List[int] = [-1, -2, -3], cropping_size=None) -> torch.Tensor:
    t = _centered(torch.fft.ifftn, image, im_dim)
    return crop_tensor(t, cropping_size, im_dim) if cropping_size is not None else t


def make_nufft_plan(spec: NufftSpec, input: Optional[torch.Tensor] = None):
    """Builds a plan from a spec; transforms ``input`` to the plan's domain (projection_funcs.py:168-196)."""
    plan = spec._cls()(**spec.get_nufft_kwargs())
    if input is None:
        return plan
    if spec.input_in_fourier:
        input = centered_fft3(input, padding_size=tuple(v * spec.upsample_factor for v in input.shape[-3:]))
    return plan, input


def lowpass_volume(volume: torch.Tensor, cutoff: float, lowpass_shape: str = "rect") -> torch.Tensor:
    """Fourier low-pass (projection_funcs.py:272-294)."""
    return _pass_volume(volume, cutoff, lowpass_shape, low=True)


def highpass_volume(volume: torch.Tensor, cutoff: float, highpass_shape: str = "rect") -> torch.Tensor:
    """Fourier high-pass (projection_funcs.py:297-318)."""
    return _pass_volume(volume, cutoff, highpass_shape, low=False)


def _pass_volume(volume, cutoff, shape, low):
    fv = centered_fft3(volume)
    L = volume.shape[-1]
    ax = torch.arange(-(L // 2), (L + 1) // 2, device=volume.device, dtype=torch.float32)
    if shape == "rect":
        m1 = (ax.abs() < cutoff) if low else (ax.abs() > cutoff)
        mask = m1[:, None, None] & m1[None, :, None] & m1[None, None, :]
    else:
        r = torch.sqrt(ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2)
        mask = (r < cutoff) if low else (r > cutoff)
    return centered_ifft3(fv * mask.unsqueeze(0)).real


def nyquist_mask(L: int, device, dtype=torch.float32) -> torch.Tensor:
    """1 everywhere except row 0 and column 0 for even L (projection_funcs.py:355-358)."""
    m = torch.ones((L, L), device=device, dtype=dtype)
    if L % 2 == 0:
        m[0, :] = 0
        m[:, 0] = 0
    return m


def vol_forward(
    volume: torch.Tensor,
    plan: Union[BaseNufftPlan, List[BaseNufftPlan]],
    filters: Optional[torch.Tensor] = None,
    fourier_domain: bool = False,
) -> torch.Tensor:
    """Projects volumes to images (projection_funcs.py:321-370).

    volume (r, L, L, L) real (NufftPlan) or (r, vL, vL, vL) centred FFT
    (NufftPlanDiscretized); returns (n, r, L, L), squeezed to (r, L, L) for a
    single image, spatial (real) unless ``fourier_domain``.
    """
    if isinstance(plan, (list, tuple)):
        L = plan[0].sz[-1]
        out = torch.zeros((len(plan), volume.shape[0], L, L), dtype=volume.dtype, device=volume.device)
        for i, p in enumerate(plan):
            out[i] = vol_forward(volume, p, filters[i] if filters is not None else None)
        return out
    L = plan.sz[-1]
    proj = nufft_forward(volume, plan)
    proj = proj.reshape((*volume.shape[:-3], -1, proj.shape[-2], proj.shape[-1])).transpose(0, 1)
    n = proj.shape[0]
    if proj.shape[-1] % 2 == 0:
        proj = proj * nyquist_mask(proj.shape[-1], proj.device, proj.real.dtype)
    if filters is not None:
        proj = proj * filters.unsqueeze(1)
    if n == 1:
        proj = proj.squeeze(0)
    out = centered_ifft2(proj).real if not fourier_domain else proj
    return out / L


def im_backward(
    image: torch.Tensor, plan: BaseNufftPlan, filters: Optional[torch.Tensor] = None, fourier_domain: bool = False
) -> torch.Tensor:
    """Back-projects images to a volume, the adjoint of ``vol_forward`` (projection_funcs.py:373-399)."""
    L = image.shape[-1]
    im_fft = centered_fft2(image / L**2) if not fourier_domain else image.clone()
    if filters is not None:
        im_fft = im_fft * filters
    if L % 2 == 0:
        im_fft = im_fft * nyquist_mask(L, im_fft.device, im_fft.real.dtype)
    back = nufft_adjoint(im_fft, plan)
    return back.real / L if not fourier_domain else back / L


def preprocess_image_batch(
    images: torch.Tensor,
    nufft_plan: BaseNufftPlan,
    filters: torch.Tensor,
    pose: Tuple[torch.Tensor, torch.Tensor],
    mean_volume: torch.Tensor,
    mask: Optional[torch.Tensor] = None,
    mask_threshold: Optional[float] = None,
    softening_kernel_fourier: Optional[torch.Tensor] = None,
    fourier_domain: bool = False,
) -> torch.Tensor:
    """Phase-shifts images, subtracts the projected mean and applies a soft mask (projection_funcs.py:199-252)."""
    pts_rot, phase_shift = pose
    nufft_plan.setpts(pts_rot)
    if not fourier_domain:
        images = centered_fft2(images)
    images = images * phase_shift
    mean_forward = vol_forward(mean_volume, nufft_plan, filters=filters, fourier_domain=True).squeeze(1)
    images = images - mean_forward
    if mask is not None:
        images = centered_ifft2(images).real
        with torch.no_grad():
            mproj = vol_forward(mask, nufft_plan, filters=None, fourier_domain=False).squeeze(1)
            mproj = (mproj > mask_threshold).to(images.dtype)
            soft = centered_ifft2(centered_fft2(mproj) * softening_kernel_fourier.to(mproj.device)).real.to(images.dtype)
        images = images * soft
        if fourier_domain:
            images = centered_fft2(images)
    elif not fourier_domain:
        images = centered_ifft2(images).real
    return images


def get_mask_threshold(mask: torch.Tensor, nufft_plan: BaseNufftPlan) -> float:
    """10th percentile (linear interpolation, as numpy's default) of the projected mask values
    above 10^-1.5 (projection_funcs.py:255-269); computed on the device."""
    vals = vol_forward(mask, nufft_plan).reshape(-1)
    vals = vals[vals > 10 ** (-1.5)]
    if vals.numel() == 0:
        raise ValueError("the projected mask has no value above 10^-1.5")
    s = torch.sort(vals.double()).values
    pos = 0.1 * (s.numel() - 1)
    lo = int(math.floor(pos))
    hi = min(lo + 1, s.numel() - 1)
    return float(s[lo] + (s[hi] - s[lo]) * (pos - lo))

