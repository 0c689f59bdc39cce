"""Low-rank covariance parameterisation (cov3d/covar.py:9-458).

``Covar`` stores normalised eigenvector directions and log sqrt-eigenvalues;
its forward returns V = vectors * exp(log_sqrt_eigenvals) in the spatial
domain, or fft3c(pad_{us*L}(V / grid_correction)) in the Fourier domain, as in
covar.py:440-449.
"""

from __future__ import annotations

from typing import Any, Dict, List, Optional, Tuple, Union

import torch

from .projection_funcs import centered_fft3, centered_ifft3, fft3c_pad_real, fft3c_pad_real_supported


def sinc_grid_correction(L: int, upsampling_factor: int, nufft_disc: str) -> torch.Tensor:
    """Separable sinc (nearest) or sinc^2 (bilinear) correction (covar.py:45-65, mean.py:27-40)."""
    assert nufft_disc in ("bilinear", "nearest"), "nufft_disc must be 'bilinear' or 'nearest'"
    pos = torch.arange(-(L // 2), (L - 1) // 2 + 1, dtype=torch.float64) / L
    pos = torch.pi * pos / upsampling_factor
    s = torch.ones_like(pos)
    nz = pos != 0
    s[nz] = torch.sin(pos[nz]) / pos[nz]
    s[s < 1e-6] = 1
    if nufft_disc == "bilinear":
        s = s**2
    s = s.to(torch.float32)
    return s[:, None, None] * s[None, :, None] * s[None, None, :]


class VolumeBase(torch.nn.Module):
    def __init__(self, resolution: int, dtype=torch.float32, fourier_domain: bool = False, upsampling_factor: int = 2):
        super().__init__()
        self.resolution = resolution
        self.dtype = dtype
        self._in_spatial_domain = not fourier_domain
        self.upsampling_factor = upsampling_factor
        self.grid_correction = None

    def init_grid_correction(self, nufft_disc: str) -> None:
        if nufft_disc not in ("bilinear", "nearest"):
            self.grid_correction = None
            return
        self.grid_correction = sinc_grid_correction(self.resolution, self.upsampling_factor, nufft_disc).to(
            self.device
        )

    def to(self, *args, **kwargs):
        super().to(*args, **kwargs)
        if self.grid_correction is not None:
            self.grid_correction = self.grid_correction.to(*args, **kwargs)
        return self

    def _axis_weights(self):
        """1 / the per-axis factor of a separable grid correction (None without one), or False when
        the correction is not separable.  The sinc corrections are s[i] s[j] s[k] with s = 1 at the
        centre, so s is the correction's line through the centre."""
        gc = self.grid_correction
        if gc is None:
            return None
        key = (id(gc), gc.data_ptr(), gc.device, gc.dtype, tuple(gc.shape))
        cached = getattr(self, "_axis_cache", None)
        if cached is not None and cached[0] == key:
            return cached[1]
        c = gc.shape[-1] // 2
        s = gc[:, c, c]
        sep = bool(torch.allclose(s[:, None, None] * s[None, :, None] * s[None, None, :], gc, rtol=1e-6, atol=0))
        w = (1.0 / s.double()).to(torch.float32) if sep else False
        self._axis_cache = (key, w)
        return w

    def _fourier_padded(self, v: torch.Tensor) -> torch.Tensor:
        """centered_fft3(v / grid_correction, padding_size=(us L,) * 3) (covar.py:440-446, :188-199),
        on the native pruned transforms for float32 volumes on the device."""
        nf = self.resolution * self.upsampling_factor
        w = self._axis_weights()
        if w is not False and fft3c_pad_real_supported(v, nf):
            return fft3c_pad_real(v, nf, w)
        if self.grid_correction is not None:
            v = v / self.grid_correction
        return centered_fft3(v, padding_size=(nf,) * 3)

    def state_dict(self, *args, **kwargs):
        sd = super().state_dict(*args, **kwargs)
        sd.update(
            {
                "_in_spatial_domain": self._in_spatial_domain,
                "grid_correction": self.grid_correction.to("cpu") if self.grid_correction is not None else None,
            }
        )
        return sd

    def load_state_dict(self, state_dict, *args, **kwargs):
        state_dict = dict(state_dict)
        self._in_spatial_domain = state_dict.pop("_in_spatial_domain")
        self.grid_correction = state_dict.pop("grid_correction")
        return super().load_state_dict(state_dict, *args, **kwargs)


class _ScaleRows(torch.autograd.Function):
    """v * exp(l) per row (covar.py:45-65's vectors * exp(log_sqrt_eigenvals)).

    The gradient of l is r dot products over whole volumes; they are summed in
    blocks of 4096 first, which spreads the reduction over the device (one
    reduction per row runs on a handful of workgroups).
    """

    @staticmethod
    def forward(ctx, v, l):
        s = torch.exp(l)
        ctx.save_for_backward(v, s)
        return v * s.reshape((-1,) + (1,) * (v.ndim - 1))

    @staticmethod
    def backward(ctx, g):
        v, s = ctx.saved_tensors
        r = v.shape[0]
        sb = s.reshape((-1,) + (1,) * (v.ndim - 1))
        gv = g * sb if ctx.needs_input_grad[0] else None
        gl = None
        if ctx.needs_input_grad[1]:
            prod = (g * v).reshape(r, -1)
            n = prod.shape[1]
            blk = 4096 if n % 4096 == 0 else 1
            gl = prod.reshape(r, n // blk, blk).sum(-1).sum(-1) * s
        return gv, gl


class Covar(VolumeBase):
    """Rank-r covariance Sigma = V^T V with V = vectors * exp(log_sqrt_eigenvals) (covar.py:278-458)."""

    def __init__(
        self,
        resolution: int,
        rank: int,
        dtype: torch.dtype = torch.float32,
        pixel_var_estimate: Union[float, torch.Tensor] = 1,
        fourier_domain: bool = False,
        upsampling_factor: int = 2,
        vectors: Optional[torch.Tensor] = None,
    ) -> None:
        super().__init__(resolution, dtype, fourier_domain, upsampling_factor)
        self.rank = rank
        self.pixel_var_estimate = pixel_var_estimate
        if vectors is None:
            # a scalar pixel variance gives white vectors; a (radial or full) PSD shapes them (covar.py:316-322)
            psd_given = isinstance(pixel_var_estimate, torch.Tensor) and pixel_var_estimate.ndim > 0
            vectors = (self.init_random_vectors_from_psd(rank, pixel_var_estimate) if psd_given
                       else self.init_random_vectors(rank))
        else:
            vectors = vectors.clone()
        v, l = self._split(vectors)
        self.vectors = torch.nn.Parameter(v)
        self.log_sqrt_eigenvals = torch.nn.Parameter(l)

    def init_random_vectors(self, num_vectors: int) -> torch.Tensor:
        """White Gaussian vectors with the estimated pixel variance (covar.py:377-387)."""
        return torch.randn((num_vectors,) + (self.resolution,) * 3, dtype=self.dtype) * (
            float(self.pixel_var_estimate) ** 0.5)

    def init_random_vectors_from_psd(self, num_vectors: int, psd: torch.Tensor) -> torch.Tensor:
        """Gaussian vectors whose spectrum follows ``psd`` (covar.py:389-405): white noise, centred
        FFT scaled by L^-1.5 (so unit white noise has unit spectral density), times sqrt(psd), back
        to space.  A 1-D ``psd`` is radial and expanded over the Fourier shells."""
        from .fsc_utils import expand_fourier_shell

        L = self.resolution
        psd = torch.as_tensor(psd)
        if psd.ndim == 1:
            psd = expand_fourier_shell(psd, L, 3)
        vectors = torch.randn((num_vectors,) + (L,) * 3, dtype=self.dtype, device=psd.device)
        vf = centered_fft3(vectors) / (L**1.5) * torch.sqrt(psd.to(vectors.dtype))
        return centered_ifft3(vf).real.to(self.dtype)

    @staticmethod
    def _split(vectors: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        n = vectors.reshape(vectors.shape[0], -1).norm(dim=1)
        return vectors / n.reshape(-1, 1, 1, 1), torch.log(n)

    @property
    def device(self) -> torch.device:
        return self.vectors.device

    def set_vectors(self, new_vectors: torch.Tensor) -> None:
        v, l = self._split(new_vectors)
        self.vectors.data.copy_(v)
        self.log_sqrt_eigenvals.data.copy_(l)

    def get_vectors_spatial_domain(self) -> torch.Tensor:
        return _ScaleRows.apply(self.vectors, self.log_sqrt_eigenvals)

    def get_vectors_fourier_domain(self) -> torch.Tensor:
        return self._fourier_padded(self.get_vectors_spatial_domain())

    def get_vectors(self) -> torch.Tensor:
        return self.get_vectors_spatial_domain() if self._in_spatial_domain else self.get_vectors_fourier_domain()

    def forward(self, dummy_var: Optional[Any] = None) -> torch.Tensor:
        return self.get_vectors()

    @property
    def eigenvecs(self) -> Tuple[torch.Tensor, torch.Tensor]:
        """Orthonormal eigenvolumes and eigenvalues from an SVD of V (covar.py:407-430)."""
        with torch.no_grad():
            v = self.get_vectors_spatial_domain().reshape(self.rank, -1)
            _, s, vh = torch.linalg.svd(v, full_matrices=False)
            return vh.reshape((self.rank,) + (self.resolution,) * 3), s**2

    def grad_lr_factor(self) -> List[Dict[str, Union[torch.nn.Parameter, float]]]:
        return [{"params": self.vectors, "lr": 1}, {"params": self.log_sqrt_eigenvals, "lr": 100}]

    def orthogonal_projection(self) -> None:
        with torch.no_grad():
            v = self.get_vectors_spatial_domain().reshape(self.rank, -1)
            _, s, vh = torch.linalg.svd(v, full_matrices=False)
            self.set_vectors((s.reshape(-1, 1) * vh).reshape((self.rank,) + (self.resolution,) * 3))


class Mean(VolumeBase):
    """Mean volume as a normalised volume and its log amplitude (covar.py:115-235)."""

    def __init__(self, volume_init: torch.Tensor, resolution: int, dtype: torch.dtype = torch.float32,
                 fourier_domain: bool = False, volume_mask: Optional[torch.Tensor] = None,
                 upsampling_factor: int = 2) -> None:
        super().__init__(resolution, dtype, fourier_domain, upsampling_factor)
        v, l = self._split(volume_init.reshape(1, resolution, resolution, resolution).to(dtype))
        self.volume = torch.nn.Parameter(v)
        self.log_volume_amplitude = torch.nn.Parameter(l)
        self.volume_mask = volume_mask

    @staticmethod
    def _split(volume: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        a = volume.reshape(volume.shape[0], -1).norm(dim=1)
        return volume / a.reshape(-1, 1, 1, 1), torch.log(a)

    def set_mean(self, volume: torch.Tensor) -> None:
        v, l = self._split(volume.reshape(1, *volume.shape[-3:]).to(self.volume.dtype))
        self.volume.data.copy_(v)
        self.log_volume_amplitude.data.copy_(l)

    @property
    def device(self) -> torch.device:
        return self.volume.device

    def to(self, *args, **kwargs):
        super().to(*args, **kwargs)
        if self.volume_mask is not None:
            self.volume_mask = self.volume_mask.to(*args, **kwargs)
        return self

    def get_volume_spatial_domain(self) -> torch.Tensor:
        return self.volume * torch.exp(self.log_volume_amplitude).reshape(-1, 1, 1, 1)

    def get_volume_fourier_domain(self) -> torch.Tensor:
        return self._fourier_padded(self.get_volume_spatial_domain())

    def forward(self, dummy_var: Optional[Any] = None) -> torch.Tensor:
        return self.get_volume_spatial_domain() if self._in_spatial_domain else self.get_volume_fourier_domain()

    def get_volume_mask(self) -> Optional[torch.Tensor]:
        if self.volume_mask is None:
            return None
        if self._in_spatial_domain:
            return self.volume_mask
        m = self.volume_mask / self.grid_correction if self.grid_correction is not None else self.volume_mask
        return centered_fft3(m, padding_size=(self.resolution * self.upsampling_factor,) * 3)
