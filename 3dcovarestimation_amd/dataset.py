"""Device-resident particle dataset, batch iterator and synthetic generator.

``CovarDataset`` keeps the interface the trainers use from the reference's
cov3d/dataset.py:32-515 (``images``, ``filters``, ``noise_var``,
``resolution``, ``__getitem__`` -> (images, pts_rot, filters, idx),
``to_fourier_domain`` / ``to_spatial_domain``, ``half_split``).  Instead of
materialising pts_rot (N, 3, L^2) on the host (dataset.py:146-164: 19.7 GB at
L=128, N=100k) it stores one rotation matrix per image in HBM and generates
the rotated slice frequencies per batch with ``cov3d_pts_rot``; CTFs can be
stored or generated per batch with ``cov3d_ctf_eval``.

``create_dataloader`` yields contiguous, unshuffled batches like
BatchIndexSampler(shuffle=False) (dataset.py:798-895).
``make_synthetic`` builds the benchmark workloads (Gaussian-blob mean and
eigenvolumes, Haar-uniform SO(3) poses, radial CTF, white noise at an SNR).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native
from .projection_funcs import centered_fft2, centered_ifft2


def pts_rot_from_rotations(rots: torch.Tensor, L: int) -> torch.Tensor:
    """(B, 3, 3) float32 rotations on a HIP device -> (B, 3, L*L) wrapped zyx frequencies."""
    _native.require_device(rots, "rotations")
    rots = rots.to(torch.float32).contiguous()
    out = torch.empty((rots.shape[0], 3, L * L), dtype=torch.float32, device=rots.device)
    _native.check(_native.load().cov3d_pts_rot(rots.data_ptr(), out.data_ptr(), rots.shape[0], L,
                                                _native.stream_ptr(rots.device)))
    return out


def ctf_from_params(params: torch.Tensor, L: int) -> torch.Tensor:
    """(B, 7) [dU, dV, ang, kV, Cs, amp, pixel] on a HIP device -> (B, L, L) CTF."""
    _native.require_device(params, "ctf params")
    params = params.to(torch.float32).contiguous()
    out = torch.empty((params.shape[0], L, L), dtype=torch.float32, device=params.device)
    _native.check(_native.load().cov3d_ctf_eval(params.data_ptr(), out.data_ptr(), params.shape[0], L,
                                                 _native.stream_ptr(params.device)))
    return out


def rotvec_to_rotmat(rotvecs: torch.Tensor) -> torch.Tensor:
    """Rodrigues' formula (poses.py:44-67)."""
    theta = torch.norm(rotvecs, dim=-1, keepdim=True)
    k = rotvecs / (theta + 1e-6)
    K = torch.zeros((rotvecs.shape[0], 3, 3), device=rotvecs.device, dtype=rotvecs.dtype)
    K[:, 0, 1], K[:, 0, 2] = -k[:, 2], k[:, 1]
    K[:, 1, 0], K[:, 1, 2] = k[:, 2], -k[:, 0]
    K[:, 2, 0], K[:, 2, 1] = -k[:, 1], k[:, 0]
    eye = torch.eye(3, device=rotvecs.device, dtype=rotvecs.dtype).unsqueeze(0)
    return eye + torch.sin(theta).unsqueeze(-1) * K + (1 - torch.cos(theta).unsqueeze(-1)) * (K @ K)


def haar_rotations(n: int, generator: torch.Generator, dtype=torch.float32) -> torch.Tensor:
    """Haar-uniform SO(3) rotations from normalised Gaussian quaternions."""
    q = torch.randn((n, 4), generator=generator, dtype=torch.float64)
    q = q / q.norm(dim=1, keepdim=True)
    w, x, y, z = q.unbind(1)
    R = torch.stack(
        [
            1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
            2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
            2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y),
        ],
        dim=1,
    ).reshape(n, 3, 3)
    return R.to(dtype)


class CovarDataset:
    """Particle stack with per-image rotations and CTFs, resident on one device."""

    def __init__(
        self,
        images: torch.Tensor,
        rotations: torch.Tensor,
        noise_var: float,
        filters: Optional[torch.Tensor] = None,
        ctf_params: Optional[torch.Tensor] = None,
        offsets: Optional[torch.Tensor] = None,
        mask: Optional[torch.Tensor] = None,
    ) -> None:
        if filters is None and ctf_params is None:
            raise ValueError("need filters or ctf_params")
        self.images = images
        self.rotations = rotations
        self.filters = filters
        self.ctf_params = ctf_params
        self.noise_var = float(noise_var)
        self.resolution = images.shape[-1]
        self.offsets = offsets if offsets is not None else torch.zeros((len(images), 2), device=rotations.device)
        self.contrasts = torch.ones(len(images), device=rotations.device)
        self.mask = mask
        self._in_spatial_domain = not images.is_complex()
        self.dtype = images.real.dtype if images.is_complex() else images.dtype

    def __len__(self) -> int:
        return self.images.shape[0]

    @property
    def device(self) -> torch.device:
        """The device the batches are computed on (where the rotations live)."""
        return self.rotations.device

    def get_filters(self, idx) -> torch.Tensor:
        if self.filters is not None:
            return self.filters[idx]
        p = self.ctf_params[idx]
        return ctf_from_params(p.reshape(-1, 7), self.resolution).reshape(*p.shape[:-1], self.resolution,
                                                                            self.resolution)

    def __getitem__(self, idx):
        rots = self.rotations[idx]
        single = rots.ndim == 2
        pts = pts_rot_from_rotations(rots.reshape(-1, 3, 3), self.resolution)
        if single:
            pts = pts[0]
        return self.images[idx], pts, self.get_filters(idx), idx

    def to_fourier_domain(self) -> None:
        """images <- fft2c(images), noise_var *= L^2 (dataset.py:426-433)."""
        if self._in_spatial_domain:
            self.images = centered_fft2(self.images)
            self.noise_var *= self.resolution**2
            self._in_spatial_domain = False

    def to_spatial_domain(self) -> None:
        """Inverse of to_fourier_domain (dataset.py:435-439)."""
        if not self._in_spatial_domain:
            self.images = centered_ifft2(self.images).real
            self.noise_var /= self.resolution**2
            self._in_spatial_domain = True

    def preprocess(self, mean_volume: torch.Tensor, mask: Optional[torch.Tensor] = None,
                   batch_size: int = 1024, softening_radius: float = 5.0) -> None:
        """Removes the translations and the projected mean from every image, and applies the
        soft projected mask when ``mask`` is given (dataset.py:232-281, preprocess_from_modules).

        Per batch, on the device: the images' spectra times exp(-i (g_x o_x + g_y o_y)) for the
        stored offsets o (the pose module negates them, poses.py:263-264), minus the CTF-weighted
        projection of ``mean_volume`` (1, L, L, L); with a mask, the residual image times the
        projected mask thresholded at ``get_mask_threshold`` (computed on the first batch's
        poses) and blurred by the soft-edged disc of radius ``softening_radius``.  Offsets are
        zero afterwards.  The images stay in their current domain.
        """
        from .nufft_plan import NufftPlan
        from .poses import get_phase_shift_grid, offset_to_phase_shift
        from .projection_funcs import get_mask_threshold, preprocess_image_batch
        from .utils import soft_edged_kernel

        L = self.resolution
        dev = self.device
        if mean_volume.ndim == 3:
            mean_volume = mean_volume.unsqueeze(0)
        if mask is not None and mask.ndim == 3:
            mask = mask.unsqueeze(0)
        plan = NufftPlan((L,) * 3, batch_size=1, device=dev)
        grid = get_phase_shift_grid(L, dtype=self.dtype, device=dev)
        soft = soft_edged_kernel(radius=softening_radius, L=L, dim=2, in_fourier=True).to(dev)
        soft = soft.to(torch.complex64 if self.dtype == torch.float32 else torch.complex128)
        with torch.no_grad():
            thr = 0.0
            if mask is not None:
                plan.setpts(self[slice(0, min(batch_size, len(self)))][1])
                thr = get_mask_threshold(mask, plan)
            for i in range(0, len(self), batch_size):
                sl = slice(i, min(i + batch_size, len(self)))
                images, pts, filters, _ = self[sl]
                phase = offset_to_phase_shift(-self.offsets[sl].to(self.dtype), phase_shift_grid=grid)
                self._store(sl, preprocess_image_batch(images, plan, filters, (pts, phase), mean_volume, mask,
                                                       thr, soft, fourier_domain=not self._in_spatial_domain))
        self.offsets.zero_()

    def _store(self, idx, images: torch.Tensor) -> None:
        self.images[idx] = images

    def update_pose(self, pose_module, batch_size: int = 16384) -> None:
        """Rotations and offsets (and contrasts) from a PoseModule (dataset.py:503-515)."""
        from .poses import rotvec_to_rotmat

        with torch.no_grad():
            for i in range(0, len(self), batch_size):
                sl = slice(i, min(i + batch_size, len(self)))
                self.rotations[sl] = rotvec_to_rotmat(pose_module.get_rotvecs()[sl]).to(self.rotations.dtype)
                self.offsets[sl] = pose_module.get_offsets()[sl].to(self.offsets.dtype)
                if pose_module.use_contrast:
                    self.contrasts[sl] = pose_module.get_contrasts()[sl].reshape(-1).to(self.contrasts.dtype)

    def get_total_gain(self, batch_size: int = 4096) -> torch.Tensor:
        """diag(sum_i P_i^T P_i) on the (L,)^3 grid (dataset.py:322-356): the nearest-neighbour
        (upsampling 1) back-projection of CTF^2 with the Nyquist row and column masked, summed
        over the stack, |.|, divided by L^2."""
        from .nufft_plan import NufftPlanDiscretized
        from .projection_funcs import im_backward

        L = self.resolution
        plan = NufftPlanDiscretized((L,) * 3, upsample_factor=1, mode="nearest")
        gain = torch.zeros((L,) * 3, device=self.device, dtype=self.dtype)
        for i in range(0, len(self), batch_size):
            _, pts, filters, _ = self[slice(i, min(i + batch_size, len(self)))]
            plan.setpts(pts)
            cf = torch.complex(filters, torch.zeros_like(filters))
            gain += im_backward(cf, plan, filters, fourier_domain=True).squeeze().abs()
        return gain / L

    def get_total_covar_gain(self, batch_size: int = 16384) -> torch.Tensor:
        """Shell x shell gain of the covariance least-squares problem (dataset.py:358-421):
        sum_i a_i a_i^T, a_i the Fourier-shell average of image i's nearest-neighbour
        (upsampling 1) back-projection of CTF_i^2 / L^2 (Nyquist row and column zeroed for even L).

        The reference forms each image's L^3 back-projection (execute_adjoint_unaggregated) and
        averages it over shells.  Every contribution is non-negative, so the shell average is
        the sum of CTF_i^2 / L^2 over the image's pixels whose nearest voxel lies in the shell,
        divided by the shell size: one pass over the pixels, no per-image volume.  The nearest
        voxel is rounded exactly as the slice kernels do (grid_sample, align_corners=True, half
        to even; points outside the volume are dropped).
        """
        from .fsc_utils import shell_index
        from .nufft_plan import NufftPlanDiscretized

        L = self.resolution
        dev = self.device
        shell_of, size, n = shell_index(L, 3, dev)
        plan = NufftPlanDiscretized((L,) * 3, upsample_factor=1, mode="nearest")
        acc = torch.zeros((n, n), device=dev, dtype=torch.float64)
        for i in range(0, len(self), batch_size):
            _, pts, filters, _ = self[slice(i, min(i + batch_size, len(self)))]
            B = filters.shape[0]
            plan.setpts(pts)
            grid = plan._grid_flat.to(device=dev, dtype=torch.float32)
            f2 = (filters.to(torch.float32) ** 2).clone()
            if L % 2 == 0:
                f2[:, 0, :] = 0
                f2[:, :, 0] = 0
            vox = torch.round(((grid + 1) / 2) * float(L - 1)).to(torch.int64)  # (x, y, z) per point
            ok = ((vox >= 0) & (vox < L)).all(dim=1)
            flat = (vox[:, 2] * L + vox[:, 1]) * L + vox[:, 0]
            sh = shell_of[torch.where(ok, flat, torch.zeros_like(flat))]
            img = torch.arange(B, device=dev).repeat_interleave(L * L)
            w = torch.where(ok, f2.reshape(-1).double() / L**2, torch.zeros((), dtype=torch.float64, device=dev))
            a = torch.zeros((B * n,), device=dev, dtype=torch.float64)
            a.index_add_(0, img * n + sh, w)
            a = a.reshape(B, n) / size
            acc += a.T @ a
        return acc.to(self.dtype)

    def _subset(self, idx: torch.Tensor) -> "CovarDataset":
        sub = CovarDataset(
            self.images[idx], self.rotations[idx], self.noise_var,
            filters=None if self.filters is None else self.filters[idx],
            ctf_params=None if self.ctf_params is None else self.ctf_params[idx],
            offsets=self.offsets[idx], mask=self.mask,
        )
        sub._in_spatial_domain = self._in_spatial_domain
        return sub

    def half_split(self, permute: bool = True) -> Tuple["CovarDataset", "CovarDataset", torch.Tensor]:
        """Two half-sets, the images permutation[:n // 2] and permutation[n // 2:] of a random
        permutation (identity order with ``permute=False``), and that permutation
        (dataset.py:302-320)."""
        n = len(self)
        permutation = torch.randperm(n) if permute else torch.arange(n)
        idx = permutation.to(self.device)
        return self._subset(idx[: n // 2]), self._subset(idx[n // 2:]), permutation

    def filter_gain(self, batch_size: int = 4096) -> torch.Tensor:
        """Mean CTF^2 over the stack, (L, L)."""
        acc = None
        for i in range(0, len(self), batch_size):
            f = self.get_filters(slice(i, min(i + batch_size, len(self))))
            s = (f.float() ** 2).sum(dim=0)
            acc = s if acc is None else acc + s
        return acc / len(self)


class StreamingCovarDataset(CovarDataset):
    """A particle stack whose images live in pinned host memory and are streamed to the device.

    Rotations, CTF parameters (or filters), offsets and contrasts stay on the device, so pts_rot
    and the CTF are made there per batch as for ``CovarDataset``; only the images (N, L, L) float32
    travel, 4 L^2 bytes per image (256 KiB at L = 256).  ``create_dataloader`` streams contiguous
    batches through two device buffers on a copy stream, the next batch's copy overlapping the
    current batch's compute (covar_sgd.py:240-244 and dataset.py:850-895 move each batch with a
    pinned DataLoader).  ``to_fourier_domain`` keeps the host images real and applies the centred
    FFT to each batch on the device instead (dataset.py:426-433), scaling noise_var by L^2 as there.
    """

    def __init__(self, host_images: torch.Tensor, rotations: torch.Tensor, noise_var: float,
                 filters: Optional[torch.Tensor] = None, ctf_params: Optional[torch.Tensor] = None,
                 offsets: Optional[torch.Tensor] = None, mask: Optional[torch.Tensor] = None) -> None:
        if host_images.is_cuda or host_images.is_complex():
            raise ValueError("host_images must be real images in host memory")
        if not host_images.is_pinned():
            host_images = host_images.pin_memory()  # a new page-locked copy
        super().__init__(host_images, rotations, noise_var, filters=filters, ctf_params=ctf_params,
                         offsets=offsets, mask=mask)
        self._fourier = False
        self._in_spatial_domain = True
        self._copy_stream = None

    @property
    def copy_stream(self):
        if self._copy_stream is None:
            self._copy_stream = torch.cuda.Stream(self.device)  # default priority
        return self._copy_stream

    def _finish(self, images: torch.Tensor) -> torch.Tensor:
        return centered_fft2(images) if self._fourier else images

    def _host_index(self, idx):
        return idx.cpu() if isinstance(idx, torch.Tensor) else idx

    def __getitem__(self, idx):
        rots = self.rotations[idx]
        single = rots.ndim == 2
        images = self._finish(self.images[self._host_index(idx)].to(self.device, non_blocking=True))
        pts = pts_rot_from_rotations(rots.reshape(-1, 3, 3), self.resolution)
        if single:
            pts = pts[0]
        return images, pts, self.get_filters(idx), idx

    def _store(self, idx, images: torch.Tensor) -> None:
        if images.is_complex():
            images = centered_ifft2(images).real
        self.images[self._host_index(idx)] = images.to("cpu")

    def to_fourier_domain(self) -> None:
        if not self._fourier:
            self._fourier = True
            self._in_spatial_domain = False
            self.noise_var *= self.resolution**2

    def to_spatial_domain(self) -> None:
        if self._fourier:
            self._fourier = False
            self._in_spatial_domain = True
            self.noise_var /= self.resolution**2

    def _subset(self, idx: torch.Tensor) -> "StreamingCovarDataset":
        sub = StreamingCovarDataset(
            self.images[self._host_index(idx)], self.rotations[idx], self.noise_var,
            filters=None if self.filters is None else self.filters[idx],
            ctf_params=None if self.ctf_params is None else self.ctf_params[idx],
            offsets=self.offsets[idx], mask=self.mask,
        )
        if self._fourier:
            sub._fourier, sub._in_spatial_domain = True, False
        return sub


class StreamingBatchIterator:
    """Contiguous batches of a StreamingCovarDataset through two device image buffers.

    The copy of batch i+1 is queued on the dataset's copy stream as soon as batch i is handed out;
    it waits for the compute stream to finish with that buffer (the step of batch i-1) and runs
    while the step of batch i computes.  The compute stream waits for a batch's copy before using
    it.  A batch's images must not be kept past the next batch (its buffer is then refilled).
    """

    def __init__(self, dataset: StreamingCovarDataset, batch_size: int) -> None:
        self.dataset = dataset
        self.batch_size = int(batch_size)

    def __len__(self) -> int:
        return math.ceil(len(self.dataset) / self.batch_size)

    def __iter__(self):
        ds, B = self.dataset, self.batch_size
        n = len(ds)
        dev = ds.device
        L = ds.resolution
        comp = torch.cuda.current_stream(dev)
        copy = ds.copy_stream
        bufs = [torch.empty((B, L, L), dtype=ds.images.dtype, device=dev) for _ in range(2)]
        for buf in bufs:
            buf.record_stream(copy)
        done = [None, None]
        ready = [None, None]

        def issue(i: int, k: int) -> None:
            a, b = i * B, min((i + 1) * B, n)
            if done[k] is not None:
                copy.wait_event(done[k])
            with torch.cuda.stream(copy):
                bufs[k][: b - a].copy_(ds.images[a:b], non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(copy)
            ready[k] = ev

        nb = len(self)
        if nb:
            issue(0, 0)
        for i in range(nb):
            k = i % 2
            if i + 1 < nb:
                issue(i + 1, 1 - k)
            comp.wait_event(ready[k])
            a, b = i * B, min((i + 1) * B, n)
            sl = slice(a, b)
            images = ds._finish(bufs[k][: b - a])
            pts = pts_rot_from_rotations(ds.rotations[sl], L)
            yield images, pts, ds.get_filters(sl), sl
            ev = torch.cuda.Event()
            ev.record(comp)  # after everything the consumer queued for this batch
            done[k] = ev


class BatchIterator:
    """Contiguous unshuffled batches of a dataset (BatchIndexSampler(shuffle=False))."""

    def __init__(self, dataset, batch_size: int, idx: Optional[torch.Tensor] = None) -> None:
        self.dataset = dataset
        self.batch_size = int(batch_size)
        self.idx = idx

    def __len__(self) -> int:
        n = len(self.dataset) if self.idx is None else len(self.idx)
        return math.ceil(n / self.batch_size)

    def __iter__(self):
        n = len(self.dataset) if self.idx is None else len(self.idx)
        for i in range(0, n, self.batch_size):
            if self.idx is None:
                yield self.dataset[i : min(i + self.batch_size, n)]
            else:
                yield self.dataset[self.idx[i : i + self.batch_size]]


def create_dataloader(dataset, batch_size: int, idx: Optional[torch.Tensor] = None, **unused):
    """Batch iterator over a dataset (dataset.py:850-895): contiguous unshuffled batches; a
    host-streamed dataset without an index subset streams through double device buffers."""
    if isinstance(dataset, StreamingCovarDataset) and idx is None:
        return StreamingBatchIterator(dataset, batch_size)
    return BatchIterator(dataset, batch_size, idx)


# ---------------------------------------------------------------------------
# Synthetic workloads
# ---------------------------------------------------------------------------


def gaussian_blob_volume(L: int, n_blobs: int, gen: torch.Generator, radius_frac: float = 0.6,
                         sigma_range=(0.04, 0.08)) -> torch.Tensor:
    """Sum of isotropic Gaussian blobs with centres uniform in a ball of radius radius_frac*L/2."""
    ax = torch.arange(L, dtype=torch.float64) - L // 2
    zz, yy, xx = torch.meshgrid(ax, ax, ax, indexing="ij")
    vol = torch.zeros((L, L, L), dtype=torch.float64)
    R = radius_frac * L / 2
    for _ in range(n_blobs):
        d = torch.randn(3, generator=gen, dtype=torch.float64)
        d = d / d.norm()
        rad = R * torch.rand(1, generator=gen, dtype=torch.float64) ** (1 / 3)
        c = d * rad
        s = (sigma_range[0] + (sigma_range[1] - sigma_range[0]) * torch.rand(1, generator=gen,
                                                                            dtype=torch.float64)) * L
        vol += torch.exp(-((zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2) / (2 * s**2))
    return vol.to(torch.float32)


@dataclass
class SyntheticConfig:
    L: int
    n_images: int
    rank: int
    eigenvalues: Sequence[float]
    pixel_size: float
    snr: float
    defocus_um: Tuple[float, float]
    shift_px: float = 0.0
    seed: int = 1000
    voltage_kv: float = 300.0
    cs_mm: float = 2.7
    amplitude_contrast: float = 0.1
    amp_jitter: float = 0.0  # per-image clean-signal amplitude 1 + amp_jitter * N(0, 1), unknown to the solver


def config_by_name(name: str) -> SyntheticConfig:
    """The benchmark workloads: L32 (CPU-sized), L64, L128 (SGD), L128-mean."""
    if name == "L32":
        return SyntheticConfig(32, 2000, 4, [4, 2, 1, 0.5], 3.0, 0.1, (1.0, 2.5), 0.0, 1001)
    if name == "L64":
        return SyntheticConfig(64, 20000, 10, list(np.geomspace(10, 0.1, 10)), 3.0, 0.05, (0.8, 3.0), 2.0, 1002)
    if name == "L128":
        return SyntheticConfig(128, 100000, 20, list(np.geomspace(10, 0.1, 20)), 1.5, 0.02, (0.5, 3.0), 0.0, 1003)
    if name == "L128-mean":
        return SyntheticConfig(128, 100000, 0, [], 1.5, 0.02, (0.5, 3.0), 0.0, 1004, amp_jitter=0.05)
    raise KeyError(name)


@dataclass
class SyntheticData:
    dataset: CovarDataset
    mean: torch.Tensor
    eigenvecs: torch.Tensor  # (r, L, L, L) orthonormal
    eigenvals: torch.Tensor
    latents: Optional[torch.Tensor] = None
    extra: dict = field(default_factory=dict)


def make_synthetic(cfg: SyntheticConfig, device, n_images: Optional[int] = None, subtract_mean: bool = True,
                   batch: int = 1024, store_filters: bool = True, host_images: bool = False) -> SyntheticData:
    """Generates images y_i = T_{o_i}[CTF_i * A_i(mean + sum_k z_ik u_k)] + noise on the device.

    Projections use the gfx950 type-2 NUFFT (vol_forward on a NufftPlan).  With
    ``cfg.shift_px > 0`` each clean image is translated by an offset o_i uniform in
    [-shift_px, shift_px]^2 (its centred spectrum times ``offset_to_phase_shift(o_i)``), and o_i is
    stored in ``dataset.offsets``: the reference's preprocessing and mean back-projection multiply
    by ``offset_to_phase_shift(-o_i)`` to undo it (mean.py:200-204, poses.py:263-264).
    With ``subtract_mean`` the projected mean is left out of the images, as the reference's
    preprocessing removes it before covariance estimation; it is left out before the shift, so
    such a dataset only needs the phase step of ``CovarDataset.preprocess`` (a zero mean volume).
    With ``host_images`` the images are kept in pinned host memory (a ``StreamingCovarDataset``;
    the device holds one batch at a time while generating) and the CTFs are made per batch.
    """
    from .nufft_plan import NufftPlan
    from .poses import get_phase_shift_grid, offset_to_phase_shift
    from .projection_funcs import vol_forward

    dev = torch.device(device)
    n = cfg.n_images if n_images is None else n_images
    L, r = cfg.L, cfg.rank
    gen = torch.Generator().manual_seed(cfg.seed)
    mean = gaussian_blob_volume(L, 8, gen)
    if r > 0:
        pert = torch.stack([gaussian_blob_volume(L, 4, gen) for _ in range(r)]).reshape(r, -1)
        q, _ = torch.linalg.qr(pert.T.double())
        eigvecs = q.T.reshape(r, L, L, L).to(torch.float32)
        eigvals = torch.tensor(cfg.eigenvalues, dtype=torch.float32)
    else:
        eigvecs = torch.zeros((0, L, L, L))
        eigvals = torch.zeros(0)
    rots = haar_rotations(n, gen).to(dev)
    du = torch.empty(n, dtype=torch.float64).uniform_(cfg.defocus_um[0], cfg.defocus_um[1], generator=gen) * 1e4
    ctf_params = torch.stack(
        [du, du, torch.zeros(n, dtype=torch.float64), torch.full((n,), cfg.voltage_kv, dtype=torch.float64),
         torch.full((n,), cfg.cs_mm, dtype=torch.float64), torch.full((n,), cfg.amplitude_contrast, dtype=torch.float64),
         torch.full((n,), cfg.pixel_size, dtype=torch.float64)], dim=1).to(torch.float32).to(dev)
    latents = (torch.randn((n, r), generator=gen) * eigvals.sqrt()).to(dev) if r > 0 else None
    amp = (1.0 + cfg.amp_jitter * torch.randn(n, generator=gen)).to(dev) if cfg.amp_jitter > 0 else None
    offsets = None
    if cfg.shift_px > 0:
        offsets = ((torch.rand((n, 2), generator=gen, dtype=torch.float64) * 2 - 1) * cfg.shift_px).to(
            torch.float32).to(dev)
        shift_grid = get_phase_shift_grid(L, dtype=torch.float32, device=dev)
    vols = torch.cat([mean.unsqueeze(0), eigvecs], dim=0).to(dev)
    plan = NufftPlan((L,) * 3, batch_size=r + 1, device=dev)
    if host_images:
        images = torch.empty((n, L, L), dtype=torch.float32, pin_memory=True)
        store_filters = False
    else:
        images = torch.empty((n, L, L), dtype=torch.float32, device=dev)
    filters = torch.empty((n, L, L), dtype=torch.float32, device=dev) if store_filters else None
    s1 = torch.zeros((), dtype=torch.float64, device=dev)
    s2 = torch.zeros((), dtype=torch.float64, device=dev)
    for i in range(0, n, batch):
        j = min(i + batch, n)
        plan.setpts(pts_rot_from_rotations(rots[i:j], L))
        f = ctf_from_params(ctf_params[i:j], L)
        proj = vol_forward(vols, plan, f)  # (b, r+1, L, L)
        if proj.ndim == 3:
            proj = proj.unsqueeze(0)
        img = proj[:, 1:].mul(latents[i:j, :, None, None]).sum(1) if r > 0 else torch.zeros_like(proj[:, 0])
        if not subtract_mean:
            img = img + proj[:, 0]
        if amp is not None:
            img = img * amp[i:j, None, None]
        if offsets is not None:
            img = centered_ifft2(centered_fft2(img) * offset_to_phase_shift(offsets[i:j], phase_shift_grid=shift_grid)).real
        if host_images:
            s1 += img.double().sum()
            s2 += (img.double() ** 2).sum()
            images[i:j].copy_(img)
        else:
            images[i:j] = img
        if filters is not None:
            filters[i:j] = f
    if host_images:
        cnt = n * L * L
        signal_var = float((s2 - s1 * s1 / cnt) / max(cnt - 1, 1))
    else:
        signal_var = float(images.var())
    noise_var = signal_var / cfg.snr if signal_var > 0 else 1.0
    g = torch.Generator(device=dev).manual_seed(cfg.seed + 7)
    for i in range(0, n, 8 * batch):
        j = min(i + 8 * batch, n)
        noise = math.sqrt(noise_var) * torch.randn((j - i, L, L), device=dev, generator=g)
        if host_images:
            images[i:j].copy_(images[i:j].to(dev) + noise)
        else:
            images[i:j] += noise
    if host_images:
        ds = StreamingCovarDataset(images, rots, noise_var, ctf_params=ctf_params, offsets=offsets)
    else:
        ds = CovarDataset(images, rots, noise_var, filters=filters, ctf_params=ctf_params, offsets=offsets)
    return SyntheticData(ds, mean.to(dev), eigvecs.to(dev), eigvals.to(dev), latents,
                         extra={} if amp is None else {"amplitudes": amp})
