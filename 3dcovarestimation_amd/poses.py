"""Per-image poses: in-plane translations as Fourier phase shifts, rotations as rotation vectors,
and the pose module the pose-refining trainer optimises (cov3d/poses.py).

The image preprocessing step (projection_funcs.py:199-252) multiplies each image's centred
spectrum by exp(i (g_x o_x + g_y o_y)); the pose module passes the negated stored offsets
(poses.py:263-264), so that a stored offset o undoes a translation of the particle by o.
"""

from __future__ import annotations

import math
from typing import Optional, Tuple

import torch

from .projection_funcs import centered_fft2, centered_ifft2


def get_phase_shift_grid(resolution: int, dtype: torch.dtype = torch.float32,
                         device: Optional[torch.device] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """Frequency grids (g_x, g_y), each (L, L) (poses.py:68-87).

    g = 2 pi ceil(k) / L for k = -L/2, -L/2 + 1, ..., L/2 - 1 (the centred DFT frequencies);
    g_x varies along the last axis and g_y along the first (``meshgrid(..., indexing="xy")``).
    """
    k = torch.ceil(torch.arange(-resolution / 2, resolution / 2, dtype=torch.float64))
    g = (k * (2 * math.pi / resolution)).to(dtype)
    gy, gx = torch.meshgrid(g, g, indexing="ij")
    dev = device if device is not None else torch.device("cpu")
    return gx.to(dev), gy.to(dev)


def offset_to_phase_shift(offsets: torch.Tensor, resolution: Optional[int] = None,
                          phase_shift_grid: Optional[Tuple[torch.Tensor, torch.Tensor]] = None) -> torch.Tensor:
    """exp(i (g_x o_x + g_y o_y)) for offsets (N, 2) -> (N, L, L) complex (poses.py:92-125).

    Raises ValueError when neither ``resolution`` nor ``phase_shift_grid`` is given, as the
    reference does.
    """
    if phase_shift_grid is None:
        if resolution is None:
            raise ValueError("Either resolution or phase_shift_grid must be provided.")
        phase_shift_grid = get_phase_shift_grid(resolution, dtype=offsets.dtype, device=offsets.device)
    gx, gy = phase_shift_grid
    phase = gx.unsqueeze(0) * offsets[:, 0].reshape(-1, 1, 1) + gy.unsqueeze(0) * offsets[:, 1].reshape(-1, 1, 1)
    return torch.polar(torch.ones_like(phase), phase)


def rotvec_to_rotmat(rotvecs: torch.Tensor) -> torch.Tensor:
    """Rodrigues' formula, differentiable (poses.py:44-67): R = I + sin(t) K + (1 - cos(t)) K^2 for the
    unit axis k = v / (|v| + 1e-6) and angle t = |v|."""
    theta = torch.norm(rotvecs, dim=-1, keepdim=True)
    k = rotvecs / (theta + 1e-6)
    zero = torch.zeros_like(k[:, 0])
    K = torch.stack([zero, -k[:, 2], k[:, 1], k[:, 2], zero, -k[:, 0], -k[:, 1], k[:, 0], zero], dim=1).reshape(-1, 3, 3)
    eye = torch.eye(3, device=rotvecs.device, dtype=rotvecs.dtype).unsqueeze(0)
    return eye + torch.sin(theta).unsqueeze(-1) * K + (1 - torch.cos(theta).unsqueeze(-1)) * (K @ K)


class PoseModule(torch.nn.Module):
    """Per-image rotations (rotation vectors), in-plane offsets and optional contrasts as sparse
    embeddings (poses.py:127-330), so that SparseAdam updates only the images of a batch.

    forward(index) returns the rotated, wrapped slice frequencies (B, 3, L^2) (zyx order, as
    cov3d_pts_rot makes them from rotation matrices) and the phase shift of the negated offsets
    (B, L, L) [and the contrasts (B, 1)], all differentiable in the parameters.
    """

    def __init__(self, init_rotvecs, offsets, resolution: int, dtype: torch.dtype = torch.float32,
                 use_contrast: bool = False) -> None:
        super().__init__()
        self.resolution = resolution
        self.dtype = dtype
        self.use_contrast = use_contrast
        init_rotvecs = torch.as_tensor(init_rotvecs, dtype=dtype)
        offsets = torch.as_tensor(offsets, dtype=dtype)
        if init_rotvecs.ndim != 2 or init_rotvecs.shape[1] != 3:
            raise ValueError("rotation vectors must have shape (N, 3)")
        if offsets.ndim != 2 or offsets.shape[1] != 2:
            raise ValueError("offsets must have shape (N, 2)")
        if init_rotvecs.shape[0] != offsets.shape[0]:
            raise ValueError("rotation vectors and offsets must have the same number of rows")
        n = init_rotvecs.shape[0]
        self.rotvec = torch.nn.Embedding(n, 3, sparse=True, _weight=init_rotvecs.clone())
        self.offsets = torch.nn.Embedding(n, 2, sparse=True, _weight=offsets.clone())
        if use_contrast:
            self.contrasts = torch.nn.Embedding(n, 1, sparse=True, _weight=torch.ones((n, 1), dtype=dtype))
        L = resolution
        g = torch.ceil(torch.arange(-L / 2, L / 2, dtype=torch.float64))[:L] / (L / 2)
        gy, gx = torch.meshgrid(g, g, indexing="ij")
        grid = math.pi * torch.stack([gx.reshape(-1), gy.reshape(-1), torch.zeros(L * L, dtype=torch.float64)])
        self.register_buffer("xy_rot_grid", grid.to(dtype), persistent=False)
        px, py = get_phase_shift_grid(L, dtype=dtype)
        self.register_buffer("phase_shift_grid_x", px, persistent=False)
        self.register_buffer("phase_shift_grid_y", py, persistent=False)

    @property
    def device(self) -> torch.device:
        return self.xy_rot_grid.device

    def forward(self, index: torch.Tensor, ds_resolution: Optional[int] = None):
        if ds_resolution is not None and ds_resolution != self.resolution:
            raise ValueError("downsampled pose grids are not supported")
        R = rotvec_to_rotmat(self.rotvec(index))
        n = R.shape[0]
        pts = torch.flip((R.reshape(n * 3, 3) @ self.xy_rot_grid).reshape(n, 3, -1), dims=[1])
        pts = torch.remainder(pts + torch.pi, 2 * torch.pi) - torch.pi
        phase = offset_to_phase_shift(-self.offsets(index), phase_shift_grid=(self.phase_shift_grid_x,
                                                                               self.phase_shift_grid_y))
        if self.use_contrast:
            return pts, phase, self.contrasts(index)
        return pts, phase

    def get_rotvecs(self) -> torch.Tensor:
        return self.rotvec.weight.data

    def set_rotvecs(self, rotvecs: torch.Tensor, idx: Optional[torch.Tensor] = None) -> None:
        with torch.no_grad():
            if idx is None:
                self.rotvec.weight.copy_(rotvecs)
            else:
                self.rotvec.weight[idx] = rotvecs

    def get_offsets(self) -> torch.Tensor:
        return self.offsets.weight.data

    def set_offsets(self, offsets: torch.Tensor, idx: Optional[torch.Tensor] = None) -> None:
        with torch.no_grad():
            if idx is None:
                self.offsets.weight.copy_(offsets)
            else:
                self.offsets.weight[idx] = offsets

    def get_contrasts(self) -> torch.Tensor:
        return self.contrasts.weight.data

    def set_contrasts(self, contrasts: torch.Tensor, idx: Optional[torch.Tensor] = None) -> None:
        with torch.no_grad():
            if idx is None:
                self.contrasts.weight.copy_(contrasts)
            else:
                self.contrasts.weight[idx] = contrasts

    def _rows(self, idx: torch.Tensor) -> "PoseModule":
        """A new module on this module's device holding copies of rows ``idx``."""
        idx = idx.to(self.rotvec.weight.device)
        m = PoseModule(self.rotvec.weight.data[idx].detach().clone(), self.offsets.weight.data[idx].detach().clone(),
                       self.resolution, dtype=self.dtype, use_contrast=self.use_contrast)
        if self.use_contrast:
            m.set_contrasts(self.contrasts.weight.data[idx].detach().clone())
        return m.to(self.device)

    def split_module(self, permutation: Optional[torch.Tensor] = None) -> Tuple["PoseModule", "PoseModule"]:
        """Two modules holding the rows permutation[:n // 2] and permutation[n // 2:] (identity order
        when no permutation is given), for the two half-sets of the same split (poses.py:350-389)."""
        n = self.rotvec.weight.shape[0]
        if permutation is None:
            permutation = torch.arange(n)
        permutation = torch.as_tensor(permutation, dtype=torch.long)
        return self._rows(permutation[: n // 2]), self._rows(permutation[n // 2:])

    @staticmethod
    def _joined(module1: "PoseModule", module2: "PoseModule", permutation: torch.Tensor):
        """The rows of two split modules put back in the original order: row permutation[k] of the
        result is row k of the two modules stacked."""
        if module1.use_contrast != module2.use_contrast:
            raise ValueError("both modules must have the same use_contrast setting")
        dev = module1.rotvec.weight.device
        perm = torch.as_tensor(permutation, dtype=torch.long).to(dev)

        def place(a, b):
            stacked = torch.cat([a.detach(), b.detach().to(dev)], dim=0)
            out = torch.empty_like(stacked)
            out[perm] = stacked
            return out

        rot = place(module1.rotvec.weight.data, module2.rotvec.weight.data)
        off = place(module1.offsets.weight.data, module2.offsets.weight.data)
        con = (place(module1.contrasts.weight.data, module2.contrasts.weight.data) if module1.use_contrast
               else None)
        return rot, off, con

    @staticmethod
    def merge_modules(module1: "PoseModule", module2: "PoseModule", permutation: torch.Tensor) -> "PoseModule":
        """The module that ``split_module(permutation)`` split into module1 and module2
        (poses.py:391-429)."""
        rot, off, con = PoseModule._joined(module1, module2, permutation)
        m = PoseModule(rot, off, module1.resolution, dtype=module1.dtype, use_contrast=module1.use_contrast)
        if con is not None:
            m.set_contrasts(con)
        return m.to(module1.device)

    def update_from_modules(self, module1: "PoseModule", module2: "PoseModule", permutation: torch.Tensor) -> None:
        """Takes the rows of two split modules back in place (poses.py:431-461)."""
        rot, off, con = PoseModule._joined(module1, module2, permutation)
        self.set_rotvecs(rot.to(self.rotvec.weight.device))
        self.set_offsets(off.to(self.offsets.weight.device))
        if con is not None and self.use_contrast:
            self.set_contrasts(con.to(self.contrasts.weight.device))


def estimate_image_offsets_newton(images: torch.Tensor, reference: torch.Tensor, mask: Optional[torch.Tensor] = None,
                                  init_offsets: Optional[torch.Tensor] = None, in_fourier_domain: bool = False,
                                  obj_func=None, iterations: int = 10, beta: float = 0.2, max_ls_steps: int = 3,
                                  step_size_limit: float = 0.5) -> torch.Tensor:
    """Per-image in-plane offsets by Newton's method (poses.py:500-560).

    The objective (default |phase-shifted image - reference|^2 per image, or ``obj_func`` of the
    phase-shifted Fourier images returning one value per image) is separable over images, so its
    Hessian in the (N, 2) offsets is block diagonal: two Hessian-vector products of the summed
    gradient give every image's 2 x 2 block.  Each iteration takes the Newton step per image
    (gradient step where the block is not positive definite), limited to ``step_size_limit``
    pixels, with up to ``max_ls_steps`` backtracking halvings (factor ``beta``) where the objective
    does not decrease.
    """
    if not in_fourier_domain:
        images = centered_fft2(images * mask) if mask is not None else centered_fft2(images)
        reference = centered_fft2(reference * mask) if mask is not None else centered_fft2(reference)
    elif mask is not None:
        images = centered_fft2(centered_ifft2(images).real * mask)
        reference = centered_fft2(centered_ifft2(reference).real * mask)
    dtype = images.real.dtype
    n, L = images.shape[0], images.shape[-1]
    off = (torch.zeros((n, 2), dtype=dtype, device=images.device) if init_offsets is None
           else init_offsets.detach().clone().to(dtype))
    grid = get_phase_shift_grid(L, dtype=dtype, device=images.device)
    if obj_func is None:
        def obj_func(shifted):
            return (shifted - reference).abs().pow(2).sum(dim=(-1, -2))

    def f(o):
        return obj_func(images * offset_to_phase_shift(-o, phase_shift_grid=grid)).real

    for _ in range(iterations):
        o = off.clone().requires_grad_(True)
        val = f(o)
        (g,) = torch.autograd.grad(val.sum(), o, create_graph=True)
        h0 = torch.autograd.grad(g[:, 0].sum(), o, retain_graph=True)[0]
        h1 = torch.autograd.grad(g[:, 1].sum(), o)[0]
        g = g.detach()
        H = torch.stack([h0, h1], dim=1).detach()  # (n, 2, 2): H[:, i, :] = d g_i / d o
        det = H[:, 0, 0] * H[:, 1, 1] - H[:, 0, 1] * H[:, 1, 0]
        pd = (H[:, 0, 0] > 0) & (det > 0)
        Hs = torch.where(pd[:, None, None], H, torch.eye(2, dtype=dtype, device=H.device).expand_as(H))
        step = -torch.linalg.solve(Hs, g.unsqueeze(-1)).squeeze(-1)
        norm = step.norm(dim=1, keepdim=True)
        step = step * torch.clamp(step_size_limit / torch.clamp(norm, min=1e-12), max=1.0)
        val = val.detach()
        t = torch.ones((n, 1), dtype=dtype, device=off.device)
        with torch.no_grad():
            for _ in range(max_ls_steps):
                better = f(off + t * step) < val
                if bool(better.all()):
                    break
                t = torch.where(better.unsqueeze(1), t, t * beta)
            keep = f(off + t * step) < val
            off = torch.where(keep.unsqueeze(1), off + t * step, off)
    return off.detach()
