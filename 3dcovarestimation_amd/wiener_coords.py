"""Wiener / MAP latent coordinates (cov3d/wiener_coords.py:85-203).

``compute_latentMAP_batch`` keeps the reference signature and returns
(z, M, P^H y / noise_var) with M = P^H P / noise_var + Lambda^{-1} and
z = M^{-1} P^H y / noise_var.  For real projected eigenvectors (the NUFFT
path) it runs the fused gfx950 kernel (Gram + FP64 Cholesky); complex
inputs take the plain torch formulation.
"""

from __future__ import annotations

import math
from typing import Any, Optional, Tuple, Union

import torch

from . import ml_ops
from .nufft_plan import NufftPlan, NufftPlanDiscretized
from .projection_funcs import centered_fft2, centered_fft3, vol_forward


def compute_latentMAP_batch(
    images: torch.Tensor,
    eigen_forward: torch.Tensor,
    noise_var: float,
    eigenvals_inv: Optional[torch.Tensor] = None,
) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Plain torch form of wiener_coords.py:168-203 (complex or real, differentiable)."""
    n = images.shape[0]
    r = eigen_forward.shape[1]
    if eigenvals_inv is None:
        eigenvals_inv = torch.eye(r, device=eigen_forward.device, dtype=eigen_forward.dtype)
    images = images.reshape(n, -1, 1)
    P = eigen_forward.reshape(n, r, -1)
    m = P.conj() @ P.transpose(1, 2) / noise_var + eigenvals_inv
    proj = torch.matmul(P.conj(), images) / noise_var
    mean_m = m.diagonal(dim1=-2, dim2=-1).abs().sum(dim=1) / m.shape[-1]
    z = torch.linalg.solve(m / mean_m.reshape(-1, 1, 1), proj) / mean_m.reshape(-1, 1, 1)
    return z, m, proj


def latentMAP(
    dataset,
    eigenvecs: torch.Tensor,
    eigenvals: torch.Tensor,
    batch_size: int = 1024,
    start_ind: Optional[int] = None,
    end_ind: Optional[int] = None,
    return_coords_covar: bool = False,
    nufft_plan: type = NufftPlan,
    precision: int = ml_ops.PRECISION_BF16X3,
    **nufft_plan_kwargs: Any,
) -> Union[torch.Tensor, Tuple[torch.Tensor, torch.Tensor]]:
    """Latent coordinates of every image (wiener_coords.py:85-165).

    With ``NufftPlan`` each batch runs type-2 NUFFT + the fused ML kernel
    (Gram on MFMA, FP64 Cholesky per image); with ``NufftPlanDiscretized``
    it follows the reference's discretised projection and the torch solve.
    """
    start_ind = 0 if start_ind is None else start_ind
    end_ind = len(dataset) if end_ind is None else end_ind
    L = eigenvecs.shape[-1]
    rank = eigenvecs.shape[0]
    device = eigenvecs.device
    dtype = eigenvecs.dtype
    if eigenvals.ndim == 1:
        eigenvals = torch.diag(eigenvals)
    eig_inv = torch.inverse(eigenvals.to(device) + 1e-6 * torch.eye(rank, device=device, dtype=dtype))
    if nufft_plan is NufftPlan:
        plan = NufftPlan((L,) * 3, batch_size=rank, dtype=dtype, device=device, **nufft_plan_kwargs)
        vols = eigenvecs
    else:
        kw = {"upsample_factor": 2, "mode": "bilinear"}
        kw.update(nufft_plan_kwargs)
        plan = NufftPlanDiscretized((L,) * 3, **kw)
        vols = centered_fft3(eigenvecs, padding_size=(L * kw["upsample_factor"],) * 3)
    n_total = end_ind - start_ind
    coords = torch.zeros((n_total, rank), device=device)
    covar = torch.zeros((n_total, rank, rank), dtype=dtype) if return_coords_covar else None
    noise_var = float(dataset.noise_var)
    for i in range(0, n_total, batch_size):
        images, pts_rot, filters, _ = dataset[start_ind + i : min(start_ind + i + batch_size, end_ind)]
        images, pts_rot, filters = images.to(device), pts_rot.to(device), filters.to(device)
        plan.setpts(pts_rot)
        if isinstance(plan, NufftPlan) and dtype == torch.float32:
            c = plan.execute_forward(vols)[:rank].reshape(rank, images.shape[0], L * L)
            res = ml_ops.ml_forward(
                c, centered_fft2(images), L, noise_var, ctf=filters, lam_inv=eig_inv,
                precision=precision, want_m=return_coords_covar, want_minv=False,
            )
            z = res["z"]
            m = res.get("m")
        else:
            ef = vol_forward(vols, plan, filters)
            zz, m, _ = compute_latentMAP_batch(images, ef, noise_var, eig_inv)
            z = zz.squeeze(-1).real
        coords[i : i + images.shape[0]] = z
        if return_coords_covar:
            covar[i : i + images.shape[0]] = m.real.to("cpu")
    del plan
    return coords if not return_coords_covar else (coords, covar)
