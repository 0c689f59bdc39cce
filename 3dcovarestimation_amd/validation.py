"""Subspace agreement of SGD forms: the chosen form against an FP64 trajectory.

``sgd_subspaces`` trains one ``Covar`` per form from the same initial vectors
on the same batches (no shuffling, as the reference's loader) and returns the
orthonormal bases of the learned eigen-subspaces.  The FP64 form runs the
plain autograd objective (cov3d/covar_sgd.py:1238-1286) on an FP64 NUFFT plan
(eps 1e-12) with FP64 parameters, images and CTFs; its points are the same
FP32 pts_rot as every other form, so the comparison isolates arithmetic.

``subspace_distance`` is the RMS sine of the principal angles between two
r-dimensional subspaces, ||(I - Q1 Q1^T) Q2||_F / sqrt(r), in [0, 1].
"""

from __future__ import annotations

import time
from typing import Dict, Iterable, Optional, Sequence, Tuple

import torch

from . import ml_ops
from .covar import Covar
from .covar_sgd import CovarTrainer
from .dataset import CovarDataset, create_dataloader
from .nufft_plan import NufftPlan

# tag -> (form, precision, dtype, half plane)
FORMS = {
    "fp64": ("autograd", ml_ops.PRECISION_FP32, torch.float64, False),
    "chosen": ("fused", ml_ops.PRECISION_BF16X3, torch.float32, True),
    "fused_fp32": ("fused", ml_ops.PRECISION_FP32, torch.float32, True),
    "fused_fp32_full": ("fused", ml_ops.PRECISION_FP32, torch.float32, False),
    "chosen_full": ("fused", ml_ops.PRECISION_BF16X3, torch.float32, False),
    "plain_fp32": ("autograd", ml_ops.PRECISION_FP32, torch.float32, False),
}


def subspace_distance(Q1: torch.Tensor, Q2: torch.Tensor) -> float:
    """RMS sine of the principal angles between span(Q1) and span(Q2); Q orthonormal (n, r)."""
    r = Q2.shape[1]
    return float(torch.linalg.norm(Q2 - Q1 @ (Q1.T @ Q2)) / r**0.5)


def basis(covar: Covar) -> torch.Tensor:
    ev, _ = covar.eigenvecs
    q, _ = torch.linalg.qr(ev.reshape(covar.rank, -1).double().T)
    return q


def _fp64_dataset(ds: CovarDataset) -> CovarDataset:
    filt = ds.filters.double() if ds.filters is not None else None
    params = ds.ctf_params if filt is None else None
    return CovarDataset(ds.images.double(), ds.rotations, ds.noise_var, filters=filt, ctf_params=params)


def sgd_subspaces(dataset: CovarDataset, rank: int, v0: torch.Tensor, batch_size: int, epochs: int = 1,
                  lr: float = 1e-5, forms: Sequence[str] = ("fp64", "chosen", "plain_fp32"),
                  n_images: Optional[int] = None, fp64_eps: float = 1e-12
                  ) -> Tuple[Dict[str, torch.Tensor], Dict[str, float]]:
    """Bases of the subspaces learned by each form, and each form's training seconds."""
    dev = dataset.device
    L = dataset.resolution
    idx = torch.arange(n_images if n_images is not None else len(dataset), device=dev)
    d64 = None
    out, secs = {}, {}
    for tag in forms:
        form, prec, dt, half = FORMS[tag]
        ds = dataset
        if dt == torch.float64:
            if d64 is None:
                d64 = _fp64_dataset(dataset)
            ds = d64
        covar = Covar(L, rank, dtype=dt, pixel_var_estimate=1.0, vectors=v0.to(dt)).to(dev)
        trainer = CovarTrainer(covar, create_dataloader(ds, batch_size, idx), dev)
        trainer.setup_training(nufft_disc="nufft", objective_func="ml", lr=lr, form=form, precision=prec)
        trainer.half_plane = half
        if dt == torch.float64:
            trainer.nufft_plans = NufftPlan((L,) * 3, batch_size=rank, eps=fp64_eps, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        trainer.train_epochs(epochs)
        torch.cuda.synchronize()
        secs[tag] = time.perf_counter() - t0
        out[tag] = basis(covar)
        del trainer, covar
    return out, secs


def subspace_errors(bases: Dict[str, torch.Tensor], ref: str = "fp64") -> Dict[str, float]:
    return {k: subspace_distance(bases[ref], v) for k, v in bases.items() if k != ref}
