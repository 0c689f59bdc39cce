"""Covariance SGD: objectives, trainer and entry point (cov3d/covar_sgd.py).

Objectives keep the reference signatures (covar_sgd.py:1190-1399):
``cost`` (LS, spatial), ``cost_maximum_liklihood`` (ML, spatial),
``cost_fourier_domain`` / ``raw_cost_fourier_domain`` (LS, Fourier) and
``cost_maximum_liklihood_fourier_domain`` /
``raw_cost_maximum_liklihood_fourier_domain`` (ML, Fourier).

``cost_maximum_liklihood`` on a ``NufftPlan`` has two forms:

* ``form="autograd"``: the reference algorithm in plain FP32 torch ops
  (vol_forward, matmul, solve, logdet) differentiated by autograd; the NUFFT
  and its type-1 adjoint still run on the gfx950 kernels;
* ``form="fused"`` (default): ``ml_ops.MLCostNufft`` - split-BF16 MFMA Gram,
  FP64 Cholesky and the hand-written adjoint (gradient images + one type-1
  NUFFT), with the same value and gradient.

``CovarTrainer`` / ``trainCovar`` follow covar_sgd.py:48-545 and :1479-1612
(Adam with per-group lr factors, lr scaled by batch size, ReduceLROnPlateau
with patience 0, stop after 4 reductions, no shuffling).
"""

from __future__ import annotations

import copy
import logging
import time

import numpy as np
from typing import Any, Dict, List, Optional, Tuple, Union

import torch

from . import ml_ops
from .covar import Covar
from .fsc_utils import covar_fsc, expand_fourier_shell
from .newton_opt import BlockwiseLBFGS
from .nufft_plan import NufftPlan, NufftPlanDiscretized
from .projection_funcs import centered_fft2, centered_fft3, vol_forward
from .wiener_coords import compute_latentMAP_batch

logger = logging.getLogger(__name__)

_DEFAULT_FORM = "fused"


def set_default_form(form: str) -> None:
    """Selects the implementation of the spatial ML objective: "fused" or "autograd"."""
    global _DEFAULT_FORM
    if form not in ("fused", "autograd"):
        raise ValueError("form must be 'fused' or 'autograd'")
    _DEFAULT_FORM = form


def _ml_regulariser(vols, fourier_reg, noise_var, reg_scale, in_fourier):
    rank = vols.shape[0]
    vf = vols if in_fourier else centered_fft3(vols)
    vf = vf * torch.sqrt(fourier_reg / noise_var)
    return reg_scale * torch.sum(torch.norm(vf.reshape(rank, -1), dim=1) ** 2) / 2


def cost(vols, images, nufft_plans, filters, noise_var, reg_scale=0, fourier_reg=None,
         apply_mean_const_term=False, mean_aggregate=True):
    """Least-squares objective, spatial domain (covar_sgd.py:1190-1235)."""
    B = images.shape[0]
    rank = vols.shape[0]
    L = vols.shape[-1]
    P = vol_forward(vols, nufft_plans, filters).reshape(B, rank, -1)
    y = images.reshape(B, 1, -1)
    py = torch.matmul(P, y.transpose(1, 2))
    pp = torch.matmul(P, P.transpose(1, 2))
    val = -2 * torch.sum(py**2, dim=(1, 2)) + torch.sum(pp**2, dim=(1, 2))
    val = val + 2 * noise_var * torch.sum(torch.diagonal(pp, dim1=1, dim2=2), dim=1)
    if apply_mean_const_term:
        ny = torch.norm(y, dim=(1, 2)) ** 2
        val = val + ny**2 - (2 * noise_var * ny + (noise_var * L) ** 2)
    val = torch.mean(val, dim=0) if mean_aggregate else val
    if fourier_reg is not None and reg_scale != 0:
        vf = (centered_fft3(vols) * torch.sqrt(fourier_reg)).reshape(rank, -1)
        val = val + reg_scale * torch.sum((vf @ vf.conj().T).abs() ** 2)
    return val


def _cost_ml_autograd(vols, images, nufft_plans, filters, noise_var, apply_mean_const_term, mean_aggregate):
    B = images.shape[0]
    rank = vols.shape[0]
    P = vol_forward(vols, nufft_plans, filters).reshape(B, rank, -1)
    y = images.reshape(B, -1, 1)
    py = torch.matmul(P, y)
    m = torch.eye(rank, device=vols.device, dtype=vols.dtype).unsqueeze(0) + P @ P.transpose(1, 2) / noise_var
    mean_m = m.diagonal(dim1=-2, dim2=-1).abs().sum(dim=1) / m.shape[-1]
    pyt = torch.linalg.solve(m / mean_m.reshape(-1, 1, 1), py) / mean_m.reshape(-1, 1, 1)
    exp_term = -1 / (noise_var**2) * torch.matmul(py.transpose(1, 2).conj(), pyt).squeeze()
    noise_term = torch.logdet(m)
    if apply_mean_const_term:
        exp_term = exp_term + 1 / noise_var * torch.norm(y, dim=(1, 2)) ** 2
    return 0.5 * torch.mean(exp_term + noise_term) if mean_aggregate else 0.5 * (exp_term + noise_term)


def _only_vols_need_grad(images, filters, plan) -> bool:
    """The fused objectives return gradients of the volumes only: images, filters (contrasts) or
    points (rotations) that need gradients take the autograd form."""
    pts = getattr(plan, "points", None) if isinstance(plan, NufftPlan) else getattr(plan, "_grid_flat", None)
    return not (images.requires_grad or (filters is not None and filters.requires_grad)
                or (pts is not None and pts.requires_grad))


def cost_maximum_liklihood(vols, images, nufft_plans, filters, noise_var, reg_scale=0, fourier_reg=None,
                           apply_mean_const_term=False, mean_aggregate=True, form=None,
                           precision=ml_ops.PRECISION_BF16X3, half_plane=True):
    """Maximum-likelihood objective, spatial domain (covar_sgd.py:1238-1286)."""
    form = form or _DEFAULT_FORM
    fused_ok = (
        form == "fused" and isinstance(nufft_plans, NufftPlan) and mean_aggregate
        and vols.dtype == torch.float32 and vols.shape[0] <= 31 and vols.ndim == 4
        and _only_vols_need_grad(images, filters, nufft_plans)
    )
    if fused_ok:
        L = vols.shape[-1]
        yhat = centered_fft2(images.reshape(-1, L, L).to(torch.float32))
        val = ml_ops.MLCostNufft.apply(vols, yhat, nufft_plans, filters, float(noise_var),
                                       bool(apply_mean_const_term), int(precision), bool(half_plane))
    else:
        val = _cost_ml_autograd(vols, images, nufft_plans, filters, noise_var, apply_mean_const_term, mean_aggregate)
    if fourier_reg is not None and reg_scale != 0:
        val = val + _ml_regulariser(vols, fourier_reg, noise_var, reg_scale, in_fourier=False)
    return val


def raw_cost_fourier_domain(projected_vols, images, noise_var, apply_mean_const_term=False, mean_aggregate=True):
    """LS objective on Fourier-domain projections (covar_sgd.py:1323-1347)."""
    B = images.shape[0]
    L = images.shape[-1]
    y = images.reshape(B, 1, -1)
    P = projected_vols.reshape(B, -1, L**2)
    py = torch.matmul(P, y.transpose(1, 2).conj())
    pp = torch.matmul(P, P.transpose(1, 2).conj())
    val = -2 * torch.sum(py.abs() ** 2, dim=(1, 2)) + torch.sum(pp.abs() ** 2, dim=(1, 2))
    val = val + 2 * noise_var * torch.sum(torch.diagonal(pp, dim1=1, dim2=2).real, dim=1)
    if apply_mean_const_term:
        ny = torch.norm(y, dim=(1, 2)) ** 2
        val = val + ny**2 - (2 * noise_var * ny + (noise_var * L) ** 2)
    return torch.mean(val, dim=0) if mean_aggregate else val


def cost_fourier_domain(vols, images, nufft_plans, filters, noise_var, reg_scale=0, fourier_reg=None,
                        apply_mean_const_term=False, mean_aggregate=True):
    """LS objective, Fourier domain, divided by L^4 (covar_sgd.py:1290-1320)."""
    rank = vols[0].shape[0] if isinstance(vols, tuple) else vols.shape[0]
    L = images.shape[-1]
    P = vol_forward(vols, nufft_plans, filters, fourier_domain=True)
    val = raw_cost_fourier_domain(P, images, noise_var, apply_mean_const_term, mean_aggregate)
    if fourier_reg is not None and reg_scale != 0:
        vf = (vols * torch.sqrt(fourier_reg)).reshape(rank, -1)
        val = val + reg_scale * torch.sum((vf @ vf.conj().T).abs() ** 2)
    return val / (L**4)


def raw_cost_maximum_liklihood_fourier_domain(projected_eigenvecs, images, noise_var, apply_mean_const_term=False,
                                              mean_aggregate=True):
    """ML objective on Fourier-domain projections (covar_sgd.py:1381-1399)."""
    z, m, pimg = compute_latentMAP_batch(images, projected_eigenvecs, noise_var)
    exp_term = -torch.matmul(pimg.transpose(1, 2).conj(), z).squeeze()
    noise_term = torch.logdet(m)
    if apply_mean_const_term:
        exp_term = exp_term + 1 / noise_var * torch.norm(images, dim=(1, 2)) ** 2
    tot = exp_term + noise_term
    return 0.5 * torch.mean(tot).real if mean_aggregate else 0.5 * tot.real


def cost_maximum_liklihood_fourier_domain(vols, images, nufft_plans, filters, noise_var, reg_scale=0,
                                          fourier_reg=None, apply_mean_const_term=False, mean_aggregate=True,
                                          form=None, precision=ml_ops.PRECISION_BF16X3, half_plane=True):
    """ML objective, Fourier domain (covar_sgd.py:1350-1378).

    ``form="fused"`` (default) on a NufftPlanDiscretized runs ``ml_ops.MLCostDiscretized``: native
    slice gather, split-BF16 MFMA Gram with FP64 Cholesky, hand-written adjoint and native slice
    scatter; ``form="autograd"`` is the reference's torch formulation differentiated by autograd."""
    form = form or _DEFAULT_FORM
    fused_ok = (
        form == "fused" and isinstance(nufft_plans, NufftPlanDiscretized) and mean_aggregate
        and not isinstance(vols, tuple) and vols.dtype == torch.complex64 and vols.ndim == 4
        and vols.shape[0] <= 31 and images.is_complex() and _only_vols_need_grad(images, filters, nufft_plans)
    )
    if fused_ok:
        val = ml_ops.MLCostDiscretized.apply(vols, images.to(torch.complex64), nufft_plans, filters,
                                             float(noise_var), bool(apply_mean_const_term), int(precision),
                                             bool(half_plane))
    else:
        P = vol_forward(vols, nufft_plans, filters, fourier_domain=True)
        val = raw_cost_maximum_liklihood_fourier_domain(P, images, noise_var, apply_mean_const_term,
                                                        mean_aggregate)
    if fourier_reg is not None and reg_scale != 0:
        val = val + _ml_regulariser(vols, fourier_reg, noise_var, reg_scale, in_fourier=True)
    return val


def frobeniusNorm(vecs):
    vecs = vecs.reshape(vecs.shape[0], -1)
    g = vecs @ vecs.transpose(0, 1).conj()
    return torch.sqrt(torch.sum(g.abs() ** 2))


class CovarTrainer:
    """Minibatch trainer of a ``Covar`` model (covar_sgd.py:48-545)."""

    def __init__(self, covar: Covar, train_data, device: torch.device, save_path: Optional[str] = None,
                 gt_data=None, training_log_freq: int = 50) -> None:
        self.device = torch.device(device)
        self.train_data = train_data
        self._covar = covar.to(self.device)
        self.batch_size = train_data.batch_size
        self.isDDP = False
        self.save_path = save_path
        self.logTraining = True
        self.training_log_freq = training_log_freq
        self.gt_data = gt_data
        self.training_log = {"epoch_ind": [], "lr_history": [], "cost_val": [], "epoch_run_time": []}
        self.num_reduced_lr_before_stop = 4
        self.scheduler_patiece = 0
        self.fourier_reg = None
        self.form = _DEFAULT_FORM
        self.precision = ml_ops.PRECISION_BF16X3
        self.half_plane = True

    @property
    def dataset(self):
        return self.train_data.dataset

    @property
    def covar(self) -> Covar:
        return self._covar

    @property
    def noise_var(self):
        return self.dataset.noise_var

    def to(self, *args, **kwargs):
        self._covar = self._covar.to(*args, **kwargs)
        if self.fourier_reg is not None:
            self.fourier_reg = self.fourier_reg.to(*args, **kwargs)
        return self

    def run_batch(self, images: torch.Tensor, pts_rot: torch.Tensor, filters: torch.Tensor) -> torch.Tensor:
        self.nufft_plans.setpts(pts_rot)
        self.optimizer.zero_grad()
        kw = {}
        if self.cost_func is cost_maximum_liklihood:
            kw = {"form": self.form, "precision": self.precision, "half_plane": self.half_plane}
        elif self.cost_func is cost_maximum_liklihood_fourier_domain:
            kw = {"form": self.form, "precision": self.precision, "half_plane": self.half_plane}
        cost_val = self.cost_func(self._covar(None), images, self.nufft_plans, filters, self.noise_var,
                                  self.reg_scale, self.fourier_reg, **kw)
        cost_val.backward()
        self.optimizer.step()
        if self.use_orthogonal_projection:
            self.covar.orthogonal_projection()
        return cost_val

    def prepare_batch(self, batch):
        images, pts_rot, filters, _ = batch
        return images.to(self.device), pts_rot.to(self.device), filters.to(self.device)

    def run_epoch(self, epoch: int) -> None:
        self.cost_in_epoch = torch.tensor(0, device=self.device, dtype=torch.float32)
        for batch in self.train_data:
            images, pts_rot, filters = self.prepare_batch(batch)
            cost_val = self.run_batch(images, pts_rot, filters)
            with torch.no_grad():
                self.cost_in_epoch += cost_val.detach() * images.shape[0]

    def get_trainable_parameters(self):
        return self.covar.grad_lr_factor()

    def setup_training(self, lr: Optional[float] = None, momentum: float = 0.9, optim_type: str = "Adam",
                       reg: float = 1, nufft_disc: str = "bilinear", orthogonal_projection: bool = False,
                       objective_func: str = "ml", form: Optional[str] = None,
                       precision: int = ml_ops.PRECISION_BF16X3) -> None:
        """covar_sgd.py:302-355; ``form``/``precision`` select the ML implementation."""
        self.use_orthogonal_projection = orthogonal_projection
        if lr is None:
            lr = 1e-1 if optim_type == "Adam" else 1e-2
        self.lr = lr * self.batch_size
        self.optim_type = optim_type
        self.momentum = momentum
        self.form = form or _DEFAULT_FORM
        self.precision = precision
        self.restart_optimizer()
        vol_shape = (self.covar.resolution,) * 3
        self.optimize_in_fourier_domain = nufft_disc != "nufft"
        self.objective_func = objective_func
        if self.optimize_in_fourier_domain:
            self.nufft_plans = NufftPlanDiscretized(vol_shape, upsample_factor=self.covar.upsampling_factor,
                                                    mode=nufft_disc)
            self.dataset.to_fourier_domain()
            self.covar._in_spatial_domain = False
            self.cost_func = cost_fourier_domain if objective_func == "ls" else cost_maximum_liklihood_fourier_domain
        else:
            self.nufft_plans = NufftPlan(vol_shape, batch_size=self.covar.rank, dtype=self.covar.dtype,
                                         device=self.device)
            self.dataset.to_spatial_domain()
            self.covar._in_spatial_domain = True
            self.cost_func = cost if objective_func == "ls" else cost_maximum_liklihood
        self.covar.init_grid_correction(nufft_disc)
        self.reg_scale = reg / len(self.dataset)
        self.epoch_index = 0

    def restart_optimizer(self) -> None:
        groups = self.get_trainable_parameters()
        for g in groups:
            g["lr"] *= self.lr
        if self.optim_type == "SGD":
            self.optimizer = torch.optim.SGD(groups, momentum=self.momentum)
        else:
            # the fused Adam (one kernel per step for all parameters) on the device, the same update
            on_device = all(p.device.type == "cuda" for g in groups for p in g["params"])
            self.optimizer = torch.optim.Adam(groups, fused=on_device or None)
        self.scheduler = torch.optim.lr_scheduler.ReduceLROnPlateau(self.optimizer, patience=self.scheduler_patiece)

    def train_epochs(self, max_epochs: int, restart_optimizer: bool = False) -> None:
        if restart_optimizer:
            self.restart_optimizer()
        for _ in range(max_epochs):
            t0 = time.time()
            self.run_epoch(self.epoch_index)
            t1 = time.time()
            logger.info(f"Epoch runtime: {t1 - t0:.2f} seconds")
            self.scheduler.step(self.cost_in_epoch)
            mask = getattr(self.dataset, "mask", None)
            if mask is not None:
                with torch.no_grad():
                    self.covar.vectors.data.copy_(self.covar.vectors.data * (mask.to(self.device) > 0.3))
            self.training_log["epoch_ind"].append(self.epoch_index)
            self.training_log["epoch_run_time"].append(t1 - t0)
            self.training_log["cost_val"].append(float(self.cost_in_epoch))
            self.training_log["lr_history"].append(self.scheduler.get_last_lr()[0])
            if self.save_path is not None:
                self.save_result()
            self.epoch_index += 1
            if self.scheduler.get_last_lr()[0] <= self.lr * (self.scheduler.factor**self.num_reduced_lr_before_stop):
                break

    def train(self, max_epochs: int, **training_kwargs: Any) -> None:
        self.setup_training(**training_kwargs)
        self.train_epochs(max_epochs)
        self.complete_training()

    def complete_training(self) -> None:
        if self.optimize_in_fourier_domain:
            self.dataset.to_spatial_domain()

    def compute_fourier_reg_term(self, eigenvecs: torch.Tensor) -> None:
        """noise_var / upsampled shell PSD of the current eigenvectors (covar_sgd.py:420-430)."""
        from .fsc_utils import rpsd, upsample_and_expand_fourier_shell

        us = self.covar.upsampling_factor
        eig_rpsd = rpsd(*eigenvecs) * (us**3)
        self.fourier_reg = self.noise_var / upsample_and_expand_fourier_shell(
            eig_rpsd, self.covar.resolution * us, 3
        )

    def update_fourier_reg_halfsets(self, fourier_reg: torch.Tensor) -> None:
        """covar_sgd.py:432-446: resample the L-grid regulariser onto the upsampled grid in Fourier mode."""
        from .fsc_utils import average_fourier_shell, upsample_and_expand_fourier_shell

        fourier_reg = fourier_reg.to(self.device)
        if self.optimize_in_fourier_domain:
            us = self.covar.upsampling_factor
            radial = average_fourier_shell(fourier_reg) / (us**3)
            fourier_reg = upsample_and_expand_fourier_shell(radial, self.covar.resolution * us, 3)
        self.fourier_reg = fourier_reg

    @property
    def vectorsGT(self) -> Optional[torch.Tensor]:
        """Ground-truth eigenvectors when ``gt_data`` carries them (covar_sgd.py:158-172)."""
        gt = self.gt_data
        return getattr(gt, "eigenvecs", None) if gt is not None else None

    def results_dict(self) -> dict:
        """The saved result, laid out as the reference's ``training_results.bin``
        (covar_sgd.py:524-535): the Covar state dict (``vectors``, ``log_sqrt_eigenvals``,
        ``_in_spatial_domain``, ``grid_correction``), ``vectorsGT``, ``fourier_reg`` and the
        training-log lists at the top level.  Tensors are moved to the host, so the file loads
        with ``torch.load(path, weights_only=True)`` on any machine."""
        def host(x):
            return x.detach().cpu() if isinstance(x, torch.Tensor) else x

        ckp = {k: host(v) for k, v in self.covar.state_dict().items()}
        ckp["vectorsGT"] = host(self.vectorsGT)
        ckp["fourier_reg"] = host(self.fourier_reg)
        ckp.update({k: [host(x) for x in v] if isinstance(v, list) else host(v) for k, v in self.training_log.items()})
        return ckp

    def save_result(self) -> None:
        """torch.save of ``results_dict()`` to ``save_path``, creating its folder (covar_sgd.py:537-545)."""
        import os

        folder = os.path.split(self.save_path)[0]
        if folder:
            os.makedirs(folder, exist_ok=True)
        torch.save(self.results_dict(), self.save_path)


def _project_mean_out(eigenvecs: torch.Tensor, mean: torch.Tensor) -> torch.Tensor:
    """Removes the mean direction from the eigenvolumes and re-orthogonalises them (utils.py:303-317)."""
    shape = eigenvecs.shape
    r = shape[0]
    v = eigenvecs.reshape(r, -1)
    m = mean.reshape(1, -1)
    v = v - (v @ m.conj().T) @ m / torch.norm(m) ** 2
    _, S, V = torch.linalg.svd(v, full_matrices=False)
    return (S.reshape(-1, 1) * V).reshape(shape)


class CovarPoseTrainer(CovarTrainer):
    """Covariance SGD with pose refinement (covar_sgd.py:548-1118).

    Each step (``run_batch``, covar_sgd.py:822-891) takes the batch's rotations, offsets (and
    contrasts) from a ``PoseModule``, preprocesses the images with them (phase shift, projected
    mean subtraction, soft projected mask) and evaluates the covariance objective with the
    preprocessed images, so one backward pass gives the gradient of the eigenvolumes and of the
    rotation vectors (through the NUFFT's point gradient); Adam updates the eigenvolumes and
    SparseAdam the rotation vectors of the batch (lr ratio 3).  Every ``mean_update_frequency``
    epochs the offsets are re-estimated per image by Newton's method on the ML objective, the
    dataset takes the refined poses and the mean is reconstructed from half-sets.

    The objective runs in its autograd form here (``form="autograd"``): its projections return the
    point gradient the rotations need, through the NUFFT's point derivative (``nufft_disc="nufft"``)
    or the trilinear slice's grid derivative (``"bilinear"``, the reference's default; ``"nearest"``
    has none, as with grid_sample).
    """

    def __init__(self, covar: Covar, train_data, device: torch.device, mean, pose, save_path: Optional[str] = None,
                 gt_data=None, training_log_freq: int = 50) -> None:
        super().__init__(covar, train_data, device, save_path, gt_data, training_log_freq)
        self.mean = mean.to(self.device)
        self.pose = pose.to(self.device)
        self.pose_lr_ratio = 3
        self.num_rep = 1
        self.set_pose_grad_req(True)
        self._updated_idx = torch.zeros(len(self.dataset), device=self.device)
        self.mean_fourier_reg = None
        self.scheduler_patiece = 3
        self.downsample_factor = min(3, int(np.log2(self.covar.resolution / 32)))
        self.mean_est_method = "Reconstruction"
        self.offset_est_method = "Newton"
        self.rotation_est_method = "SGD"
        self.mean_update_frequency = 5
        self.use_eigen_in_contrast_from_epoch = 0
        for prm in self.mean.parameters():
            prm.requires_grad = False
        for prm in self.pose.offsets.parameters():
            prm.requires_grad = False
        if self.pose.use_contrast:
            for prm in self.pose.contrasts.parameters():
                prm.requires_grad = False

    @property
    def downsample_size(self) -> int:
        return int(2 ** (-self.downsample_factor) * self.pose.resolution)

    def get_mean_module(self):
        return self.mean

    def get_pose_module(self):
        return self.pose

    def set_pose_grad_req(self, grad: bool) -> None:
        for prm in self.pose.parameters():
            prm.requires_grad = grad
        for prm in self.mean.parameters():
            prm.requires_grad = grad

    def get_trainable_parameters(self):
        return super().get_trainable_parameters() + [{"params": self.mean.parameters(), "lr": 1}]

    def prepare_batch(self, batch):
        images, _, filters, idx = batch
        if isinstance(idx, slice):
            idx = torch.arange(idx.start or 0, idx.stop, device=self.device)
        return images.to(self.device), idx.to(self.device), filters.to(self.device)

    def run_batch(self, images: torch.Tensor, idx: torch.Tensor, filters: torch.Tensor) -> torch.Tensor:
        from .projection_funcs import lowpass_volume, preprocess_image_batch

        for _ in range(self.num_rep):
            self.optimizer.zero_grad()
            self.pose_optimizer.zero_grad()
            if self.pose.use_contrast:
                pts_rot, phase_shift, contrasts = self.pose(idx)
                filters = filters * contrasts.reshape(-1, 1, 1)
            else:
                pts_rot, phase_shift = self.pose(idx)
            pre = preprocess_image_batch(images, self.nufft_plans, filters, (pts_rot, phase_shift), self.mean(None),
                                         self.mask, self.mask_threshold, self.softening_kernel_fourier,
                                         fourier_domain=self.optimize_in_fourier_domain)
            kw = ({"form": "autograd"} if self.cost_func in (cost_maximum_liklihood,
                                                              cost_maximum_liklihood_fourier_domain) else {})
            cost_val = self.cost_func(self._covar(None), pre, self.nufft_plans, filters, self.noise_var,
                                      self.reg_scale, self.fourier_reg, apply_mean_const_term=True, **kw)
            cost_val.backward()
            self.optimizer.step()
            self.pose_optimizer.step()
        if self.use_orthogonal_projection:
            self.covar.orthogonal_projection()
        with torch.no_grad():
            self.covar.vectors.data.copy_(lowpass_volume(self.covar.vectors.data, self.downsample_size // 2))
        self._updated_idx[idx] = 1
        return cost_val

    def run_epoch(self, epoch: int) -> None:
        super().run_epoch(epoch)
        if self.pose.use_contrast:
            with torch.no_grad():
                v = _project_mean_out(self.covar.get_vectors_spatial_domain(), self.mean.get_volume_spatial_domain())
                self.covar.vectors.data.copy_(v / torch.exp(self.covar.log_sqrt_eigenvals).reshape(-1, 1, 1, 1))
        if epoch % self.mean_update_frequency == self.mean_update_frequency - 1:
            if self.offset_est_method == "Newton":
                self.correct_offsets()
            if self.mean_est_method == "Reconstruction":
                from .mean import reconstruct_mean_from_halfsets

                self.dataset.update_pose(self.pose)
                self.mean.set_mean(reconstruct_mean_from_halfsets(self.dataset, mask=self.mean.volume_mask))

    def correct_offsets(self) -> None:
        """Newton re-estimation of every image's offset on the ML objective with the current mean,
        eigenvolumes and rotations (covar_sgd.py:703-783)."""
        from .poses import estimate_image_offsets_newton
        from .projection_funcs import centered_fft2, centered_ifft2, vol_forward

        for batch in self.train_data:
            images, idx, filters = self.prepare_batch(batch)
            with torch.no_grad():
                pts = self.pose(idx)[0]
                self.nufft_plans.setpts(pts)
                fourier = self.optimize_in_fourier_domain
                mean_fwd = vol_forward(self.mean(None), self.nufft_plans, filters, fourier_domain=fourier).squeeze(1)
                proj = vol_forward(self._covar(None), self.nufft_plans, filters, fourier_domain=fourier)
                soft = None
                if self.mask is not None:
                    mf = vol_forward(self.mask, self.nufft_plans, filters=None).squeeze(1)
                    soft = centered_ifft2(centered_fft2((mf > self.mask_threshold).to(mf.dtype)) *
                                          self.softening_kernel_fourier).real

            def obj(shifted):
                if not fourier:
                    shifted = centered_ifft2(shifted).real
                return self.raw_cost_func(proj, shifted - mean_fwd, self.noise_var, apply_mean_const_term=True,
                                          mean_aggregate=False)

            off = estimate_image_offsets_newton(images, mean_fwd, mask=soft, init_offsets=self.pose.get_offsets()[idx],
                                                in_fourier_domain=fourier, obj_func=obj)
            self.pose.set_offsets(off, idx=idx)

    def setup_training(self, **kwargs) -> None:
        from .projection_funcs import get_mask_threshold
        from .utils import soft_edged_kernel

        super().setup_training(**kwargs)
        self.mean.init_grid_correction(kwargs.get("nufft_disc"))
        self.mean._in_spatial_domain = not self.optimize_in_fourier_domain
        self.mask = self.mean.get_volume_mask()
        # the per-image objectives of the offset search (generic in the domain, as the reference uses them)
        self.raw_cost_func = (raw_cost_fourier_domain if kwargs.get("objective_func") == "ls"
                              else raw_cost_maximum_liklihood_fourier_domain)
        self.softening_kernel_fourier = soft_edged_kernel(radius=5, L=self.mean.resolution, dim=2,
                                                          in_fourier=True).to(self.device)
        self.mask_threshold = None
        if self.mask is not None:
            with torch.no_grad():
                idx = torch.arange(min(self.batch_size, len(self.dataset)), device=self.device)
                self.nufft_plans.setpts(self.pose(idx)[0])
                self.mask_threshold = get_mask_threshold(self.mask, self.nufft_plans)

    def restart_optimizer(self) -> None:
        super().restart_optimizer()
        if self.rotation_est_method == "SGD":
            self.pose_optimizer = torch.optim.SparseAdam([{"params": list(self.pose.parameters()),
                                                           "lr": self.pose_lr_ratio * self.lr}])
        elif self.rotation_est_method == "LBFGS":
            # covar_sgd.py:980-981: a row-wise L-BFGS on the rotation vectors, steps limited to 0.1
            self.pose_optimizer = BlockwiseLBFGS([{"params": list(self.pose.parameters()), "lr": 1,
                                                   "step_size_limit": 1e-1}])
        else:
            raise ValueError(f"unknown rotation_est_method {self.rotation_est_method!r}")

    def set_pose_module(self, pose) -> None:
        """Takes a new pose module (covar_sgd.py:645-658), e.g. the one merged from two half-set
        modules; offsets and contrasts stay out of the gradient as in ``__init__``."""
        self.pose = pose.to(self.device)
        for prm in self.pose.offsets.parameters():
            prm.requires_grad = False
        if self.pose.use_contrast:
            for prm in self.pose.contrasts.parameters():
                prm.requires_grad = False
        self._updated_idx = torch.zeros(len(self.dataset), device=self.device)

    def train_epochs(self, max_epochs: int, **training_kwargs) -> None:
        while self.downsample_factor >= 0:
            super().train_epochs(max_epochs, **training_kwargs)
            self.downsample_factor -= 1
        self.downsample_factor = 0

    def results_dict(self) -> dict:
        ckp = super().results_dict()
        ckp["mean"] = {k: v.detach().cpu() if isinstance(v, torch.Tensor) else v
                       for k, v in self.mean.state_dict().items()}
        ckp["pose"] = {k: v.detach().cpu() for k, v in self.pose.state_dict().items()}
        return ckp


def load_training_results(path: str) -> dict:
    """Reads a ``training_results.bin`` written by ``CovarTrainer.save_result`` (tensors, lists,
    numbers, None only), with the loader that executes nothing from the file."""
    return torch.load(path, map_location="cpu", weights_only=True)


def covar_from_results(results: dict, dtype: torch.dtype = torch.float32) -> Covar:
    """Rebuilds the ``Covar`` module from a results dictionary (its state-dict entries)."""
    keys = ("vectors", "log_sqrt_eigenvals", "_in_spatial_domain", "grid_correction")
    v = results["vectors"]
    covar = Covar(v.shape[-1], v.shape[0], dtype=dtype)
    covar.load_state_dict({k: results[k] for k in keys})
    return covar


def trainCovar(covar_model, dataset, batch_size, optimize_pose=False, mean_model=None, pose=None, savepath=None,
               gt_data=None, **kwargs):
    """Entry point (covar_sgd.py:1479-1612): optional regulariser rounds, then a final training pass."""
    from .dataset import create_dataloader

    use_halfsets = kwargs.pop("use_halfsets")
    num_reg_update_iters = kwargs.pop("num_reg_update_iters", None) or 0
    num_epochs = kwargs.pop("max_epochs")
    device = covar_model.device
    if not use_halfsets:
        loader = create_dataloader(dataset, batch_size=batch_size)
        if optimize_pose:
            trainer = CovarPoseTrainer(covar_model, loader, device, mean_model, pose, savepath, gt_data=gt_data)
        else:
            trainer = CovarTrainer(covar_model, loader, device, savepath, gt_data=gt_data)
        trainer.setup_training(**kwargs)
        for _ in range(num_reg_update_iters):
            trainer.train_epochs(num_epochs, restart_optimizer=True)
            ev, el = trainer.covar.eigenvecs
            trainer.compute_fourier_reg_term(ev * (el**0.5).reshape(-1, 1, 1, 1))
            trainer.covar.orthogonal_projection()
        trainer.train_epochs(num_epochs, restart_optimizer=True)
        trainer.complete_training()
        return trainer
    covar_copy = copy.deepcopy(covar_model)
    with torch.no_grad():
        covar_copy.set_vectors(torch.randn_like(covar_copy.get_vectors_spatial_domain()) *
                               float(covar_model.pixel_var_estimate) ** 0.5)
    half1, half2, permutation = dataset.half_split()
    dl1 = create_dataloader(half1, batch_size=batch_size)
    dl2 = create_dataloader(half2, batch_size=batch_size)
    if not optimize_pose:
        t1 = CovarTrainer(covar_model, dl1, device, savepath, gt_data=gt_data)
        t2 = CovarTrainer(covar_copy, dl2, device, None, gt_data=gt_data)
    else:
        # covar_sgd.py:1557-1572: each half-set refines the poses of its own images, with its own
        # copy of the mean
        pose1, pose2 = pose.split_module(permutation)
        gt1, gt2 = gt_data.half_split(permutation) if hasattr(gt_data, "half_split") else (gt_data, gt_data)
        t1 = CovarPoseTrainer(covar_model, dl1, device, mean_model, pose1, savepath, gt_data=gt1)
        t2 = CovarPoseTrainer(covar_copy, dl2, device, copy.deepcopy(mean_model), pose2, None, gt_data=gt2)
    t1.setup_training(**kwargs)
    t2.setup_training(**kwargs)
    for _ in range(num_reg_update_iters):
        t1.train_epochs(num_epochs, restart_optimizer=True)
        t2.train_epochs(num_epochs, restart_optimizer=True)
        update_fourier_reg(t1, t2)
    t2.complete_training()
    t1.train_data = create_dataloader(dataset, batch_size=batch_size)
    if optimize_pose:
        # covar_sgd.py:1603-1607: the full set continues from the two half-sets' poses
        pose.update_from_modules(pose1, pose2, permutation)
        t1.set_pose_module(pose)
        t1.gt_data = gt_data
    t1.train(max_epochs=num_epochs, **kwargs)
    return t1


def update_fourier_reg(trainer1: CovarTrainer, trainer2: CovarTrainer) -> None:
    """Half-set shell regulariser from the covariance FSC (covar_sgd.py:1121-1156).  The gain is
    the mean of the two half-sets' shell x shell covariance gains (the trainer's filter_gain,
    covar_sgd.py:175-181, dataset.py:358-421)."""
    L = trainer1.covar.resolution
    gain = (trainer1.dataset.get_total_covar_gain() + trainer2.dataset.get_total_covar_gain()) / 2
    e1, l1 = trainer1.covar.eigenvecs
    e2, l2 = trainer2.covar.eigenvecs
    reg, fsc = compute_updated_fourier_reg(e1 * (l1**0.5).reshape(-1, 1, 1, 1), e2 * (l2**0.5).reshape(-1, 1, 1, 1),
                                           gain, getattr(trainer1, "fourier_reg", None), L,
                                           trainer1.optimize_in_fourier_domain, mask=trainer1.dataset.mask)
    trainer1.update_fourier_reg_halfsets(reg)
    trainer2.update_fourier_reg_halfsets(reg)
    trainer1.training_log["covariance_fsc_halfset"] = fsc


def compute_updated_fourier_reg(eigenvecs1, eigenvecs2, filter_gain, current_fourier_reg, L,
                                optimize_in_fourier_domain, mask=None):
    """reg = 1 / ((FSC / (1 - FSC)) * shell gain^-1), rank-1 shell approximation (covar_sgd.py:1159-1187).

    ``filter_gain`` is the (S, S) shell x shell covariance gain; its inverse is held constant
    beyond shell L // 2 (both axes).  ``current_fourier_reg`` is accepted for the reference's
    signature and, as there, does not enter the result."""
    corr = 1 / filter_gain
    corr[:, L // 2:] = corr[L // 2, L // 2]
    corr[L // 2:, :] = corr[L // 2, L // 2]
    if mask is not None:
        mask = mask.to(eigenvecs1.device)
        eigenvecs1 = eigenvecs1 * mask
        eigenvecs2 = eigenvecs2 * mask
    fsc = covar_fsc(eigenvecs1, eigenvecs2)
    fsc = fsc.clamp(1e-6, 1 - 1e-6)
    reg = 1 / ((fsc / (1 - fsc)) * corr)
    reg[reg < 0] = 0
    reg = expand_fourier_shell(reg.diag().sqrt().unsqueeze(0), L, 3)
    if not optimize_in_fourier_domain:
        reg = reg / L**2
    return reg, fsc
