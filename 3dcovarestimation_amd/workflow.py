"""The on-disk summary of a covariance run (cov3d/workflow.py:474-518): ``recorded_data.pkl``.

``record_data`` computes the latent coordinates of every image with the estimated eigenvolumes
(and with ground-truth eigenvolumes when given) through ``latentMAP`` and writes the dictionary
the reference's ``analyze`` step reads, with the reference's keys and numpy values:

    eigen_est, eigenval_est, coords_est, coords_covar_inv_est, mean_est, particles_path, ctf_path,
    poses_path, ind, data_sign_inverted, lazy
    [+ eigenvals_GT, eigenvectors_GT, coords_GT, coords_covar_inv_GT with ground truth]

Only writing is done here; the file holds numpy arrays, strings, booleans and None.
"""

from __future__ import annotations

import os
import pickle
from typing import Optional, Sequence

import numpy as np
import torch

from .wiener_coords import latentMAP


def _abspath(p: Optional[str]) -> Optional[str]:
    return os.path.abspath(p) if p is not None else None


def record_data(output_dir: str, covar, dataset, mean_volume: Optional[torch.Tensor] = None,
                gt_eigenvecs: Optional[torch.Tensor] = None, input_paths: Sequence[Optional[str]] = (None,) * 3,
                indices: Optional[np.ndarray] = None, data_sign_inverted: bool = False, lazy: bool = False,
                batch_size: int = 1024, filename: str = "recorded_data.pkl") -> dict:
    """Writes ``output_dir/recorded_data.pkl`` (workflow.py:474-518) and returns the dictionary.

    ``covar``: the trained ``Covar``; its eigen-decomposition gives eigen_est / eigenval_est.
    ``gt_eigenvecs``: ground-truth eigenvolumes scaled by the square roots of their eigenvalues,
    (r, L^3) or (r, L, L, L), as the reference's GTData carries them; their norms squared are the
    ground-truth eigenvalues.
    """
    L = covar.resolution
    eigen_est, eigenval_est = covar.eigenvecs
    dev = eigen_est.device
    coords_est, coords_covar_inv_est = latentMAP(dataset, eigen_est, eigenval_est.to(dev), batch_size=batch_size,
                                                 return_coords_covar=True)
    n = len(dataset)
    data = {
        "eigen_est": eigen_est.detach().cpu().numpy(),
        "eigenval_est": eigenval_est.detach().cpu().numpy(),
        "coords_est": coords_est.cpu().numpy(),
        "coords_covar_inv_est": coords_covar_inv_est.cpu().numpy(),
        "mean_est": mean_volume.detach().cpu().numpy() if mean_volume is not None else None,
        "particles_path": _abspath(input_paths[0]),
        "ctf_path": _abspath(input_paths[1]),
        "poses_path": _abspath(input_paths[2]),
        "ind": np.arange(n) if indices is None else np.asarray(indices),
        "data_sign_inverted": bool(data_sign_inverted),
        "lazy": bool(lazy),
    }
    if gt_eigenvecs is not None:
        gt = gt_eigenvecs.reshape(gt_eigenvecs.shape[0], -1).to(dev, torch.float32)
        eigenvals_gt = torch.norm(gt, dim=1) ** 2
        eigenvectors_gt = (gt / torch.sqrt(eigenvals_gt).unsqueeze(1)).reshape((-1, L, L, L))
        coords_gt, coords_covar_inv_gt = latentMAP(dataset, eigenvectors_gt, eigenvals_gt, batch_size=batch_size,
                                                   return_coords_covar=True)
        data.update({
            "eigenvals_GT": eigenvals_gt.cpu().numpy(),
            "eigenvectors_GT": eigenvectors_gt.cpu().numpy(),
            "coords_GT": coords_gt.cpu().numpy(),
            "coords_covar_inv_GT": coords_covar_inv_gt.cpu().numpy(),
        })
    os.makedirs(output_dir, exist_ok=True)
    with open(os.path.join(output_dir, filename), "wb") as fid:
        pickle.dump(data, fid)
    return data
