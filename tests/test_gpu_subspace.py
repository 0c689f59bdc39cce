"""Eigen-subspace agreement of the chosen SGD form (split-BF16 contractions, hand
adjoint, Hermitian half plane) and the plain FP32 autograd form with an FP64
trajectory (FP64 parameters and images, FP64 NUFFT at eps 1e-12), from the same
start on the same batches.  Bar: RMS sine of the principal angles <= 1e-3."""

import pytest
import torch

from tests.conftest import pkg

pytestmark = pytest.mark.gpu


def test_chosen_form_subspace_matches_fp64_trajectory(dev):
    ds = pkg("dataset")
    val = pkg("validation")
    cfg = ds.config_by_name("L32")  # L=32, 2000 images, r=4, SNR 0.1
    data = ds.make_synthetic(cfg, dev)
    signal_var = max(float(data.dataset.images.var()) - data.dataset.noise_var, 1e-6)
    v0 = torch.randn((cfg.rank,) + (cfg.L,) * 3, generator=torch.Generator().manual_seed(0)) * (signal_var / cfg.L) ** 0.5
    bases, secs = val.sgd_subspaces(data.dataset, cfg.rank, v0, batch_size=256, epochs=3,
                                    forms=("fp64", "chosen", "fused_fp32", "plain_fp32"))
    err = val.subspace_errors(bases)
    print("subspace error vs FP64:", err, "seconds:", secs)
    assert err["chosen"] < 1e-3
    assert err["fused_fp32"] < 1e-4
    assert err["plain_fp32"] < 1e-4
    # the trajectories moved: the learned subspace differs from the start
    q0, _ = torch.linalg.qr(v0.reshape(cfg.rank, -1).double().T.to(dev))
    assert val.subspace_distance(q0, bases["fp64"]) > 1e-2

