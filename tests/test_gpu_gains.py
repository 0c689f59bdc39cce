"""Gain tensors (dataset.py:322-421) and the half-set regulariser that consumes them
(covar_sgd.py:1121-1187, :1479-1612)."""

import numpy as np
import pytest
import torch

from tests.conftest import oracle, pkg

pytestmark = pytest.mark.gpu


def _dataset(dev, L, B, rots, seed=0):
    o = oracle()
    dsm = pkg("dataset")
    rng = np.random.default_rng(seed)
    params = np.tile(np.array([[12000.0, 15000.0, 0.4, 300.0, 2.7, 0.1, 3.0]]), (B, 1))
    params[:, 0] += rng.uniform(0, 8000, B)
    params[:, 1] = params[:, 0] + rng.uniform(0, 3000, B)
    ctf = o.ctf(params, L)
    ds = dsm.CovarDataset(torch.zeros((B, L, L), device=dev), torch.tensor(rots, dtype=torch.float32, device=dev), 1.0,
                          filters=torch.tensor(ctf, dtype=torch.float32, device=dev))
    return ds, ctf


@pytest.mark.parametrize("L", [16, 15])
def test_gains_on_the_central_plane(dev, L):
    """Identity poses put every pixel k on voxel (L//2, k_y + L//2, k_x + L//2): the total gain is
    sum_i CTF_i^2 / L^2 on that plane (Nyquist row and column zero for even L) and zero elsewhere;
    the shell gain is sum_i a_i a_i^T with a_i[s] = sum over pixels of 2-D radius in shell s of
    CTF_i^2 / L^2, divided by the 3-D shell size."""
    B = 24
    ds, ctf = _dataset(dev, L, B, np.tile(np.eye(3), (B, 1, 1)))
    f2 = ctf.astype(np.float64) ** 2
    if L % 2 == 0:
        f2[:, 0, :] = 0
        f2[:, :, 0] = 0
    gain = ds.get_total_gain(batch_size=10).double().cpu().numpy()
    expect = np.zeros((L,) * 3)
    expect[L // 2] = f2.sum(0) / L**2
    assert np.abs(gain - expect).max() < 1e-5 * expect.max()
    # shells: radius of the centred integer grid, shell floor(r + 0.5), capped at n - 1
    ax = np.arange(L) - L // 2
    r3 = np.sqrt(ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2)
    n = int(np.ceil(r3.max()))
    size = np.bincount(np.minimum(np.floor(r3 + 0.5), n - 1).astype(int).ravel(), minlength=n)
    r2 = np.sqrt(ax[:, None] ** 2 + ax[None, :] ** 2)
    s2 = np.minimum(np.floor(r2 + 0.5), n - 1).astype(int)
    A = np.stack([np.bincount(s2.ravel(), weights=(f2[i] / L**2).ravel(), minlength=n) for i in range(B)]) / size
    cg = ds.get_total_covar_gain(batch_size=7).double().cpu().numpy()
    ref = A.T @ A
    assert cg.shape == (n, n)
    assert np.abs(cg - ref).max() < 1e-5 * ref.max()


@pytest.mark.parametrize("L", [16, 15])
def test_covar_gain_matches_per_image_backprojections(dev, L):
    """The one-pass shell gain equals the reference's construction: per-image nearest-neighbour
    back-projections of CTF^2 (execute_adjoint_unaggregated), averaged over shells, summed as
    outer products; random poses, so pixels collide in voxels and leave the volume."""
    B = 40
    rng = np.random.default_rng(L)
    R = np.linalg.qr(rng.standard_normal((B, 3, 3)))[0]
    ds, _ = _dataset(dev, L, B, R, seed=L)
    npl = pkg("nufft_plan")
    fsc = pkg("fsc_utils")
    plan = npl.NufftPlanDiscretized((L,) * 3, upsample_factor=1, mode="nearest")
    ref = None
    for i in range(0, B, 16):
        _, pts, filters, _ = ds[slice(i, min(i + 16, B))]
        plan.setpts(pts)
        f = filters.clone()
        if L % 2 == 0:
            f[:, 0, :] = 0
            f[:, :, 0] = 0
        g = plan.execute_adjoint_unaggregated(torch.complex(f**2, torch.zeros_like(f))).abs() / L**2
        a = fsc.average_fourier_shell(*g).double()
        ref = a.T @ a if ref is None else ref + a.T @ a
    cg = ds.get_total_covar_gain(batch_size=13).double()
    assert torch.allclose(cg, ref, rtol=1e-5, atol=1e-6 * float(ref.max()))


def test_half_set_training_updates_the_regulariser(dev):
    """trainCovar with half-sets runs a regulariser round: the shell x shell gain of each half,
    the covariance FSC of the two halves' eigenvectors, a non-negative shell regulariser."""
    dsm = pkg("dataset")
    covar_mod = pkg("covar")
    sgd = pkg("covar_sgd")
    cfg = dsm.SyntheticConfig(L=16, n_images=1024, rank=2, eigenvalues=[4.0, 1.0], pixel_size=3.0, snr=1.0,
                              defocus_um=(1.0, 2.5), seed=3)
    data = dsm.make_synthetic(cfg, dev)
    torch.manual_seed(0)
    signal_var = float(data.dataset.images.var()) - data.dataset.noise_var
    covar = covar_mod.Covar(16, 2, pixel_var_estimate=max(signal_var, 1e-6) / 16).to(dev)
    trainer = sgd.trainCovar(covar, data.dataset, 256, use_halfsets=True, num_reg_update_iters=1, max_epochs=1,
                             nufft_disc="nufft", objective_func="ml", lr=1e-5)
    fsc = trainer.training_log["covariance_fsc_halfset"]
    n = pkg("fsc_utils").shell_index(16, 3, dev)[2]
    assert fsc.shape == (n, n)
    reg = trainer.fourier_reg
    assert reg is not None and reg.shape == (16, 16, 16)
    assert torch.isfinite(reg).all() and (reg >= 0).all()
