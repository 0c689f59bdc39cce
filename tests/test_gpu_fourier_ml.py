"""The fused Fourier-domain ML objective on the discretised projection (ml_ops.MLCostDiscretized:
native slice gather, fused Gram + FP64 Cholesky, hand adjoint, native slice scatter) against the
reference's torch formulation differentiated by autograd (cov3d/covar_sgd.py:1350-1399), through
the Covar's Fourier-domain forward (covar.py:440-446) so that the gradient reaches the real
eigenvolume parameters; plus the recorded_data.pkl writer (workflow.py:474-518)."""

import pickle

import numpy as np
import pytest
import torch

from tests.conftest import pkg, rel_l2

pytestmark = pytest.mark.gpu


def _fourier_batch(dev, L, r, B, seed):
    ds = pkg("dataset")
    cfg = ds.SyntheticConfig(L, B, r, list(np.geomspace(4, 0.5, r)), 3.0, 0.2, (0.8, 2.5), shift_px=0.0, seed=seed)
    data = ds.make_synthetic(cfg, dev, batch=B)
    data.dataset.to_fourier_domain()
    ims, pts, filt, _ = data.dataset[0:B]
    return data, ims, pts, filt


@pytest.mark.parametrize("L,r,B,mode,us", [(16, 4, 32, "bilinear", 2), (17, 3, 24, "nearest", 2),
                                           (16, 4, 32, "bilinear", 1), (64, 10, 64, "bilinear", 2)])
@pytest.mark.parametrize("precision,tol", [(0, 1e-4), (1, 1e-3)])
@pytest.mark.parametrize("half", [True, False])
def test_fused_fourier_ml_matches_autograd(dev, L, r, B, mode, us, precision, tol, half):
    """The fused Fourier-domain ML cost (full plane, or half plane with the boundary pairs exact)
    against the reference's autograd formulation: cost and gradients of the real eigenvolumes and
    eigenvalues."""
    sgd = pkg("covar_sgd")
    cv = pkg("covar")
    npl = pkg("nufft_plan")
    data, ims, pts, filt = _fourier_batch(dev, L, r, B, 40 + L + us)
    nv = data.dataset.noise_var
    plan = npl.NufftPlanDiscretized((L,) * 3, upsample_factor=us, mode=mode)
    plan.setpts(pts)
    torch.manual_seed(7)
    base = cv.Covar(L, r, upsampling_factor=us, pixel_var_estimate=0.05).to(dev)
    out = {}
    for form in ("fused", "autograd"):
        covar = cv.Covar(L, r, upsampling_factor=us, vectors=base.get_vectors_spatial_domain().detach()).to(dev)
        covar.init_grid_correction(mode)
        cost = sgd.cost_maximum_liklihood_fourier_domain(covar.get_vectors_fourier_domain(), ims, plan, filt, nv,
                                                         form=form, precision=precision,
                                                         apply_mean_const_term=True, half_plane=half)
        cost.backward()
        out[form] = (cost.item(), covar.vectors.grad.cpu().numpy(), covar.log_sqrt_eigenvals.grad.cpu().numpy())
    ec = abs(out["fused"][0] - out["autograd"][0]) / abs(out["autograd"][0])
    eg = rel_l2(out["fused"][1], out["autograd"][1])
    el = rel_l2(out["fused"][2], out["autograd"][2])
    print(f"L={L} r={r} {mode} us={us} precision={precision} half={half}: cost {ec:.2e} grad {eg:.2e} "
          f"eigenvalue grad {el:.2e}")
    assert ec < tol and eg < tol and el < tol


def test_fused_fourier_trainer_matches_autograd_trainer(dev):
    """Two epochs of the bilinear trainer (setup_training(nufft_disc="bilinear")) with the fused and
    the autograd Fourier-domain objective from the same start land on the same subspace."""
    sgd = pkg("covar_sgd")
    cv = pkg("covar")
    ds = pkg("dataset")
    val = pkg("validation")
    cfg = ds.SyntheticConfig(16, 2048, 3, [4.0, 2.0, 1.0], 3.0, 1.0, (0.8, 2.5), seed=77)
    data = ds.make_synthetic(cfg, dev)
    torch.manual_seed(8)
    v0 = torch.randn((3, 16, 16, 16)) * 0.05
    bases = {}
    for form in ("fused", "autograd"):
        covar = cv.Covar(16, 3, vectors=v0).to(dev)
        tr = sgd.CovarTrainer(covar, ds.create_dataloader(data.dataset, batch_size=256), dev)
        tr.setup_training(nufft_disc="bilinear", objective_func="ml", lr=1e-5, form=form, precision=0)
        tr.train_epochs(2)
        tr.complete_training()
        bases[form] = val.basis(covar)
    d = val.subspace_distance(bases["autograd"], bases["fused"])
    print("fused vs autograd bilinear trainer subspace distance", d)
    assert d < 1e-3


def test_record_data_writes_the_reference_layout(dev, tmp_path):
    wf = pkg("workflow")
    cv = pkg("covar")
    ds = pkg("dataset")
    cfg = ds.SyntheticConfig(16, 300, 2, [4.0, 1.0], 3.0, 1.0, (0.8, 2.5), seed=78)
    data = ds.make_synthetic(cfg, dev)
    gt = data.eigenvecs * data.eigenvals.sqrt().reshape(-1, 1, 1, 1)
    covar = cv.Covar(16, 2, vectors=gt.cpu()).to(dev)
    out = wf.record_data(str(tmp_path), covar, data.dataset, mean_volume=data.mean, gt_eigenvecs=gt,
                         input_paths=("p.mrcs", None, "poses.pkl"), batch_size=128)
    with open(tmp_path / "recorded_data.pkl", "rb") as fid:  # written by this test's own call
        saved = pickle.load(fid)
    keys = {"eigen_est", "eigenval_est", "coords_est", "coords_covar_inv_est", "mean_est", "particles_path",
            "ctf_path", "poses_path", "ind", "data_sign_inverted", "lazy", "eigenvals_GT", "eigenvectors_GT",
            "coords_GT", "coords_covar_inv_GT"}
    assert set(saved) == keys and set(out) == keys
    assert saved["coords_est"].shape == (300, 2) and saved["coords_covar_inv_est"].shape == (300, 2, 2)
    assert saved["ctf_path"] is None and saved["particles_path"].endswith("p.mrcs")
    assert np.allclose(np.sort(saved["eigenval_est"]), np.sort(saved["eigenvals_GT"]), rtol=1e-4)
    # the estimate is the ground truth here, so the coordinates agree up to each eigenvector's sign
    for k in range(2):
        a, b = saved["coords_est"][:, k], saved["coords_GT"][:, k]
        assert min(rel_l2(a, b), rel_l2(-a, b)) < 1e-3
