"""The saved training result (covar_sgd.py:524-545, covar.py:82-112): the reference's
``training_results.bin`` layout, loadable without executing anything from the file."""

import torch

from tests.conftest import pkg


def test_training_results_layout_and_round_trip(tmp_path):
    dsm = pkg("dataset")
    covar_mod = pkg("covar")
    sgd = pkg("covar_sgd")
    L, r, B = 8, 2, 6
    torch.manual_seed(0)
    ds = dsm.CovarDataset(torch.randn(B, L, L), torch.eye(3).repeat(B, 1, 1), 1.0, filters=torch.ones(B, L, L))
    covar = covar_mod.Covar(L, r)

    class GT:
        eigenvecs = torch.randn(r, L, L, L)

    path = tmp_path / "out" / "training_results.bin"
    trainer = sgd.CovarTrainer(covar, dsm.create_dataloader(ds, batch_size=4), "cpu", save_path=str(path), gt_data=GT)
    trainer.fourier_reg = torch.rand(L, L, L)
    trainer.training_log["cost_val"] += [3.0, 2.5]
    trainer.training_log["epoch_ind"] += [0, 1]
    trainer.training_log["covariance_fsc_halfset"] = torch.rand(5, 5)
    trainer.save_result()
    res = sgd.load_training_results(str(path))
    for k in ("vectors", "log_sqrt_eigenvals", "_in_spatial_domain", "grid_correction", "vectorsGT", "fourier_reg",
              "cost_val", "epoch_ind", "lr_history", "epoch_run_time", "covariance_fsc_halfset"):
        assert k in res, k
    assert res["cost_val"] == [3.0, 2.5]
    assert torch.equal(res["vectorsGT"], GT.eigenvecs)
    assert torch.equal(res["fourier_reg"], trainer.fourier_reg)
    back = sgd.covar_from_results(res)
    assert torch.equal(back.vectors, covar.vectors) and torch.equal(back.log_sqrt_eigenvals, covar.log_sqrt_eigenvals)
    assert torch.allclose(back.eigenvecs[0], covar.eigenvecs[0]) and torch.allclose(back.eigenvecs[1], covar.eigenvecs[1])
