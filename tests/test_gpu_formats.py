"""A dataset written as cryoDRGN files and read back (formats.py; reference source.py:52-105,
:476-542): rotations, offsets and the per-image CTF images agree with the original dataset."""

import numpy as np
import pytest
import torch

from tests.conftest import pkg

pytestmark = pytest.mark.gpu


def test_cryodrgn_round_trip_dataset(dev, tmp_path):
    ds = pkg("dataset")
    fm = pkg("formats")
    L, n = 32, 40
    cfg = ds.SyntheticConfig(L, n, 2, [4.0, 1.0], 3.0, 1.0, (0.8, 2.5), shift_px=1.5, seed=9)
    data = ds.make_synthetic(cfg, dev, subtract_mean=False)
    d0 = data.dataset
    fm.save_poses(str(tmp_path / "poses.pkl"), d0.rotations.cpu().numpy(), d0.offsets.cpu().numpy(), L)
    # the writer stores the kernel's angle in degrees as it is (as the reference's simulation
    # writer does); store the angle that reads back to the same CTF: dfang = -ang - 90 deg
    p7 = d0.ctf_params.cpu().numpy().astype(np.float64).copy()
    p7[:, 2] = -p7[:, 2] - np.pi / 2
    fm.save_ctf(str(tmp_path / "ctf.pkl"), p7, L)
    idx = np.arange(3, n, 2)
    d1 = fm.covar_dataset_from_cryodrgn(d0.images, str(tmp_path / "poses.pkl"), str(tmp_path / "ctf.pkl"),
                                        d0.noise_var, indices=idx)
    t = torch.as_tensor(idx, device=dev)
    assert torch.allclose(d1.rotations, d0.rotations[t], atol=1e-7)
    assert torch.allclose(d1.offsets, d0.offsets[t], atol=1e-5)
    assert torch.equal(d1.images, d0.images[t])
    f0 = d0.get_filters(t)
    f1 = d1.get_filters(torch.arange(len(idx), device=dev))
    assert (f1 - f0).abs().max() < 1e-4
