"""Image preprocessing on the device (projection_funcs.py:199-269, dataset.py:232-281):
translations, the projected mean and the soft projected mask, checked against the FP64 NUDFT
projection of the oracle and against the shift theorem."""

import numpy as np
import pytest
import torch

from tests.conftest import oracle, pkg

pytestmark = pytest.mark.gpu


def _dataset(dev, L, B, seed, with_offsets):
    o = oracle()
    dsm = pkg("dataset")
    rng = np.random.default_rng(seed)
    R = np.linalg.qr(rng.standard_normal((B, 3, 3)))[0]
    mean = rng.standard_normal((1, L, L, L)) / L
    ctf = o.ctf(np.tile(np.array([[12000, 15000, 0.4, 300, 2.7, 0.1, 3.0]]), (B, 1)), L)
    proj = o.vol_forward_nudft(mean, o.pts_rot(R, L), ctf)[:, 0]  # (B, L, L) FP64 oracle
    offsets = rng.integers(-4, 5, size=(B, 2)).astype(np.float64) if with_offsets else np.zeros((B, 2))
    # the particle of image i is translated by -offsets[i]: stored offsets undo it
    images = np.stack([np.roll(proj[i], shift=(-int(offsets[i, 1]), -int(offsets[i, 0])), axis=(0, 1))
                       for i in range(B)])
    ds = dsm.CovarDataset(torch.tensor(images, dtype=torch.float32, device=dev),
                          torch.tensor(R, dtype=torch.float32, device=dev), 1.0,
                          filters=torch.tensor(ctf, dtype=torch.float32, device=dev),
                          offsets=torch.tensor(offsets, dtype=torch.float32, device=dev))
    return ds, torch.tensor(mean, dtype=torch.float32, device=dev), images


@pytest.mark.parametrize("L,fourier", [(16, False), (15, False), (16, True)])
def test_preprocess_removes_translation_and_projected_mean(dev, L, fourier):
    """Images made of the oracle's projection of the mean, translated by integer offsets, are
    zero after preprocessing (the HIP projection and the phase shifts undo them), in either
    image domain, in batches that do not divide the stack; the offsets are zero afterwards."""
    ds, mean, images = _dataset(dev, L, 10, seed=L, with_offsets=True)
    if fourier:
        ds.to_fourier_domain()
    before = torch.linalg.vector_norm(ds.images).item()
    ds.preprocess(mean, batch_size=4)
    assert torch.linalg.vector_norm(ds.images).item() < 1e-5 * before
    assert (ds.offsets == 0).all()


def test_mask_threshold_and_soft_mask(dev):
    """get_mask_threshold: the 10th percentile of the projected-mask values above 10^-1.5.  The
    soft projected mask of a centred ball keeps the image centre (deep inside the projected disc)
    unchanged, zeroes the corners (far outside), and never amplifies."""
    L, B = 32, 8
    pf = pkg("projection_funcs")
    npl = pkg("nufft_plan")
    ds, _, _ = _dataset(dev, L, B, seed=3, with_offsets=False)
    g = torch.arange(L, device=dev, dtype=torch.float32) - L // 2
    rr = torch.sqrt(g[:, None, None] ** 2 + g[None, :, None] ** 2 + g[None, None, :] ** 2)
    mask = (rr < 0.4 * L).float().unsqueeze(0)
    plan = npl.NufftPlan((L,) * 3, batch_size=1, device=dev)
    plan.setpts(ds[slice(0, B)][1])
    thr = pf.get_mask_threshold(mask, plan)
    vals = pf.vol_forward(mask, plan).reshape(-1).double().cpu().numpy()
    kept = vals[vals > 10 ** (-1.5)]
    assert thr > 10 ** (-1.5)
    frac_below = (kept < thr).mean()
    assert abs(frac_below - 0.1) <= 1.0 / len(kept) + 1e-12
    # soft mask on constant images, zero mean
    ds.images = torch.ones((B, L, L), device=dev)
    ds.preprocess(torch.zeros((1, L, L, L), device=dev), mask=mask, batch_size=B)
    out = ds.images
    assert (out <= 1 + 1e-5).all() and (out >= -1e-5).all()
    assert torch.allclose(out[:, L // 2 - 2:L // 2 + 3, L // 2 - 2:L // 2 + 3], torch.ones(1, device=dev), atol=1e-5)
    assert out[:, :3, :3].abs().max() < 1e-5
