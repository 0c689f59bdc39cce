"""Host-streamed datasets (dataset.StreamingCovarDataset): images in pinned host memory, copied to
the device per batch through two buffers on a copy stream (covar_sgd.py:240-244,
dataset.py:850-895).  The trainer, the latent coordinates and the mean reconstruction must give the
same results as on the device-resident copy of the same data."""

import pytest
import torch

from tests.conftest import pkg, rel_l2

pytestmark = pytest.mark.gpu


def _pair(dev, n=1500, L=16, r=3, shift=0.0, seed=91):
    ds = pkg("dataset")
    cfg = ds.SyntheticConfig(L, n, r, [4.0, 2.0, 1.0][:r], 3.0, 1.0, (0.8, 2.5), shift_px=shift, seed=seed)
    data = ds.make_synthetic(cfg, dev, subtract_mean=False)
    res = data.dataset
    st = ds.StreamingCovarDataset(res.images.cpu(), res.rotations, res.noise_var, ctf_params=res.ctf_params,
                                  offsets=res.offsets.clone())
    assert st.images.is_pinned() and not st.images.is_cuda
    return data, res, st


def test_streamed_batches_equal_resident_batches(dev):
    ds = pkg("dataset")
    data, res, st = _pair(dev)
    it_r = iter(ds.create_dataloader(res, 256))
    n = 0
    for (ir, pr, fr, _), (is_, ps, fs, _) in zip(it_r, ds.create_dataloader(st, 256)):
        assert torch.equal(ir, is_) and torch.equal(pr, ps) and torch.allclose(fr, fs)
        n += 1
    assert n == 6  # 1500 images in batches of 256: five full, one of 220
    st.to_fourier_domain()
    res.to_fourier_domain()
    assert st.noise_var == res.noise_var
    a, b = res[100:300][0], st[100:300][0]
    assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("nufft_disc", ["nufft", "bilinear"])
def test_streamed_training_equals_resident_training(dev, nufft_disc):
    sgd = pkg("covar_sgd")
    cv = pkg("covar")
    ds = pkg("dataset")
    data, res, st = _pair(dev)
    out = {}
    for tag, dset in (("resident", res), ("streamed", st)):
        torch.manual_seed(0)
        covar = cv.Covar(16, 3, vectors=torch.randn((3, 16, 16, 16)) * 0.05).to(dev)
        tr = sgd.CovarTrainer(covar, ds.create_dataloader(dset, batch_size=256), dev)
        tr.setup_training(nufft_disc=nufft_disc, objective_func="ml", lr=1e-5)
        tr.train_epochs(2)
        tr.complete_training()
        out[tag] = (covar.get_vectors_spatial_domain().detach(), tr.training_log["cost_val"])
    print(nufft_disc, out["resident"][1], out["streamed"][1])
    assert rel_l2(out["streamed"][0].cpu().numpy(), out["resident"][0].cpu().numpy()) < 1e-5
    assert dset._in_spatial_domain


def test_streamed_latents_and_mean_equal_resident(dev):
    wc = pkg("wiener_coords")
    mean = pkg("mean")
    data, res, st = _pair(dev, shift=1.5)
    z_r = wc.latentMAP(res, data.eigenvecs, data.eigenvals, batch_size=512)
    z_s = wc.latentMAP(st, data.eigenvecs, data.eigenvals, batch_size=512)
    assert rel_l2(z_s.cpu().numpy(), z_r.cpu().numpy()) < 1e-6
    m_r = mean.reconstruct_mean(res, batch_size=512)
    m_s = mean.reconstruct_mean(st, batch_size=512)
    # the streamed images get their centred FFT per batch rather than all at once (rounding-level
    # differences), which the padded system's weakly sampled corner modes amplify in the CG solve
    assert rel_l2(m_s.cpu().numpy(), m_r.cpu().numpy()) < 5e-4


def test_make_synthetic_host_images(dev):
    ds = pkg("dataset")
    cfg = ds.SyntheticConfig(16, 1000, 2, [4.0, 1.0], 3.0, 0.5, (0.8, 2.5), seed=92)
    a = ds.make_synthetic(cfg, dev)
    b = ds.make_synthetic(cfg, dev, host_images=True)
    assert isinstance(b.dataset, ds.StreamingCovarDataset) and b.dataset.images.is_pinned()
    assert abs(a.dataset.noise_var - b.dataset.noise_var) < 1e-6 * a.dataset.noise_var
    assert rel_l2(b.dataset.images.numpy(), a.dataset.images.cpu().numpy()) < 1e-5
