"""End-to-end GPU checks: ML objective forms, hand adjoint, SGD trainer,
latent coordinates, mean reconstruction and (P)CG."""

import numpy as np
import pytest
import torch

from tests.conftest import oracle, pkg, rel_l2

pytestmark = pytest.mark.gpu


def _batch(dev, L=12, r=3, B=6, seed=0):
    o = oracle()
    rng = np.random.default_rng(seed)
    R = np.linalg.qr(rng.standard_normal((B, 3, 3)))[0]
    pts = torch.tensor(o.pts_rot(R, L), device=dev)
    vols = rng.standard_normal((r, L, L, L)) / L
    images = rng.standard_normal((B, L, L))
    ctf = o.ctf(np.tile(np.array([[12000, 15000, 0.4, 300, 2.7, 0.1, 3.0]]), (B, 1)), L)
    return pts, vols, images, ctf


def _oracle_cost_grad(plan, vols, images, ctf, noise_var):
    o = oracle()
    B, L = images.shape[0], images.shape[-1]
    r = vols.shape[0]
    pts = plan.points.cpu().numpy().reshape(3, B, -1).transpose(1, 0, 2).astype(np.float64)
    P = o.vol_forward_nudft(vols, pts, ctf).reshape(B, r, -1)
    cost, G = o.ml_cost_grad(P, images.reshape(B, -1), noise_var)
    grad = np.stack([o.im_backward_nudft(G[:, t].reshape(B, L, L), pts, ctf) for t in range(r)])
    return cost, grad


@pytest.mark.parametrize("form,precision,half,tol", [("fused", 0, True, 1e-4), ("fused", 0, False, 1e-4),
                                                     ("fused", 1, True, 1e-3), ("autograd", 1, True, 1e-4)])
@pytest.mark.parametrize("L", [12, 13])
def test_ml_cost_and_gradient_vs_fp64_oracle(dev, form, precision, half, tol, L):
    npl = pkg("nufft_plan")
    sgd = pkg("covar_sgd")
    r, B = 3, 6
    pts, vols64, images64, ctf64 = _batch(dev, L, r, B)
    plan = npl.NufftPlan((L,) * 3, batch_size=r, device=dev)
    plan.setpts(pts)
    vols = torch.tensor(vols64, dtype=torch.float32, device=dev, requires_grad=True)
    cost = sgd.cost_maximum_liklihood(vols, torch.tensor(images64, dtype=torch.float32, device=dev), plan,
                                      torch.tensor(ctf64, dtype=torch.float32, device=dev), 0.5, form=form,
                                      precision=precision, half_plane=half)
    cost.backward()
    ref_cost, ref_grad = _oracle_cost_grad(plan, vols64, images64, ctf64, 0.5)
    print(f"{form}/{precision}/half={half}/L={L}: cost {abs(cost.item() - ref_cost) / abs(ref_cost):.2e} "
          f"grad {rel_l2(vols.grad.cpu().numpy(), ref_grad):.2e}")
    assert abs(cost.item() - ref_cost) / abs(ref_cost) < tol
    assert rel_l2(vols.grad.cpu().numpy(), ref_grad) < tol


def test_hand_adjoint_matches_autograd_at_scale(dev):
    """Chosen form vs the plain FP32 autograd form on an L=64, r=10 batch (the config-2 shape)."""
    npl = pkg("nufft_plan")
    sgd = pkg("covar_sgd")
    L, r, B = 64, 10, 32
    pts, vols64, images64, ctf64 = _batch(dev, L, r, B, seed=1)
    plan = npl.NufftPlan((L,) * 3, batch_size=r, device=dev)
    plan.setpts(pts)
    out = {}
    for form in ("fused", "autograd"):
        v = torch.tensor(vols64, dtype=torch.float32, device=dev, requires_grad=True)
        c = sgd.cost_maximum_liklihood(v, torch.tensor(images64, dtype=torch.float32, device=dev), plan,
                                       torch.tensor(ctf64, dtype=torch.float32, device=dev), 2.0, form=form)
        c.backward()
        out[form] = (c.item(), v.grad.cpu().numpy())
    assert abs(out["fused"][0] - out["autograd"][0]) / abs(out["autograd"][0]) < 1e-3
    assert rel_l2(out["fused"][1], out["autograd"][1]) < 2e-3


def test_finite_difference_of_chosen_form(dev):
    """Directional derivative of the fused cost vs central differences in FP64 (step 1e-3)."""
    npl = pkg("nufft_plan")
    sgd = pkg("covar_sgd")
    L, r, B = 12, 2, 5
    pts, vols64, images64, ctf64 = _batch(dev, L, r, B, seed=2)
    plan = npl.NufftPlan((L,) * 3, batch_size=r, device=dev)
    plan.setpts(pts)
    img = torch.tensor(images64, dtype=torch.float32, device=dev)
    ctf = torch.tensor(ctf64, dtype=torch.float32, device=dev)
    v = torch.tensor(vols64, dtype=torch.float32, device=dev, requires_grad=True)
    sgd.cost_maximum_liklihood(v, img, plan, ctf, 0.5, form="fused", precision=0).backward()
    d = np.random.default_rng(3).standard_normal(vols64.shape)
    d /= np.linalg.norm(d)
    o = oracle()
    h = 1e-3
    cp, _ = _oracle_cost_grad(plan, vols64 + h * d, images64, ctf64, 0.5)
    cm, _ = _oracle_cost_grad(plan, vols64 - h * d, images64, ctf64, 0.5)
    fd = (cp - cm) / (2 * h)
    an = float((v.grad.cpu().numpy() * d).sum())
    assert abs(fd - an) < 1e-3 * abs(fd) + 1e-6


def test_sgd_trainer_reduces_cost_and_finds_subspace(dev):
    ds = pkg("dataset")
    covar_mod = pkg("covar")
    sgd = pkg("covar_sgd")
    cfg = ds.SyntheticConfig(L=16, n_images=4096, rank=2, eigenvalues=[4.0, 1.0], pixel_size=3.0, snr=1.0,
                             defocus_um=(1.0, 2.5), seed=11)
    data = ds.make_synthetic(cfg, dev)
    torch.manual_seed(0)
    signal_var = float(data.dataset.images.var()) - data.dataset.noise_var
    covar = covar_mod.Covar(16, 2, pixel_var_estimate=max(signal_var, 1e-6) / 16).to(dev)
    trainer = sgd.CovarTrainer(covar, ds.create_dataloader(data.dataset, batch_size=512), dev)
    trainer.setup_training(nufft_disc="nufft", objective_func="ml", lr=1e-5)
    trainer.train_epochs(10)
    costs = trainer.training_log["cost_val"]
    assert costs[-1] < costs[0]
    ev, _ = covar.eigenvecs
    U = data.eigenvecs.reshape(2, -1)
    V = ev.reshape(2, -1)
    s = torch.linalg.svdvals(U @ V.T)
    print("subspace cosines", s.tolist(), "costs", costs)
    assert float(s.max()) > 0.8


def test_latent_map_vs_explicit(dev):
    o = oracle()
    ds = pkg("dataset")
    wc = pkg("wiener_coords")
    cfg = ds.SyntheticConfig(L=12, n_images=64, rank=3, eigenvalues=[3.0, 2.0, 1.0], pixel_size=3.0, snr=0.5,
                             defocus_um=(1.0, 2.5), seed=12)
    data = ds.make_synthetic(cfg, dev)
    lam = data.eigenvals
    coords, covar = wc.latentMAP(data.dataset, data.eigenvecs, lam, batch_size=32, return_coords_covar=True)
    images, pts, filt, _ = data.dataset[0:64]
    P = o.vol_forward_nudft(data.eigenvecs.cpu().numpy(), pts.cpu().numpy().astype(np.float64),
                            filt.cpu().numpy().astype(np.float64)).reshape(64, 3, -1)
    lam_inv = np.linalg.inv(np.diag(lam.cpu().numpy()) + 1e-6 * np.eye(3))
    _, z, M, _ = o.ml_terms(P, images.cpu().numpy().reshape(64, -1).astype(np.float64), data.dataset.noise_var,
                            lam_inv=lam_inv)
    assert rel_l2(coords.cpu().numpy(), z) < 1e-3
    assert rel_l2(covar.numpy(), M) < 1e-3


def test_cg_matches_oracle_and_pcg_fallback(dev):
    o = oracle()
    mean = pkg("mean")
    pc = pkg("precond")
    L, nf = 8, 16
    rng = np.random.default_rng(5)
    ax = np.arange(nf) - nf // 2
    r = np.sqrt(ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2)
    lhs = ((1.0 + 100.0 / (1 + r)) * rng.uniform(0.9, 1.1, (nf,) * 3)).astype(np.complex64)
    b = (rng.standard_normal(L**3) + 0j).astype(np.complex64)
    lhs_t = torch.tensor(lhs, device=dev)
    A = lambda v: mean.toeplitz_apply(v, lhs_t, L)
    x, info = mean.complex_conjugate_gradient(A, torch.tensor(b, device=dev), rtol=1e-6, return_info=True)
    xo, it_o, _ = o.conjugate_gradient(lambda v: o.toeplitz_apply(v, lhs.astype(complex), L).ravel(),
                                       b.astype(complex), rtol=1e-6)
    assert info.final_residual < 2e-6
    assert rel_l2(x.cpu().numpy(), xo) < 1e-4
    T = pc.probe_diagonal(lhs_t, L)
    pre = pc.preconditioner_from_log_diag(torch.log(T), L)
    xp, infop = mean.complex_conjugate_gradient(A, torch.tensor(b, device=dev), rtol=1e-6, precond=pre,
                                                return_info=True)
    assert infop.final_residual < 2e-6
    assert infop.iterations < info.iterations
    bad = pc.preconditioner_from_log_diag(torch.full((L**3,), float("nan"), device=dev), L)
    assert bad is None


def test_mean_reconstruction_recovers_volume(dev):
    ds = pkg("dataset")
    mean = pkg("mean")
    cfg = ds.SyntheticConfig(L=32, n_images=8000, rank=0, eigenvalues=[], pixel_size=3.0, snr=1e6,
                             defocus_um=(1.0, 2.5), seed=13)
    data = ds.make_synthetic(cfg, dev, subtract_mean=False)
    vol = mean.reconstruct_mean(data.dataset, batch_size=1024)
    pf = pkg("projection_funcs")
    # the reconstruction keeps frequencies inside the Nyquist sphere only (mean.py:242-245)
    gt = pf.centered_ifft3(pf.centered_fft3(data.mean) * mean._sphere_mask(32, dev)).real
    err = float((vol - gt).norm() / gt.norm())
    print("mean rel err", err)
    # nearest-neighbour back-projection on the 2x grid; the reference accepts 2e-2 at L=128 (its
    # test_reconstruction.py NUFFT case) and the discretisation error shrinks with L (0.10 at L=16)
    assert err < 8e-2


def test_mean_reconstruction_undoes_offsets(dev):
    """Images translated by stored offsets (U[-2, 2] px, the config-2 shifts) give the same mean as
    the untranslated images: the back-projection multiplies each image by the phase ramp of its
    negated offset (reference mean.py:200-204).  Ignoring the offsets gives a visibly wrong mean,
    so the check is sensitive to the step."""
    ds = pkg("dataset")
    mean = pkg("mean")
    base = dict(L=32, n_images=4000, rank=0, eigenvalues=[], pixel_size=3.0, snr=1e6, defocus_um=(1.0, 2.5), seed=23)
    ref = ds.make_synthetic(ds.SyntheticConfig(**base), dev, subtract_mean=False)
    shifted = ds.make_synthetic(ds.SyntheticConfig(**base, shift_px=2.0), dev, subtract_mean=False)
    off = shifted.dataset.offsets
    assert float(off.abs().max()) > 1.5 and float(off.abs().max()) <= 2.0
    assert float((shifted.dataset.images - ref.dataset.images).norm() / ref.dataset.images.norm()) > 0.3
    v_ref = mean.reconstruct_mean(ref.dataset, batch_size=1024)
    v_shift = mean.reconstruct_mean(shifted.dataset, batch_size=1024)
    err = float((v_shift - v_ref).norm() / v_ref.norm())
    shifted.dataset.offsets = torch.zeros_like(off)
    v_blind = mean.reconstruct_mean(shifted.dataset, batch_size=1024)
    blind = float((v_blind - v_ref).norm() / v_ref.norm())
    print(f"shifted vs unshifted mean: {err:.2e}; offsets ignored: {blind:.2e}")
    # both solves stop at the CG's relative residual 1e-6 on a system whose weakly sampled corner
    # frequencies amplify that residual ~100x in the solution; ignoring the offsets costs 20%
    assert err < 1e-3
    assert blind > 1e-1


def test_halfset_mean_on_noiseless_data(dev):
    """reconstruct_mean_from_halfsets (mean.py:262-291, :359-421) on noiseless shifted images: the
    FSC of the two halves is ~1, so the shell regulariser vanishes and the result equals the
    full-set reconstruction and the ground truth to the accuracy of the nearest-neighbour grid."""
    ds = pkg("dataset")
    mean = pkg("mean")
    pf = pkg("projection_funcs")
    cfg = ds.SyntheticConfig(L=32, n_images=8000, rank=0, eigenvalues=[], pixel_size=3.0, snr=1e6,
                             defocus_um=(1.0, 2.5), shift_px=1.0, seed=24)
    data = ds.make_synthetic(cfg, dev, subtract_mean=False)
    full = mean.reconstruct_mean(data.dataset, batch_size=1024)
    halves = mean.reconstruct_mean_from_halfsets(data.dataset, batch_size=1024)
    gt = pf.centered_ifft3(pf.centered_fft3(data.mean) * mean._sphere_mask(32, dev)).real
    d_full = float((halves - full).norm() / full.norm())
    d_gt = float((halves - gt).norm() / gt.norm())
    print(f"half-set mean vs full-set mean {d_full:.2e}, vs ground truth {d_gt:.2e}")
    assert d_full < 1e-2
    assert d_gt < 8e-2


def test_config2_shifted_batch_gradient(dev):
    """Config 2 (L=64, r=10, shifts U[-2, 2] px, SNR 0.05): images with their projected mean and
    translations are preprocessed (phase shifts of the stored offsets, mean subtraction), then the
    chosen form's gradient on a 64-image batch is checked against the FP64 autograd gradient and
    against central finite differences of the FP64 cost (step 1e-3) along random directions."""
    ds = pkg("dataset")
    npl = pkg("nufft_plan")
    sgd = pkg("covar_sgd")
    cfg = ds.config_by_name("L64")
    assert cfg.shift_px == 2.0 and cfg.rank == 10
    B = 64
    data = ds.make_synthetic(cfg, dev, n_images=B, subtract_mean=False, batch=B)
    data.dataset.preprocess(data.mean)
    assert float(data.dataset.offsets.abs().max()) == 0.0
    images, pts, ctf, _ = data.dataset[0:B]
    nv = data.dataset.noise_var
    L, r = cfg.L, cfg.rank
    vols = (data.eigenvecs * data.eigenvals.sqrt().reshape(-1, 1, 1, 1)).float()
    vols = vols + 0.3 * vols.std() * torch.randn(vols.shape, generator=torch.Generator().manual_seed(8)).to(dev)
    plan = npl.NufftPlan((L,) * 3, batch_size=r, device=dev)
    plan.setpts(pts)
    v = vols.clone().requires_grad_(True)
    cost = sgd.cost_maximum_liklihood(v, images, plan, ctf, nv, form="fused")
    cost.backward()
    p64 = npl.NufftPlan((L,) * 3, batch_size=r, eps=1e-12, dtype=torch.float64, device=dev)
    p64.setpts(pts.double())

    def c64(x):
        return sgd.cost_maximum_liklihood(x, images.double(), p64, ctf.double(), nv, form="autograd")

    v64 = vols.double().clone().requires_grad_(True)
    ref = c64(v64)
    ref.backward()
    err_c = abs(cost.item() - ref.item()) / abs(ref.item())
    err_g = rel_l2(v.grad.cpu().numpy(), v64.grad.cpu().numpy())
    print(f"config 2 batch: cost {err_c:.2e}, gradient {err_g:.2e} vs FP64 autograd")
    assert err_c < 1e-3 and err_g < 2e-3
    g = torch.Generator(device=dev).manual_seed(9)
    h = 1e-3
    with torch.no_grad():
        for _ in range(3):
            d = torch.randn(vols.shape, device=dev, dtype=torch.float64, generator=g)
            d = d / d.norm() * float(vols.norm())
            fd = (c64(vols.double() + h * d).item() - c64(vols.double() - h * d).item()) / (2 * h)
            an = float((v.grad.double() * d).sum())
            print(f"directional derivative: analytic {an:.6e}, finite difference {fd:.6e}")
            assert abs(an - fd) < 2e-3 * abs(fd) + 1e-9


def test_cg_vector_kernels_match_torch(dev):
    """cov3d_cg_update / cov3d_cg_direction (the fused CG vector steps) against the reference's
    complex64 torch updates (mean.py:64-80) with the same FP64 scalars."""
    nat = pkg("_native")
    lib = nat.load()
    n = 100003
    g = torch.Generator(device=dev).manual_seed(4)
    cx = lambda: torch.complex(torch.randn(n, device=dev, generator=g), torch.randn(n, device=dev, generator=g))
    x, r, p, ap, z = cx(), cx(), cx(), cx(), cx()
    S = torch.tensor([3.7, 0.0, 2.5, 0.01, 0.0, 0.0, 1.9, 0.0], dtype=torch.float64, device=dev)
    rz, pap, rr, rz_new = (S.data_ptr() + 16 * k for k in range(4))
    alpha = (torch.complex(S[0], S[1]) / torch.complex(S[2], S[3])).to(torch.complex64)
    x_ref = x + alpha * p
    r_ref = r - alpha * ap
    st = nat.stream_ptr(dev)
    nat.check(lib.cov3d_cg_update(x.data_ptr(), r.data_ptr(), p.data_ptr(), ap.data_ptr(), n, rz, pap, rr, st))
    assert (x - x_ref).abs().max() < 1e-5 and (r - r_ref).abs().max() < 1e-5
    rr_ref = (r.to(torch.complex128).abs() ** 2).sum()
    assert abs(float(S[4]) - float(rr_ref)) < 1e-9 * float(rr_ref) and float(S[5]) == 0.0
    beta = (torch.complex(S[6], S[7]) / torch.complex(S[0], S[1])).to(torch.complex64)
    p_ref = p * beta + z
    nat.check(lib.cov3d_cg_direction(p.data_ptr(), z.data_ptr(), n, rz_new, rz, st))
    assert (p - p_ref).abs().max() < 1e-5
