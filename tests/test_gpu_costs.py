"""The least-squares and maximum-likelihood objectives, their Fourier-domain forms on the
discretised (bilinear) projection, the trainer's bilinear branch and the NUFFT point gradient,
checked with the invariants the reference's own tests use (tests/test_torch_impl.py:53-89 scaling
laws, :166-195 regulariser identity, :197-244 spatial vs Fourier LS, :246-296 Woodbury ML vs the
dense likelihood) plus a finite-difference check of the point gradient (nufft_plan.py:452-462)."""

import warnings

import numpy as np
import pytest
import torch

from tests.conftest import pkg, rel_l2

pytestmark = pytest.mark.gpu

L16, R4 = 16, 4


def _data(dev, n=64, L=L16, r=R4, seed=700, snr=0.5):
    ds = pkg("dataset")
    cfg = ds.SyntheticConfig(L, n, r, [4.0, 2.0, 1.0, 0.5][:r], 3.0, snr, (0.8, 2.5), seed=seed)
    return ds.make_synthetic(cfg, dev)


def _plan(dev, pts, L=L16, r=R4, dtype=torch.float32, eps=1e-6):
    npl = pkg("nufft_plan")
    plan = npl.NufftPlan((L,) * 3, batch_size=r, eps=eps, dtype=dtype, device=dev)
    plan.setpts(pts)
    return plan


def test_ls_cost_gradient_scaling_laws(dev):
    """Volumes and images scaled by a with the noise variance by a^2: the LS gradient scales by a^3.
    Filters and images scaled by a, noise variance by a^2 and the regulariser weight by a^4: the
    gradient scales by a^4 (reference test_cost_gradient_scaling)."""
    sgd = pkg("covar_sgd")
    fs = pkg("fsc_utils")
    data = _data(dev)
    images, pts, filt, _ = data.dataset[0:50]
    nv = data.dataset.noise_var
    plan = _plan(dev, pts)
    a = 10.0
    g = torch.Generator(device=dev).manual_seed(1)
    v0 = torch.randn((R4, L16, L16, L16), device=dev, generator=g)
    reg = nv / fs.expand_fourier_shell(fs.rpsd(*data.eigenvecs).mean(0), L16, 3)

    def grad(v, im, f, noise, reg_scale):
        v = v.clone().requires_grad_(True)
        sgd.cost(v, im, plan, f, noise, reg_scale=reg_scale, fourier_reg=reg).backward()
        return v.grad

    base = grad(v0, images, filt, nv, 0.0)
    assert rel_l2(grad(a * v0, a * images, filt, a**2 * nv, 0.0).cpu().numpy(), a**3 * base.cpu().numpy()) < 1e-4
    base_r = grad(v0, images, filt, nv, 1e-3)
    scaled = grad(v0, a * images, a * filt, a**2 * nv, 1e-3 * a**4)
    assert rel_l2(scaled.cpu().numpy(), a**4 * base_r.cpu().numpy()) < 1e-4


def test_ls_regulariser_identity(dev):
    """The regulariser reg_scale * ||(V^ sqrt(w)) (V^ sqrt(w))^H||_F^2 with w = noise_var / psd equals
    the Frobenius norm of the covariance in Fourier space weighted entrywise by
    noise_var / sqrt(psd psd^T) (reference test_fourier_reg), read off cost() as the difference
    of the objective with and without it."""
    sgd = pkg("covar_sgd")
    fs = pkg("fsc_utils")
    pf = pkg("projection_funcs")
    data = _data(dev)
    images, pts, filt, _ = data.dataset[0:16]
    plan = _plan(dev, pts)
    vols = data.eigenvecs * data.eigenvals.sqrt().reshape(-1, 1, 1, 1)
    psd = fs.expand_fourier_shell(fs.rpsd(*data.eigenvecs).sum(0), L16, 3)
    nv = 10.0
    w = nv / psd
    with torch.no_grad():
        diff = (sgd.cost(vols, images, plan, filt, 1.0, reg_scale=1.0, fourier_reg=w)
                - sgd.cost(vols, images, plan, filt, 1.0, reg_scale=0.0)).double()
        vf = pf.centered_fft3(vols.double()).reshape(R4, -1)
        coeff = nv / torch.sqrt(psd.double().reshape(-1, 1) * psd.double().reshape(1, -1))
        direct = torch.linalg.norm(coeff * (vf.T @ vf.conj())) ** 2
    assert abs(float(diff) - float(direct)) < 1e-3 * float(direct)


def test_spatial_and_fourier_ls_costs_agree(dev):
    """The LS objective on the NUFFT in the spatial domain and on the bilinear discretised projection
    in the Fourier domain (divided by L^4) agree to within the discretisation, for plain, scaled
    volumes and scaled noise (reference test_fourier_domain_cost, bar 5e-2)."""
    sgd = pkg("covar_sgd")
    cv = pkg("covar")
    npl = pkg("nufft_plan")
    data = _data(dev, n=32)
    torch.manual_seed(3)
    covar = cv.Covar(L16, R4, upsampling_factor=2).to(dev)
    ims, pts, filt, _ = data.dataset[0:16]
    nv = data.dataset.noise_var
    plan = _plan(dev, pts)
    with torch.no_grad():
        v = covar.get_vectors_spatial_domain()
        spatial = [sgd.cost(v, ims, plan, filt, nv), sgd.cost(v * 1e3, ims, plan, filt, nv),
                   sgd.cost(v, ims, plan, filt, nv * 1e6)]
        data.dataset.to_fourier_domain()
        fims, fpts, ffilt, _ = data.dataset[0:16]
        fnv = data.dataset.noise_var
        covar.init_grid_correction("bilinear")
        dplan = npl.NufftPlanDiscretized((L16,) * 3, upsample_factor=2, mode="bilinear")
        dplan.setpts(fpts)
        vf = covar.get_vectors_fourier_domain()
        fourier = [sgd.cost_fourier_domain(vf, fims, dplan, ffilt, fnv),
                   sgd.cost_fourier_domain(vf * 1e3, fims, dplan, ffilt, fnv),
                   sgd.cost_fourier_domain(vf, fims, dplan, ffilt, fnv * 1e6)]
    ratios = [float(f / s) for f, s in zip(fourier, spatial)]
    print("Fourier / spatial LS cost ratios", ratios)
    assert all(abs(q - 1) < 5e-2 for q in ratios)


def _dense_ml(vols, images, plan, filters, noise_var, fourier):
    """0.5 mean_i [y^H C^-1 y - |y|^2 / s2 + logdet C - d log s2], C = P^T conj(P) + s2 I (dense d x d),
    in FP64 (reference test_ml_cost's naive form)."""
    pf = pkg("projection_funcs")
    B, L = images.shape[0], images.shape[-1]
    d = L * L
    P = pf.vol_forward(vols, plan, filters, fourier_domain=fourier).reshape(B, vols.shape[0], d).to(
        torch.complex128 if fourier else torch.float64)
    y = images.reshape(B, d, 1).to(P.dtype)
    C = P.transpose(1, 2) @ P.conj() + noise_var * torch.eye(d, dtype=P.dtype, device=P.device)
    exp = (y.transpose(1, 2).conj() @ torch.linalg.solve(C, y)).reshape(B) - (y.abs() ** 2).sum((1, 2)) / noise_var
    logdet = torch.linalg.slogdet(C)[1] - d * np.log(noise_var)
    return 0.5 * (exp + logdet).mean().real


@pytest.mark.parametrize("form", ["fused", "autograd"])
def test_ml_woodbury_equals_dense_likelihood(dev, form):
    """cost_maximum_liklihood (both forms, NUFFT path) against the dense d x d Gaussian likelihood,
    cost and gradient (reference test_ml_cost: bar 5e-3; here 1e-3)."""
    sgd = pkg("covar_sgd")
    data = _data(dev, n=13)
    ims, pts, filt, _ = data.dataset[0:13]
    nv = data.dataset.noise_var * 100
    plan = _plan(dev, pts)
    g = torch.Generator(device=dev).manual_seed(4)
    v0 = torch.randn((R4, L16, L16, L16), device=dev, generator=g) * 0.05
    v = v0.clone().requires_grad_(True)
    c = sgd.cost_maximum_liklihood(v, ims, plan, filt, nv, form=form, precision=0)
    c.backward()
    p64 = _plan(dev, pts.double(), dtype=torch.float64, eps=1e-12)
    w = v0.double().clone().requires_grad_(True)
    ref = _dense_ml(w, ims.double(), p64, filt.double(), nv, fourier=False)
    ref.backward()
    print(form, float(c), float(ref))
    assert abs(float(c) - float(ref)) < 1e-3 * abs(float(ref))
    assert rel_l2(v.grad.cpu().numpy(), w.grad.cpu().numpy()) < 1e-3


def test_fourier_domain_ml_equals_dense_likelihood(dev):
    """cost_maximum_liklihood_fourier_domain on the bilinear discretised projection (upsampling 1)
    against the dense complex likelihood (reference test_ml_cost, second half), cost and gradient."""
    sgd = pkg("covar_sgd")
    cv = pkg("covar")
    npl = pkg("nufft_plan")
    data = _data(dev, n=13)
    data.dataset.to_fourier_domain()
    ims, pts, filt, _ = data.dataset[0:13]
    nv = data.dataset.noise_var
    torch.manual_seed(5)
    covar = cv.Covar(L16, R4, upsampling_factor=1).to(dev)
    covar.init_grid_correction("bilinear")
    plan = npl.NufftPlanDiscretized((L16,) * 3, upsample_factor=1, mode="bilinear")
    plan.setpts(pts)
    vf = covar.get_vectors_fourier_domain().detach()
    # full plane: the gradient with respect to the complex volume itself
    a = vf.clone().requires_grad_(True)
    c = sgd.cost_maximum_liklihood_fourier_domain(a, ims, plan, filt, nv, half_plane=False)
    c.backward()
    b = vf.to(torch.complex128).clone().requires_grad_(True)
    ref = _dense_ml(b, ims.to(torch.complex128), plan, filt.double(), nv, fourier=True)
    ref.backward()
    print(float(c), float(ref))
    assert abs(float(c) - float(ref)) < 1e-3 * abs(float(ref))
    assert rel_l2(a.grad.cpu().numpy(), b.grad.cpu().numpy()) < 1e-3
    # both planes: the gradient with respect to the real eigenvolumes behind the volume (the
    # half plane is exact there; its complex-volume gradient differs by an anti-Hermitian part)
    pf = pkg("projection_funcs")
    V = covar.get_vectors_spatial_domain().detach()
    gc = covar.grid_correction

    def fourier(v):
        return pf.centered_fft3(v / gc.to(v.dtype), padding_size=(L16,) * 3)

    Vb = V.double().clone().requires_grad_(True)
    _dense_ml(fourier(Vb), ims.to(torch.complex128), plan, filt.double(), nv, fourier=True).backward()
    for half in (False, True):
        Va = V.clone().requires_grad_(True)
        sgd.cost_maximum_liklihood_fourier_domain(fourier(Va), ims, plan, filt, nv, half_plane=half).backward()
        err = rel_l2(Va.grad.cpu().numpy(), Vb.grad.cpu().numpy())
        print("half" if half else "full", "plane: real-eigenvolume gradient rel err", err)
        assert err < 1e-3


@pytest.mark.parametrize("objective", ["ml", "ls"])
def test_bilinear_trainer_branch(dev, objective):
    """setup_training(nufft_disc="bilinear") (the reference default, covar_sgd.py:302-355): the
    dataset goes to the Fourier domain, the discretised plan and the Fourier-domain objective are
    used, the cost falls over the epochs, and complete_training returns the images to space."""
    sgd = pkg("covar_sgd")
    cv = pkg("covar")
    ds = pkg("dataset")
    data = _data(dev, n=2048, snr=1.0)
    torch.manual_seed(6)
    signal_var = max(float(data.dataset.images.var()) - data.dataset.noise_var, 1e-6)
    covar = cv.Covar(L16, R4, pixel_var_estimate=signal_var / L16).to(dev)
    trainer = sgd.CovarTrainer(covar, ds.create_dataloader(data.dataset, batch_size=256), dev)
    trainer.setup_training(nufft_disc="bilinear", objective_func=objective, lr=1e-5 if objective == "ml" else 1e-7)
    assert not data.dataset._in_spatial_domain and not covar._in_spatial_domain
    assert trainer.cost_func is (sgd.cost_maximum_liklihood_fourier_domain if objective == "ml"
                                 else sgd.cost_fourier_domain)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        trainer.train_epochs(4)
    trainer.complete_training()
    costs = trainer.training_log["cost_val"]
    print(objective, costs)
    assert all(np.isfinite(costs)) and costs[-1] < costs[0]
    assert data.dataset._in_spatial_domain and not data.dataset.images.is_complex()


@pytest.mark.parametrize("L", [6, 7])
def test_nufft_point_gradient_finite_differences(dev, L):
    """TorchNufftForward's gradient with respect to the points (nufft_plan.py:452-462, here with the
    plan's own mode grid; the reference hard-codes L=15) against central differences of an FP64 plan:
    the directional derivative of Re <w, type2(V; p)> along random point perturbations."""
    npl = pkg("nufft_plan")
    rng = np.random.default_rng(L)
    N = 2
    R = np.linalg.qr(rng.standard_normal((N, 3, 3)))[0]
    o = pkg("dataset")
    pts0 = o.pts_rot_from_rotations(torch.tensor(R, dtype=torch.float32, device=dev), L).double() * 0.8
    vol = torch.randn((1, L, L, L), dtype=torch.float64, device=dev, generator=torch.Generator(device=dev).manual_seed(2))
    wgt = torch.randn((1, N, L, L), dtype=torch.complex128, device=dev,
                      generator=torch.Generator(device=dev).manual_seed(3))
    plan = npl.NufftPlan((L,) * 3, batch_size=1, eps=1e-12, dtype=torch.float64, device=dev)

    def f(p):
        return (npl.TorchNufftForward.apply(vol, p, plan, True).conj() * wgt).real.sum()

    p = pts0.clone().requires_grad_(True)
    f(p).backward()
    grad = p.grad.clone()
    h = 1e-6
    for k in range(3):
        d = torch.randn(pts0.shape, dtype=torch.float64, device=dev, generator=torch.Generator(device=dev).manual_seed(10 + k))
        with torch.no_grad():
            fd = (f(pts0 + h * d) - f(pts0 - h * d)) / (2 * h)
        an = float((grad * d).sum())
        print(f"L={L}: analytic {an:.10e} finite difference {float(fd):.10e}")
        assert abs(an - float(fd)) < 1e-6 * abs(float(fd)) + 1e-9


@pytest.mark.parametrize("r", [1, 5, 20, 31])
def test_herm_solve_matches_torch(dev, r):
    """cov3d_herm_solve_c128 (one FP64 Cholesky per matrix) against torch's complex128 inverse,
    solve and log-determinant on Hermitian positive definite matrices of the ML's form."""
    ml = pkg("ml_ops")
    g = torch.Generator(device=dev).manual_seed(r)
    n, d = 37, 3 * r + 2
    P = torch.randn((n, r, d), dtype=torch.complex128, device=dev, generator=g)
    M = torch.eye(r, dtype=torch.complex128, device=dev) + P.conj() @ P.transpose(1, 2) / 0.7
    b = torch.randn((n, r), dtype=torch.complex128, device=dev, generator=g)
    minv, z, logdet = ml.herm_solve(M, b)
    ref_inv = torch.linalg.inv(M)
    torch.testing.assert_close(minv, ref_inv, rtol=1e-11, atol=1e-13)
    torch.testing.assert_close(z, torch.linalg.solve(M, b.unsqueeze(-1)).squeeze(-1), rtol=1e-11, atol=1e-12)
    torch.testing.assert_close(logdet, torch.linalg.slogdet(M)[1], rtol=1e-12, atol=1e-12)
