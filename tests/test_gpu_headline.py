"""Parity at the sizes the benchmarks run: L=128 (fine grid 256^3), r=20, batches of 1024, and the
L=256 shape (fine grid 512^3).

Three independent references pin the L=128 NUFFT:

* rocFFT: the same plan with the pruned real FFT passes switched off (COV3D_FFT_MODE=0), for all 20
  transforms;
* an exact closed form: for points on a shifted integer lattice p = 2 pi (k + delta) / L both
  NUFFT types reduce to one FFT of a phase-modulated array (``oracle.lattice_type2/1``, pinned
  against the direct sums on the CPU), so every point of a 1M-point set is checked in FP64;
* the FP64 direct NUDFT on the rotated central slices of a 512-image subset of a 1024-image
  batch: type 2 at 2048 sampled pixels, type 1 at 64 sampled output modes.

Bar: relative L2 <= 1e-5 (FP32, eps 1e-6), as for the small sizes.
"""

import warnings

import numpy as np
import pytest
import torch

from tests.conftest import oracle, pkg, rel_l2

pytestmark = pytest.mark.gpu

TOL_F32 = 1e-5
L128, RANK = 128, 20


def _slice_points(dev, L, B, seed, half=False):
    """Wrapped (3, M) points of B Haar-random central slices, image-major (the plans' layout)."""
    ds = pkg("dataset")
    g = torch.Generator().manual_seed(seed)
    pts = ds.pts_rot_from_rotations(ds.haar_rotations(B, g).to(dev), L)  # (B, 3, L*L)
    if half:
        pts = pts[:, :, pkg("ml_ops").half_plane_pixels(L, dev).long()]
    P = pts.transpose(0, 1).reshape(3, -1)
    return (torch.remainder(P + torch.pi, 2 * torch.pi) - torch.pi).contiguous()


def _plan(dev, L, ntrans, pts, n_images):
    npl = pkg("nufft_plan")
    plan = npl.NufftPlan((L,) * 3, batch_size=ntrans, eps=1e-6, device=dev)
    plan.set_points_flat(pts, n_images)
    return plan


@pytest.mark.parametrize("nrun", [RANK, RANK - 1])
def test_nf256_pruned_fft_matches_rocfft(dev, monkeypatch, nrun):
    """The 256^3 instance of the pruned real FFT passes (the L=128 benchmark grid) against rocFFT's
    R2C / C2R path of the same plan, 20 transforms, type 2 of real volumes and the real part of
    type 1; transforms beyond ``nrun`` stay zero."""
    L = L128
    pts = _slice_points(dev, L, 64, 11)
    M = pts.shape[1]
    g = torch.Generator().manual_seed(12 + nrun)
    vol = torch.randn((nrun, L, L, L), generator=g).to(dev)
    c = torch.complex(torch.randn((nrun, M), generator=g), torch.randn((nrun, M), generator=g)).to(dev)
    outs = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("COV3D_FFT_MODE", mode)
        plan = _plan(dev, L, RANK, pts, 64)
        assert plan.info()["nf"] == (256, 256, 256)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            outs[mode] = (plan.execute_forward_flat(vol).cpu().numpy(),
                          plan.execute_adjoint(c, real_output=True).cpu().numpy())
        del plan
        torch.cuda.synchronize()
    for k in (0, 1):
        err = rel_l2(outs["1"][k][:nrun], outs["0"][k][:nrun])
        print(f"nrun={nrun} {'type2' if k == 0 else 'type1'} pruned vs rocFFT: {err:.2e}")
        assert err < 2e-6
        assert np.abs(outs["1"][k][nrun:]).max(initial=0.0) == 0.0


def test_l128_lattice_closed_form(dev):
    """Type 2 and type 1 at L=128, r=20 on 2 x 2^19 points of two shifted integer lattices (random
    fractional offsets, so every stencil position occurs), against the exact FFT closed forms."""
    o = oracle()
    L = L128
    rng = np.random.default_rng(128)
    groups = []
    for _ in range(2):
        k = rng.integers(-(L // 2), L // 2, (3, 1 << 19))
        delta = rng.uniform(0, 1, 3)
        groups.append((k, delta))
    P = np.concatenate([o.lattice_points(k, d, L) for k, d in groups], axis=1)
    pts = torch.tensor(o.wrap_points(P), dtype=torch.float32, device=dev)
    plan = _plan(dev, L, RANK, pts, 1)
    g = torch.Generator().manual_seed(1)
    vol = torch.randn((RANK, L, L, L), generator=g)
    out2 = plan.execute_forward_flat(vol.to(dev)).cpu().numpy()
    M1 = groups[0][0].shape[1]
    vol64 = vol.numpy().astype(np.float64)
    err2 = [rel_l2(out2[:, i * M1:(i + 1) * M1], o.lattice_type2(vol64, k, d)) for i, (k, d) in enumerate(groups)]
    print("type2 lattice rel_l2", err2)
    assert max(err2) < TOL_F32
    c = torch.complex(torch.randn((RANK, 2 * M1), generator=g), torch.randn((RANK, 2 * M1), generator=g))
    out1 = plan.execute_adjoint(c.to(dev)).cpu().numpy()
    c64 = c.numpy().astype(np.complex128)
    ref1 = sum(o.lattice_type1(c64[:, i * M1:(i + 1) * M1], k, d, L) for i, (k, d) in enumerate(groups))
    err1 = rel_l2(out1, ref1)
    print("type1 lattice rel_l2", err1)
    assert err1 < TOL_F32


def test_l128_batch_matches_nudft_on_subset(dev):
    """The benchmark's point set (1024 Haar-random central slices, the Hermitian half-plane pixels
    the chosen ML form transforms: 8.26M points) at L=128, r=20.  Type 2 of 20 smooth-plus-noise
    volumes is checked at 4 random pixels of each of the first 512 images; type 1 of random values
    on the 512-image subset is checked at 64 random output modes."""
    o = oracle()
    L = L128
    B, sub = 1024, 512
    pts = _slice_points(dev, L, B, 21, half=True)
    npix = pts.shape[1] // B
    plan = _plan(dev, L, RANK, pts, B)
    g = torch.Generator().manual_seed(2)
    ax = torch.arange(L) - L // 2
    env = torch.exp(-(ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2) / (2 * 20.0**2))
    vol = torch.randn((RANK, L, L, L), generator=g) * (0.2 + env)
    out2 = plan.execute_forward_flat(vol.to(dev))
    rng = np.random.default_rng(3)
    cols = (np.arange(sub)[:, None] * npix + rng.integers(0, npix, (sub, 4))).ravel()
    Psub = pts[:, torch.tensor(cols, device=dev)].cpu().numpy().astype(np.float64)
    ref2 = o.nudft_type2(vol.numpy().astype(np.complex128), Psub)
    err2 = rel_l2(out2[:, torch.tensor(cols, device=dev)].cpu().numpy(), ref2)
    print(f"type2 vs NUDFT on {cols.size} points x {RANK}: {err2:.2e}")
    assert err2 < TOL_F32
    del plan
    psub = pts[:, : sub * npix].contiguous()
    plan1 = _plan(dev, L, RANK, psub, sub)
    Ms = psub.shape[1]
    c = torch.complex(torch.randn((RANK, Ms), generator=g), torch.randn((RANK, Ms), generator=g))
    out1 = plan1.execute_adjoint(c.to(dev)).reshape(RANK, -1)
    modes = rng.choice(L**3, 64, replace=False)
    ref1 = o.nudft_type1_modes(c.numpy().astype(np.complex128), psub.cpu().numpy().astype(np.float64), (L, L, L),
                               modes)
    err1 = rel_l2(out1[:, torch.tensor(modes, device=dev)].cpu().numpy(), ref1)
    print(f"type1 vs NUDFT at {modes.size} modes x {RANK} over {Ms} points: {err1:.2e}")
    assert err1 < TOL_F32


def _fp64_ml(dev, vols, images, pts, ctf, noise_var):
    """Cost and gradient of the plain autograd ML objective in FP64 (FP64 NUFFT plan at eps 1e-12)."""
    npl = pkg("nufft_plan")
    sgd = pkg("covar_sgd")
    L, r = vols.shape[-1], vols.shape[0]
    plan = npl.NufftPlan((L,) * 3, batch_size=r, eps=1e-12, dtype=torch.float64, device=dev)
    plan.setpts(pts.double())
    v = vols.double().clone().requires_grad_(True)
    cost = sgd.cost_maximum_liklihood(v, images.double(), plan, ctf.double(), noise_var, form="autograd")
    cost.backward()
    return cost.item(), v.grad


def _ml_forms(dev, vols, images, plan, ctf, noise_var, ref_cost, ref_grad):
    """(cost, gradient) relative errors of the three forms of the ML objective against a reference."""
    sgd = pkg("covar_sgd")
    res = {}
    for tag, form, prec in (("chosen", "fused", 1), ("fused_fp32", "fused", 0), ("plain_fp32", "autograd", 0)):
        v = vols.clone().requires_grad_(True)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            cost = sgd.cost_maximum_liklihood(v, images, plan, ctf, noise_var, form=form, precision=prec)
            cost.backward()
        res[tag] = (abs(cost.item() - ref_cost) / abs(ref_cost), rel_l2(v.grad.cpu().numpy(), ref_grad))
    return res


def test_ml_forms_at_l128_vs_fp64(dev):
    """The chosen form (split-BF16 Gram, hand adjoint, half plane), the fused FP32 form and the plain
    FP32 autograd form against the FP64 autograd objective at the SGD benchmark's L=128, r=20, on
    48 synthetic images generated as for the benchmark (blob eigenvolumes, CTF, SNR 0.02)."""
    ds = pkg("dataset")
    npl = pkg("nufft_plan")
    B = 48
    cfg = ds.config_by_name("L128")
    data = ds.make_synthetic(cfg, dev, n_images=B, batch=B)
    images, pts, ctf, _ = data.dataset[0:B]
    noise_var = data.dataset.noise_var
    vols = (data.eigenvecs * data.eigenvals.sqrt().reshape(-1, 1, 1, 1)).float()
    vols = vols + 0.3 * vols.std() * torch.randn(vols.shape, generator=torch.Generator().manual_seed(4)).to(dev)
    ref_cost, ref_grad = _fp64_ml(dev, vols, images, pts, ctf, noise_var)
    plan = npl.NufftPlan((L128,) * 3, batch_size=RANK, device=dev)
    plan.setpts(pts)
    res = _ml_forms(dev, vols, images, plan, ctf, noise_var, ref_cost, ref_grad.cpu().numpy())
    print(f"L=128 B={B}: rel errors (cost, gradient) vs FP64 autograd: {res}")
    assert res["fused_fp32"][0] < 1e-4 and res["fused_fp32"][1] < 1e-3
    assert res["plain_fp32"][0] < 1e-4 and res["plain_fp32"][1] < 1e-3
    assert res["chosen"][0] < 1e-3 and res["chosen"][1] < 2e-3


def test_ml_forms_at_l256(dev):
    """L=256, r=20 (the full-pipeline shape, fine grid 512^3), 4 images: the chosen and the fused FP32
    forms against the plain FP32 autograd form.  Data are made on the device: 20 random band-limited
    volumes, images as CTF-weighted projections of random latent combinations plus noise."""
    npl = pkg("nufft_plan")
    ds = pkg("dataset")
    pf = pkg("projection_funcs")
    L, B = 256, 4
    g = torch.Generator(device=dev).manual_seed(6)
    ax = torch.arange(L, device=dev) - L // 2
    env = torch.exp(-(ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2) / (2 * 40.0**2))
    vols = pf.centered_ifft3(pf.centered_fft3(torch.randn((RANK, L, L, L), device=dev, generator=g)) * env).real
    vols = vols / vols.reshape(RANK, -1).norm(dim=1).reshape(-1, 1, 1, 1) * torch.linspace(3, 0.3, RANK, device=dev).reshape(-1, 1, 1, 1)
    rots = ds.haar_rotations(B, torch.Generator().manual_seed(7)).to(dev)
    pts = ds.pts_rot_from_rotations(rots, L)
    params = torch.tensor([[15000.0, 15000.0, 0.0, 300.0, 2.7, 0.1, 1.0]], device=dev).repeat(B, 1)
    params[:, 0:2] *= torch.linspace(0.5, 3.5, B, device=dev).reshape(-1, 1) / 1.5
    ctf = ds.ctf_from_params(params, L)
    plan = npl.NufftPlan((L,) * 3, batch_size=RANK, device=dev)
    plan.setpts(pts)
    with torch.no_grad():
        proj = pf.vol_forward(vols, plan, ctf)  # (B, r, L, L)
        z = torch.randn((B, RANK), device=dev, generator=g)
        clean = (proj * z[:, :, None, None]).sum(1)
        noise_var = float(clean.var()) / 0.5
        images = clean + noise_var**0.5 * torch.randn((B, L, L), device=dev, generator=g)
    sgd = pkg("covar_sgd")
    v = vols.clone().requires_grad_(True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ref = sgd.cost_maximum_liklihood(v, images, plan, ctf, noise_var, form="autograd", precision=0)
        ref.backward()
    ref_cost, ref_grad = ref.item(), v.grad.cpu().numpy()
    res = _ml_forms(dev, vols, images, plan, ctf, noise_var, ref_cost, ref_grad)
    print(f"L=256 B={B}: rel errors (cost, gradient) vs plain FP32 autograd: {res}")
    assert res["fused_fp32"][0] < 1e-4 and res["fused_fp32"][1] < 1e-3
    assert res["chosen"][0] < 1e-3 and res["chosen"][1] < 2e-3


def test_nf512_pruned_fft_matches_rocfft_r20(dev, monkeypatch):
    """The 512^3 pruned FFT passes (L=256, the full-pipeline grid) against rocFFT for 20 transforms."""
    L = 256
    pts = _slice_points(dev, L, 4, 31)
    M = pts.shape[1]
    g = torch.Generator().manual_seed(13)
    vol = torch.randn((RANK, L, L, L), generator=g).to(dev)
    c = torch.complex(torch.randn((RANK, M), generator=g), torch.randn((RANK, M), generator=g)).to(dev)
    outs = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("COV3D_FFT_MODE", mode)
        plan = _plan(dev, L, RANK, pts, 4)
        assert plan.info()["nf"] == (512, 512, 512)
        outs[mode] = (plan.execute_forward_flat(vol).cpu().numpy(),
                      plan.execute_adjoint(c, real_output=True).cpu().numpy())
        del plan
        torch.cuda.synchronize()
    for k in (0, 1):
        err = rel_l2(outs["1"][k], outs["0"][k])
        print(f"L=256 {'type2' if k == 0 else 'type1'} pruned vs rocFFT: {err:.2e}")
        assert err < 2e-6
