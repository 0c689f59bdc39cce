"""Pose refinement (covar_sgd.CovarPoseTrainer, poses.PoseModule; reference covar_sgd.py:548-1118,
poses.py:127-330, :500-560): the pose module's slice frequencies equal the native ones, the
rotation gradient through preprocessing and the ML objective matches finite differences in FP64,
Newton's method recovers known offsets, and the trainer pulls perturbed rotations back toward the
true ones."""

import math

import numpy as np
import pytest
import torch

from tests.conftest import pkg, rel_l2

pytestmark = pytest.mark.gpu


def _rotvecs(R: torch.Tensor) -> torch.Tensor:
    """Rotation vectors of rotation matrices (axis * angle)."""
    R = R.double()
    ang = torch.arccos(torch.clamp((R.diagonal(dim1=1, dim2=2).sum(1) - 1) / 2, -1, 1))
    w = torch.stack([R[:, 2, 1] - R[:, 1, 2], R[:, 0, 2] - R[:, 2, 0], R[:, 1, 0] - R[:, 0, 1]], dim=1)
    return (w / (2 * torch.sin(ang)).unsqueeze(1) * ang.unsqueeze(1)).float()


def test_pose_module_matches_native_slices(dev):
    ds = pkg("dataset")
    ps = pkg("poses")
    L, n = 24, 50
    R = ds.haar_rotations(n, torch.Generator().manual_seed(1))
    rv = _rotvecs(R)
    pm = ps.PoseModule(rv, torch.zeros((n, 2)), L).to(dev)
    pts, phase = pm(torch.arange(n, device=dev))
    ref = ds.pts_rot_from_rotations(ps.rotvec_to_rotmat(rv.to(dev)), L)
    assert (pts - ref).abs().max() < 1e-5
    assert torch.allclose(phase, torch.ones_like(phase))


def test_rotation_gradient_matches_finite_differences(dev):
    """d cost / d rotvec through preprocessing (phase shift, mean subtraction) and the ML objective,
    FP64 end to end (FP64 NUFFT plan at eps 1e-12)."""
    ds = pkg("dataset")
    ps = pkg("poses")
    npl = pkg("nufft_plan")
    sgd = pkg("covar_sgd")
    pf = pkg("projection_funcs")
    L, r, B = 16, 2, 6
    cfg = ds.SyntheticConfig(L, B, r, [4.0, 1.0], 3.0, 1.0, (0.8, 2.5), shift_px=1.0, seed=5)
    data = ds.make_synthetic(cfg, dev, subtract_mean=False)
    images, _, ctf, _ = data.dataset[0:B]
    rv = _rotvecs(data.dataset.rotations.cpu()).double()
    pm = ps.PoseModule(rv, data.dataset.offsets.cpu().double(), L, dtype=torch.float64).to(dev)
    plan = npl.NufftPlan((L,) * 3, batch_size=r, eps=1e-12, dtype=torch.float64, device=dev)
    mplan = npl.NufftPlan((L,) * 3, batch_size=1, eps=1e-12, dtype=torch.float64, device=dev)
    vols = (data.eigenvecs * data.eigenvals.sqrt().reshape(-1, 1, 1, 1)).double()
    mean = data.mean.double().unsqueeze(0)
    nv = data.dataset.noise_var
    idx = torch.arange(B, device=dev)

    def cost():
        pts, phase = pm(idx)
        pre = pf.preprocess_image_batch(images.double(), mplan, ctf.double(), (pts, phase), mean)
        plan.setpts(pts)
        return sgd.cost_maximum_liklihood(vols, pre, plan, ctf.double(), nv, form="autograd",
                                          apply_mean_const_term=True)

    pm.zero_grad()
    cost().backward()
    g = pm.rotvec.weight.grad.to_dense().clone()
    w0 = pm.rotvec.weight.data.clone()
    h = 1e-6
    gen = torch.Generator(device=dev).manual_seed(7)
    for _ in range(3):
        d = torch.randn(w0.shape, dtype=torch.float64, device=dev, generator=gen)
        with torch.no_grad():
            pm.rotvec.weight.copy_(w0 + h * d)
            cp = float(cost())
            pm.rotvec.weight.copy_(w0 - h * d)
            cm = float(cost())
            pm.rotvec.weight.copy_(w0)
        fd = (cp - cm) / (2 * h)
        an = float((g * d).sum())
        print(f"rotation gradient: analytic {an:.8e} finite difference {fd:.8e}")
        assert abs(an - fd) < 1e-5 * abs(fd) + 1e-8


def test_newton_recovers_offsets(dev):
    """Fourier-domain images that are exact phase shifts of a reference: Newton on the squared
    difference returns the shifts."""
    ps = pkg("poses")
    pf = pkg("projection_funcs")
    L, n = 32, 8
    g = torch.Generator(device=dev).manual_seed(2)
    ax = torch.arange(L, device=dev) - L // 2
    ref = torch.exp(-(ax[:, None] ** 2 + ax[None, :] ** 2) / 30.0) * (1 + 0.3 * torch.randn((n, L, L), device=dev, generator=g))
    true = (torch.rand((n, 2), device=dev, generator=g) - 0.5) * 1.2
    grid = ps.get_phase_shift_grid(L, device=dev)
    ref_f = pf.centered_fft2(ref)
    images_f = ref_f * ps.offset_to_phase_shift(true, phase_shift_grid=grid)
    est = ps.estimate_image_offsets_newton(images_f, ref_f, in_fourier_domain=True)
    print("offset error", (est - true).abs().max().item())
    assert (est - true).abs().max() < 1e-4


def test_slice_point_gradient_matches_grid_sample(dev):
    """cov3d_slice_point_grad against torch's grid_sample grid derivative (the reference's own
    operator, nufft_plan.py:148-150) in FP64, points near and past the volume's edge included."""
    npl = pkg("nufft_plan")
    L, us, n = 12, 2, 5
    vL = L * us
    gen = torch.Generator(device=dev).manual_seed(4)
    vol = torch.randn((3, vL, vL, vL), dtype=torch.complex128, device=dev, generator=gen)
    pts = (torch.rand((n, 3, L * L), dtype=torch.float64, device=dev, generator=gen) * 2 - 1) * math.pi * 1.02
    plan = npl.NufftPlanDiscretized((L,) * 3, upsample_factor=us, mode="bilinear")
    p1 = pts.clone().requires_grad_(True)
    plan.setpts(p1)
    out = npl.nufft_forward(vol, plan)
    w = torch.randn(out.shape, dtype=out.dtype, device=dev, generator=gen)
    (out.conj() * w).real.sum().backward()

    p2 = pts.clone().requires_grad_(True)
    plan.setpts(p2)
    grid = plan.points
    vri = torch.cat((vol.real.unsqueeze(1), vol.imag.unsqueeze(1)), dim=1).reshape(1, -1, vL, vL, vL)
    ref = torch.nn.functional.grid_sample(vri, grid, mode="bilinear", align_corners=True).reshape(3, 2, n, L, L)
    ref = torch.complex(ref[:, 0], ref[:, 1])
    assert (out - ref).abs().max() < 1e-12
    (ref.conj() * w).real.sum().backward()
    err = rel_l2(p1.grad.cpu(), p2.grad.cpu())
    print("point gradient rel err vs grid_sample", err)
    assert err < 1e-12


@pytest.mark.parametrize("disc", ["nufft", "bilinear"])
def test_pose_trainer_refines_rotations(dev, disc):
    """Rotations perturbed by ~3 degrees move back toward the truth over six epochs of
    CovarPoseTrainer (mean and eigenvolumes started at the truth; pose step 7.7e-3 rad, about the
    reference's default 3 * 1e-6 * 1024); the offset re-estimation and the half-set mean update run
    every third epoch."""
    ds = pkg("dataset")
    ps = pkg("poses")
    cv = pkg("covar")
    sgd = pkg("covar_sgd")
    L, r, n = 32, 2, 1024
    cfg = ds.SyntheticConfig(L, n, r, [4.0, 1.0], 3.0, 2.0, (0.8, 2.5), shift_px=0.0, seed=6)
    data = ds.make_synthetic(cfg, dev, subtract_mean=False)
    R_true = data.dataset.rotations.clone()
    rv = _rotvecs(R_true.cpu())
    noise = torch.randn(rv.shape, generator=torch.Generator().manual_seed(3)) * math.radians(3.0) / math.sqrt(3)
    rv_init = rv + noise

    def err(rot):
        rel = rot.transpose(1, 2) @ R_true
        return float(torch.rad2deg(torch.arccos(torch.clamp((rel.diagonal(dim1=1, dim2=2).sum(1) - 1) / 2, -1, 1))).mean())

    pose = ps.PoseModule(rv_init, torch.zeros((n, 2)), L)
    mean = cv.Mean(data.mean, L)
    covar = cv.Covar(L, r, vectors=(data.eigenvecs * data.eigenvals.sqrt().reshape(-1, 1, 1, 1)).cpu())
    tr = sgd.CovarPoseTrainer(covar, ds.create_dataloader(data.dataset, batch_size=128), dev, mean, pose)
    tr.mean_update_frequency = 3
    tr.setup_training(nufft_disc=disc, objective_func="ml", lr=2e-5)
    before = err(ps.rotvec_to_rotmat(pose.get_rotvecs()))
    tr.train_epochs(6)
    after = err(ps.rotvec_to_rotmat(pose.get_rotvecs()))
    print(f"mean rotation error {before:.3f} -> {after:.3f} degrees; costs {tr.training_log['cost_val']}")
    assert after < 0.7 * before
    assert torch.isfinite(pose.get_offsets()).all() and pose.get_offsets().abs().max() < 2.0


def _perturbed_problem(dev, n=1024, L=32, seed=6, deg=3.0):
    ds = pkg("dataset")
    ps = pkg("poses")
    cv = pkg("covar")
    cfg = ds.SyntheticConfig(L, n, 2, [4.0, 1.0], 3.0, 2.0, (0.8, 2.5), shift_px=0.0, seed=seed)
    data = ds.make_synthetic(cfg, dev, subtract_mean=False)
    R_true = data.dataset.rotations.clone()
    rv = _rotvecs(R_true.cpu())
    noise = torch.randn(rv.shape, generator=torch.Generator().manual_seed(3)) * math.radians(deg) / math.sqrt(3)
    pose = ps.PoseModule(rv + noise, torch.zeros((n, 2)), L)
    mean = cv.Mean(data.mean, L)
    covar = cv.Covar(L, 2, vectors=(data.eigenvecs * data.eigenvals.sqrt().reshape(-1, 1, 1, 1)).cpu())

    def err(rot):
        rel = rot.transpose(1, 2) @ R_true
        return float(torch.rad2deg(torch.arccos(torch.clamp((rel.diagonal(dim1=1, dim2=2).sum(1) - 1) / 2, -1, 1))).mean())

    return data, pose, mean, covar, err


def test_pose_trainer_lbfgs_rotations(dev):
    """rotation_est_method = "LBFGS" (covar_sgd.py:980-981): the row-wise L-BFGS on the rotation
    vectors (steps limited to 0.1 rad) moves 3-degree perturbed rotations toward the truth."""
    ds = pkg("dataset")
    ps = pkg("poses")
    sgd = pkg("covar_sgd")
    data, pose, mean, covar, err = _perturbed_problem(dev)
    tr = sgd.CovarPoseTrainer(covar, ds.create_dataloader(data.dataset, batch_size=128), dev, mean, pose)
    tr.rotation_est_method = "LBFGS"
    tr.mean_update_frequency = 100
    tr.setup_training(nufft_disc="nufft", objective_func="ml", lr=2e-5)
    assert type(tr.pose_optimizer).__name__ == "BlockwiseLBFGS"
    before = err(ps.rotvec_to_rotmat(pose.get_rotvecs()))
    tr.train_epochs(4)
    after = err(ps.rotvec_to_rotmat(pose.get_rotvecs()))
    print(f"LBFGS mean rotation error {before:.3f} -> {after:.3f} degrees; costs {tr.training_log['cost_val']}")
    assert after < 0.8 * before


def test_train_covar_half_sets_with_pose(dev):
    """trainCovar(optimize_pose=True, use_halfsets=True) (covar_sgd.py:1541-1607): each half-set
    refines its own images' poses (PoseModule.split_module of the dataset's permutation), and the
    full-set pass continues from both halves' poses put back in place (update_from_modules)."""
    ps = pkg("poses")
    sgd = pkg("covar_sgd")
    data, pose, mean, covar, err = _perturbed_problem(dev, n=512)
    covar = covar.to(dev)  # trainCovar runs on the model's device (covar_sgd.py:1494)
    rv0 = pose.get_rotvecs().clone().to(dev)
    before = err(ps.rotvec_to_rotmat(rv0))
    seen = {}
    orig = ps.PoseModule.update_from_modules

    def spy(self, m1, m2, permutation):
        seen["halves"] = (m1.get_rotvecs().clone(), m2.get_rotvecs().clone(), permutation.clone())
        return orig(self, m1, m2, permutation)

    ps.PoseModule.update_from_modules = spy
    try:
        trainer = sgd.trainCovar(covar, data.dataset, 128, optimize_pose=True, mean_model=mean, pose=pose,
                                 use_halfsets=True, num_reg_update_iters=1, max_epochs=2, nufft_disc="nufft",
                                 objective_func="ml", lr=2e-5)
    finally:
        ps.PoseModule.update_from_modules = orig
    m1, m2, perm = seen["halves"]
    n = rv0.shape[0]
    assert m1.shape[0] == n // 2 and m2.shape[0] == n - n // 2
    # both halves moved their own rows, and the full-set trainer started from them
    assert not torch.equal(m1, rv0[perm[: n // 2].to(rv0.device)])
    assert not torch.equal(m2, rv0[perm[n // 2:].to(rv0.device)])
    assert trainer.pose is pose
    assert "covariance_fsc_halfset" in trainer.training_log
    after = err(ps.rotvec_to_rotmat(pose.get_rotvecs()))
    print(f"half-set pose rounds: mean rotation error {before:.3f} -> {after:.3f} degrees")
    assert after < before
