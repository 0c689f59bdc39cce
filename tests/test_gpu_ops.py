"""GPU parity of slice extraction, ML contractions and the data-feed / CG / MLP
kernels against the CPU oracle."""

import numpy as np
import pytest
import torch

from tests.conftest import oracle, pkg, rel_l2

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- slice extraction


@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
@pytest.mark.parametrize("L,us", [(8, 2), (9, 2), (7, 1)])
def test_slice_forward_adjoint_vs_oracle(dev, mode, L, us):
    o = oracle()
    npl = pkg("nufft_plan")
    rng = np.random.default_rng(30 + L)
    vL = L * us
    R = np.linalg.qr(rng.standard_normal((5, 3, 3)))[0]
    pts = torch.tensor(o.pts_rot(R, L), device=dev)
    plan = npl.NufftPlanDiscretized((L,) * 3, upsample_factor=us, mode=mode)
    plan.setpts(pts)
    vol = rng.standard_normal((3, vL, vL, vL)) + 1j * rng.standard_normal((3, vL, vL, vL))
    out = plan.execute_forward(torch.tensor(vol, dtype=torch.complex64, device=dev))
    grid = plan._grid_flat.double().cpu().numpy()
    ref = o.slice_forward(vol, grid, mode).reshape(3, 5, L, L)
    assert rel_l2(out.cpu().numpy(), ref) < 5e-6  # FP32 weights and sums vs the FP64 oracle
    sig = rng.standard_normal((5, L, L)) + 1j * rng.standard_normal((5, L, L))
    adj = plan.execute_adjoint(torch.tensor(sig, dtype=torch.complex64, device=dev))
    ref_adj = o.slice_adjoint(sig.reshape(1, -1), grid, (vL,) * 3, mode)
    assert adj.shape == (1, vL, vL, vL)
    assert rel_l2(adj.cpu().numpy(), ref_adj) < 5e-6
    un = plan.execute_adjoint_unaggregated(torch.tensor(sig, dtype=torch.complex64, device=dev))
    assert un.shape == (5, vL, vL, vL)
    assert rel_l2(un.sum(0).cpu().numpy(), ref_adj[0]) < 5e-6


def test_vol_forward_im_backward_adjointness_discretized(dev):
    """Dot-product test of the discretised pair in FP64 (reference test_torch_wraps.py:374-424)."""
    npl = pkg("nufft_plan")
    pf = pkg("projection_funcs")
    o = oracle()
    L = 8
    rng = np.random.default_rng(40)
    pts = torch.tensor(o.pts_rot(np.linalg.qr(rng.standard_normal((4, 3, 3)))[0], L), dtype=torch.float64,
                       device=dev)
    plan = npl.NufftPlanDiscretized((L,) * 3, upsample_factor=2, mode="nearest")
    plan.setpts(pts)
    v = torch.tensor(rng.standard_normal((1, 2 * L, 2 * L, 2 * L)) + 1j * rng.standard_normal((1, 16, 16, 16)),
                     device=dev)
    y = torch.tensor(rng.standard_normal((4, L, L)) + 1j * rng.standard_normal((4, L, L)), device=dev)
    F = torch.tensor(rng.uniform(0.1, 1, (4, L, L)), device=dev)
    for filt in (None, F):
        Av = pf.vol_forward(v, plan, filt, fourier_domain=True)
        Ay = pf.im_backward(y, plan, filt, fourier_domain=True)
        lhs = torch.vdot(Av.reshape(-1), y.reshape(-1))
        rhs = torch.vdot(v.reshape(-1), Ay.reshape(-1))
        assert abs(lhs - rhs) <= 1e-8 * abs(lhs) + 1e-4


# ---------------------------------------------------------------- ML contractions


def _ml_inputs(dev, B=6, r=5, L=16, seed=50):
    g = torch.Generator().manual_seed(seed)
    c = torch.complex(torch.randn((r, B, L * L), generator=g), torch.randn((r, B, L * L), generator=g)) * 3
    yhat = torch.complex(torch.randn((B, L, L), generator=g), torch.randn((B, L, L), generator=g)) * 5
    ctf = torch.rand((B, L, L), generator=g) * 2 - 1
    return c.to(dev), yhat.to(dev), ctf.to(dev)


def _ml_oracle(c, yhat, ctf, L, s2, lam_inv=None):
    o = oracle()
    mask = o.nyquist_mask(L).ravel()
    c = c.cpu().numpy().astype(np.complex128)
    Q = mask * ctf.cpu().numpy().reshape(ctf.shape[0], -1) * c / L**2  # (r, B, d)
    Q = np.transpose(Q, (1, 0, 2))
    yt = yhat.cpu().numpy().reshape(yhat.shape[0], -1) / L
    P = np.concatenate([Q.real, Q.imag], axis=2)
    y = np.concatenate([yt.real, yt.imag], axis=1)
    return P, y


@pytest.mark.parametrize("precision,tol", [(0, 1e-5), (1, 2e-4)])
@pytest.mark.parametrize("r", [1, 5, 20, 31])
def test_ml_forward_vs_oracle(dev, precision, tol, r):
    ml = pkg("ml_ops")
    o = oracle()
    L, s2 = 16, 0.3
    c, yhat, ctf = _ml_inputs(dev, r=r)
    res = ml.ml_forward(c, yhat, L, s2, ctf=ctf, precision=precision, want_m=True, want_minv=True, want_b=True)
    P, y = _ml_oracle(c, yhat, ctf, L, s2)
    cost_i, z, M, b = o.ml_terms(P, y, s2)
    print(f"ml_forward r={r} prec={precision}: M {rel_l2(res['m'].cpu().numpy(), M):.2e}")
    assert rel_l2(res["m"].cpu().numpy(), M) < tol
    assert rel_l2(res["b"].cpu().numpy(), b) < tol
    assert rel_l2(res["z"].cpu().numpy(), z) < 10 * tol
    assert rel_l2(res["cost"].cpu().numpy(), cost_i) < 10 * tol
    assert rel_l2(res["minv"].cpu().numpy(), np.linalg.inv(M)) < 10 * tol


def test_ml_forward_with_prior_and_const(dev):
    ml = pkg("ml_ops")
    o = oracle()
    L, s2, r = 16, 0.5, 4
    c, yhat, ctf = _ml_inputs(dev, r=r, seed=51)
    lam = torch.diag(torch.tensor([4.0, 2.0, 1.0, 0.5]))
    lam_inv = torch.inverse(lam)
    res = ml.ml_forward(c, yhat, L, s2, ctf=ctf, lam_inv=lam_inv.to(dev), mean_const=True, precision=0)
    P, y = _ml_oracle(c, yhat, ctf, L, s2)
    cost_i, z, M, b = o.ml_terms(P, y, s2, lam_inv=lam_inv.double().numpy())
    cost_i = cost_i + 0.5 * (y**2).sum(1) / s2
    assert rel_l2(res["z"].cpu().numpy(), z) < 1e-4
    assert rel_l2(res["cost"].cpu().numpy(), cost_i) < 1e-4


def test_ml_grad_images_vs_closed_form(dev):
    ml = pkg("ml_ops")
    o = oracle()
    L, s2, r, B = 16, 0.4, 6, 5
    c, yhat, ctf = _ml_inputs(dev, B=B, r=r, seed=52)
    res = ml.ml_forward(c, yhat, L, s2, ctf=ctf, precision=0)
    g = ml.ml_grad_images(c, yhat, res["z"], res["minv"], L, s2, 1.0 / B, ctf=ctf)
    P, y = _ml_oracle(c, yhat, ctf, L, s2)
    _, G = o.ml_cost_grad(P, y, s2)  # dF/d[Re Q, Im Q]
    d = L * L
    Gq = G[:, :, :d] + 1j * G[:, :, d:]  # real-pair gradient wrt Q
    mask = o.nyquist_mask(L).ravel()
    D = mask * ctf.cpu().numpy().reshape(B, -1)
    ref = np.transpose(Gq, (1, 0, 2)) * D / L**2  # dQ/dc = D / L^2
    assert rel_l2(g.cpu().numpy(), ref) < 1e-4


@pytest.mark.parametrize("r", [6, 20])
def test_ml_grad_sorted_matches_images_then_adjoint(dev, r):
    """The gradient written straight into a plan's sorted type-1 input, then the presorted adjoint,
    equals ml_grad_images followed by the plan's ordinary adjoint bit for bit (same arithmetic,
    same rows, same spreading order); half-plane pixels, CTF, a full batch and r = 6 / 20."""
    ml = pkg("ml_ops")
    npl = pkg("nufft_plan")
    dsm = pkg("dataset")
    L, s2, B = 16, 0.4, 5
    g = torch.Generator().manual_seed(53 + r)
    pix = ml.half_plane_pixels(L, dev)
    npx = pix.numel()
    c = (torch.complex(torch.randn((r, B, npx), generator=g), torch.randn((r, B, npx), generator=g)) * 3).to(dev)
    yhat = (torch.complex(torch.randn((B, L, L), generator=g), torch.randn((B, L, L), generator=g)) * 5).to(dev)
    ctf = (torch.rand((B, L, L), generator=g) * 2 - 1).to(dev)
    rots = dsm.haar_rotations(B, g).to(dev)
    pts = dsm.pts_rot_from_rotations(rots, L)
    plan = npl.NufftPlan((L,) * 3, batch_size=r, eps=1e-6, device=dev)
    plan.set_points_flat(pts[:, :, pix.long()].transpose(0, 1).reshape(3, -1).contiguous(), B)
    res = ml.ml_forward(c, yhat, L, s2, ctf=ctf, precision=1, pix_map=pix)
    gimg = ml.ml_grad_images(c, yhat, res["z"], res["minv"], L, s2, 1.0 / B, ctf=ctf, pix_map=pix)
    ref = plan.execute_adjoint(gimg.reshape(r, -1), real_output=True)
    vals, ldc, spos = plan.sorted_input(r)
    assert ldc % 2 == 0 and r <= ldc <= 24
    ml.ml_grad_sorted(c, yhat, res["z"], res["minv"], L, s2, 1.0 / B, vals, spos, ldc, ctf=ctf, pix_map=pix)
    out = plan.execute_adjoint_presorted(r, real_output=True)
    assert torch.equal(out, ref)


@pytest.mark.parametrize("L", [16, 15])
def test_half_plane_equals_full_plane_for_hermitian_data(dev, L):
    """With exactly Hermitian projections, the half-plane Gram/gradient equal the full-plane ones."""
    ml = pkg("ml_ops")
    pf = pkg("projection_funcs")
    r, B, s2 = 4, 5, 0.7
    g = torch.Generator().manual_seed(70)
    # c = fft2c of real fields, so c(-k) = conj(c(k)) on the grid (odd L) / off the Nyquist lines (even L)
    c = pf.centered_fft2(torch.randn((r, B, L, L), generator=g, dtype=torch.float64)).to(torch.complex64)
    yhat = pf.centered_fft2(torch.randn((B, L, L), generator=g, dtype=torch.float64)).to(torch.complex64)
    ctf = torch.rand((B, L, L), generator=g)
    ctf = 0.5 * (ctf + torch.flip(ctf, dims=(1, 2)).roll(shifts=(1 - L % 2, 1 - L % 2), dims=(1, 2)))
    c, yhat, ctf = c.to(dev), yhat.to(dev), ctf.to(dev)
    full = ml.ml_forward(c.reshape(r, B, -1), yhat, L, s2, ctf=ctf, precision=0, want_m=True)
    pix = ml.half_plane_pixels(L, dev)
    ch = c.reshape(r, B, -1)[:, :, pix.long()].contiguous()
    half = ml.ml_forward(ch, yhat, L, s2, ctf=ctf, precision=0, want_m=True, pix_map=pix)
    assert rel_l2(half["m"].cpu().numpy(), full["m"].cpu().numpy()) < 1e-5
    assert rel_l2(half["cost"].cpu().numpy(), full["cost"].cpu().numpy()) < 1e-5
    gf = ml.ml_grad_images(c.reshape(r, B, -1), yhat, full["z"], full["minv"], L, s2, 1.0 / B, ctf=ctf)
    gh = ml.ml_grad_images(ch, yhat, half["z"], half["minv"], L, s2, 1.0 / B, ctf=ctf, pix_map=pix)
    mult = torch.where(pix == (L // 2) * L + L // 2, 1.0, 2.0).to(dev)
    assert rel_l2(gh.cpu().numpy(), (gf[:, :, pix.long()] * mult).cpu().numpy()) < 1e-5


# ---------------------------------------------------------------- feed, CG, MLP


def test_pts_rot_and_ctf_kernels(dev):
    o = oracle()
    ds = pkg("dataset")
    rng = np.random.default_rng(60)
    L = 32
    R = np.linalg.qr(rng.standard_normal((7, 3, 3)))[0].astype(np.float32)
    pts = ds.pts_rot_from_rotations(torch.tensor(R, device=dev), L).cpu().numpy()
    ref = o.pts_rot(R, L)
    assert np.abs(pts - ref).max() < 2e-6 * np.pi * 4
    params = np.array([[15000, 14000, 0.3, 300, 2.7, 0.1, 1.5], [25000, 25000, 0.0, 200, 2.0, 0.07, 3.0]],
                      dtype=np.float32)
    h = ds.ctf_from_params(torch.tensor(params, device=dev), L).cpu().numpy()
    assert np.abs(h - o.ctf(params.astype(np.float64), L)).max() < 1e-5


def test_vdot_fp64_and_toeplitz(dev):
    o = oracle()
    mean = pkg("mean")
    rng = np.random.default_rng(61)
    x = (rng.standard_normal(5000) + 1j * rng.standard_normal(5000)).astype(np.complex64)
    y = (rng.standard_normal(5000) + 1j * rng.standard_normal(5000)).astype(np.complex64)
    v = mean.vdot_f64(torch.tensor(x, device=dev), torch.tensor(y, device=dev))
    ref = np.vdot(x.astype(np.complex128), y.astype(np.complex128))
    assert abs(complex(v) - ref) < 1e-12 * abs(ref) + 1e-9
    L, nf = 8, 16
    lhs = (rng.uniform(0.5, 2, (nf,) * 3) + 0j).astype(np.complex64)
    xv = (rng.standard_normal(L**3) + 1j * rng.standard_normal(L**3)).astype(np.complex64)
    out = mean.toeplitz_apply(torch.tensor(xv, device=dev), torch.tensor(lhs, device=dev), L)
    ref = o.toeplitz_apply(xv.astype(np.complex128), lhs.astype(np.complex128), L).ravel()
    assert rel_l2(out.cpu().numpy(), ref) < 1e-5


def test_mlp_forward_backward(dev):
    o = oracle()
    pc = pkg("precond")
    rng = np.random.default_rng(62)
    n_in = 4
    mlp = pc.DiagonalMLP(n_in, device=dev, seed=3)
    x = torch.tensor(rng.standard_normal((300, n_in)), dtype=torch.float32, device=dev)
    y = mlp.forward(x)
    ref = o.mlp_forward(mlp.params.detach().cpu().numpy().astype(np.float64), x.cpu().numpy().astype(np.float64))
    assert rel_l2(y.cpu().numpy(), ref) < 1e-5
    gy = torch.tensor(rng.standard_normal(300), dtype=torch.float32, device=dev)
    grad = mlp.backward(x, gy)
    p64 = mlp.params.detach().double().cpu().requires_grad_(True)
    ref_y = pc.mlp_torch_reference(p64, x.double().cpu())
    (ref_y * gy.double().cpu()).sum().backward()
    assert rel_l2(grad.cpu().numpy(), p64.grad.numpy()) < 1e-4


@pytest.mark.parametrize("L,nf", [(7, 64), (32, 64), (48, 128), (64, 128), (128, 256), (40, 96)])
def test_toeplitz_operator_matches_oracle(dev, L, nf):
    """cov3d_toeplitz_apply against the FP64 oracle: the pruned passes on 64/128/256-cell grids
    (odd and even L, L up to nf/2) and the rocFFT path on other grids (nf = 96)."""
    o = oracle()
    mean = pkg("mean")
    rng = np.random.default_rng(L + nf)
    lhs = (rng.standard_normal((nf,) * 3) + 1j * rng.standard_normal((nf,) * 3)).astype(np.complex64)
    x = (rng.standard_normal(L**3) + 1j * rng.standard_normal(L**3)).astype(np.complex64)
    out = mean.toeplitz_apply(torch.tensor(x, device=dev), torch.tensor(lhs, device=dev), L)
    ref = o.toeplitz_apply(x.astype(np.complex128), lhs.astype(np.complex128), L).ravel()
    err = rel_l2(out.cpu().numpy(), ref)
    print(f"Toeplitz L={L} nf={nf}: {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("L,nf", [(16, 512), (200, 512), (128, 256)])
def test_toeplitz_operator_closed_forms(dev, L, nf):
    """Sizes where the FP64 oracle is too slow (512-cell grids): a constant lhs is a multiple of the
    identity and the centred phase ramp exp(-2 pi i k.s / nf) translates x by s with zero fill."""
    mean = pkg("mean")
    g = torch.Generator(device=dev).manual_seed(L)
    x = torch.complex(torch.randn(L**3, device=dev, generator=g), torch.randn(L**3, device=dev, generator=g))
    y = mean.toeplitz_apply(x, torch.full((nf,) * 3, 1.5 + 0j, dtype=torch.complex64, device=dev), L)
    assert rel_l2(y.cpu().numpy(), 1.5 * x.cpu().numpy()) < 1e-6
    k = torch.arange(nf, device=dev, dtype=torch.float64) - nf // 2
    s = (2, -1, 3)
    ph = -2 * np.pi * (k[:, None, None] * s[0] + k[None, :, None] * s[1] + k[None, None, :] * s[2]) / nf
    ramp = torch.polar(torch.ones_like(ph), ph).to(torch.complex64)
    y = mean.toeplitz_apply(x, ramp, L).reshape(L, L, L).cpu().numpy()
    xs = x.reshape(L, L, L).cpu().numpy()
    ref = np.zeros_like(xs)
    src = tuple(slice(max(0, -d), L - max(0, d)) for d in s)
    dst = tuple(slice(max(0, d), L - max(0, -d)) for d in s)
    ref[dst] = xs[src]
    err = rel_l2(y, ref)
    print(f"Toeplitz shift L={L} nf={nf}: {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("n", [8 * 65536 - 1, 8 * 65536 + 12345, 2**21 + 3])
def test_vdot_fp64_batched_lengths(dev, n):
    """Lengths at which every thread of the 256 x 256 dot grid runs whole batches of 8 elements
    and a ragged tail, checked against numpy's complex128 vdot."""
    mean = pkg("mean")
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    y = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    v = mean.vdot_f64(torch.tensor(x, device=dev), torch.tensor(y, device=dev))
    ref = np.vdot(x.astype(np.complex128), y.astype(np.complex128))
    assert abs(complex(v) - ref) < 1e-12 * np.sqrt(n) * np.abs(x).max() * np.abs(y).max() + 1e-9
