"""The slice plan (cov3d_slice_plan_*: points sorted once per setpts, sorted-order forward, shared
sort for the gather adjoint) against the stateless kernels (cov3d_slice_forward / _adjoint), bit
for bit, in both modes and precisions, with points outside the volume, and after a second setpts
on the same plan (nufft_plan.py:94-200).  The FP32 trilinear adjoint of a plan whose volume tiles
into 16 x 4 x 4 blocks runs block-wise through LDS and sums in another order: it is checked to
rounding, with dense clusters that need many LDS chunks per block."""

import ctypes

import pytest
import torch

from tests.conftest import pkg

pytestmark = pytest.mark.gpu


def _stateless(nat, vol, grid, mode, sig=None):
    n, D, H, W = vol.shape
    npts = grid.shape[0]
    st = nat.stream_ptr(vol.device)
    code = 1 if vol.dtype == torch.complex128 else 0
    out = torch.empty((n, npts), dtype=vol.dtype, device=vol.device)
    nat.check(nat.load().cov3d_slice_forward(vol.data_ptr(), grid.data_ptr(), out.data_ptr(), n, D, H, W, npts,
                                             mode, code, st))
    back = torch.empty_like(vol)
    nat.check(nat.load().cov3d_slice_adjoint(sig.data_ptr(), grid.data_ptr(), back.data_ptr(), n, D, H, W, npts,
                                             mode, 1, npts, code, st))
    return out, back


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("mode", [0, 1])
def test_slice_plan_matches_stateless_kernels(dev, dtype, mode):
    nat = pkg("_native")
    rdt = torch.float64 if dtype == torch.complex128 else torch.float32
    g = torch.Generator(device=dev).manual_seed(3 + mode)
    D, n = 24, 3
    vol = torch.randn((n, D, D, D), dtype=dtype, device=dev, generator=g)
    lib = nat.load()
    h = ctypes.c_void_p()
    nat.check(lib.cov3d_slice_plan_create(ctypes.byref(h)))
    try:
        for npts in (5000, 777):  # a second setpts on the same plan, fewer points
            grid = ((torch.rand((npts, 3), dtype=rdt, device=dev, generator=g) * 2 - 1) * 1.1).contiguous()
            sig = torch.randn((n, npts), dtype=dtype, device=dev, generator=g)
            ref_f, ref_b = _stateless(nat, vol, grid, mode, sig)
            st = nat.stream_ptr(dev)
            nat.check(lib.cov3d_slice_plan_setpts(h, grid.data_ptr(), npts, D, D, D, mode,
                                                  1 if dtype == torch.complex128 else 0, st))
            out = torch.empty_like(ref_f)
            nat.check(lib.cov3d_slice_plan_forward(h, vol.data_ptr(), out.data_ptr(), n, st))
            back = torch.empty_like(ref_b)
            nat.check(lib.cov3d_slice_plan_adjoint(h, sig.data_ptr(), back.data_ptr(), n, st))
            torch.cuda.synchronize()
            assert torch.equal(out, ref_f)
            assert torch.equal(back, ref_b)
    finally:
        lib.cov3d_slice_plan_destroy(h)


@pytest.mark.parametrize("shape", [(32, 32, 32), (16, 8, 48)])
def test_slice_plan_tiled_adjoint_matches_gather(dev, shape):
    nat = pkg("_native")
    D, H, W = shape
    n = 20
    g = torch.Generator(device=dev).manual_seed(11)
    uniform = (torch.rand((7000, 3), device=dev, generator=g) * 2 - 1) * 1.05
    cluster = torch.randn((40000, 3), device=dev, generator=g) * 0.12  # thousands of points per block
    edge = torch.rand((3000, 3), device=dev, generator=g) * 0.1 + 0.93  # near and past the upper faces
    grid = torch.cat([uniform, cluster, edge]).contiguous()
    npts = grid.shape[0]
    sig = torch.randn((n, npts), dtype=torch.complex64, device=dev, generator=g)
    vol = torch.zeros((n, D, H, W), dtype=torch.complex64, device=dev)
    _, ref = _stateless(nat, vol, grid, 0, sig)
    lib = nat.load()
    h = ctypes.c_void_p()
    nat.check(lib.cov3d_slice_plan_create(ctypes.byref(h)))
    try:
        st = nat.stream_ptr(dev)
        nat.check(lib.cov3d_slice_plan_setpts(h, grid.data_ptr(), npts, D, H, W, 0, 0, st))
        back = torch.empty_like(ref)
        nat.check(lib.cov3d_slice_plan_adjoint(h, sig.data_ptr(), back.data_ptr(), n, st))
        torch.cuda.synchronize()
    finally:
        lib.cov3d_slice_plan_destroy(h)
    # the block around the centre holds far more than one LDS chunk (256 rows)
    cell = torch.floor((grid + 1) * torch.tensor([W - 1, H - 1, D - 1], device=dev) / 2)
    centre = torch.tensor([(W - 1) // 2, (H - 1) // 2, (D - 1) // 2], device=dev)
    assert int(((cell - centre).abs() <= 1).all(dim=1).sum()) > 4 * 256
    scale = float(ref.abs().max())
    torch.testing.assert_close(back, ref, rtol=1e-5, atol=1e-5 * scale)


@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_slice_plan_split_matches_whole(dev, dtype, mode):
    """cov3d_slice_plan_forward_split / _adjoint_split (two point sets set as one) against the plain
    forward / adjoint of the same plan: the two outputs are the whole output cut at n_a, and the
    adjoint of the two signal parts, scaled by a device scalar, is the adjoint of the joined signal
    times that scalar (0.5: exact in binary, so bit for bit)."""
    npl = pkg("nufft_plan")
    rdt = torch.float64 if dtype == torch.complex128 else torch.float32
    g = torch.Generator(device=dev).manual_seed(5)
    L, us, n = 16, 2, 4
    vL = L * us
    npts, n_a = 6000, 4700
    plan = npl.NufftPlanDiscretized((L,) * 3, upsample_factor=us, mode=mode)
    plan.set_grid_flat(((torch.rand((npts, 3), dtype=rdt, device=dev, generator=g) * 2 - 1) * 1.05).contiguous())
    vol = torch.randn((n, vL, vL, vL), dtype=dtype, device=dev, generator=g)
    whole = plan.execute_forward_flat(vol)
    a, b = plan.execute_forward_split(vol, n_a)
    assert torch.equal(a, whole[:, :n_a]) and torch.equal(b, whole[:, n_a:])
    a0, b0 = plan.execute_forward_split(vol, npts)
    assert torch.equal(a0, whole) and b0.shape == (n, 0)
    sig = torch.randn((n, npts), dtype=dtype, device=dev, generator=g)
    ref = plan._adjoint(sig, aggregated=True)
    got = plan.adjoint_split(sig[:, :n_a].contiguous(), sig[:, n_a:].contiguous())
    assert torch.equal(got, ref)
    half = torch.tensor(0.5, device=dev)
    got = plan.adjoint_split(sig[:, :n_a].contiguous(), sig[:, n_a:].contiguous(), scale=half)
    assert torch.equal(got, ref * 0.5)
    with pytest.raises(ValueError):
        plan.adjoint_split(sig[:, :n_a].contiguous(), None)


def test_discretized_plan_resyncs_on_setpts_and_volume_size(dev):
    """NufftPlanDiscretized re-sorts after setpts and for a volume of another size."""
    npl = pkg("nufft_plan")
    ds = pkg("dataset")
    L = 16
    plan = npl.NufftPlanDiscretized((L,) * 3, upsample_factor=2, mode="bilinear")
    g = torch.Generator().manual_seed(0)
    outs = []
    for seed in (1, 2):
        rots = ds.haar_rotations(6, torch.Generator().manual_seed(seed)).to(dev)
        plan.setpts(ds.pts_rot_from_rotations(rots, L))
        for vL in (2 * L, 2 * L + 2):
            vol = torch.randn((2, vL, vL, vL), generator=g, dtype=torch.complex64).to(dev)
            a = plan.execute_forward(vol)
            fresh = npl.NufftPlanDiscretized((L,) * 3, upsample_factor=2, mode="bilinear")
            fresh.setpts(ds.pts_rot_from_rotations(rots, L))
            assert torch.equal(a, fresh.execute_forward(vol))
            outs.append(a)
    assert not torch.equal(outs[0], outs[2])


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_grid_from_pts_matches_tensor_ops(dev, dtype):
    """cov3d_slice_grid_from_pts (the native half of NufftPlanDiscretized.setpts) equals the
    transpose, flip and affine map of nufft_plan.py:94-118 bit for bit."""
    npl = pkg("nufft_plan")
    ds = pkg("dataset")
    L, us = 16, 2
    rots = ds.haar_rotations(7, torch.Generator().manual_seed(4)).to(dev)
    pts = ds.pts_rot_from_rotations(rots, L).to(dtype)
    plan = npl.NufftPlanDiscretized((L,) * 3, upsample_factor=us, mode="bilinear")
    plan.setpts(pts)
    vL = L * us
    ref = pts.transpose(-2, -1).reshape(1, -1, L, L, 3).flip(-1) * ((vL / (vL - 1)) / torch.pi) + 1 / (vL - 1)
    assert torch.equal(plan._grid_flat, ref.reshape(-1, 3))
    assert plan.points.shape == (1, 7, L, L, 3) and plan.batch_points == 7


@pytest.mark.parametrize("L", [16, 15])
def test_boundary_pixels_cover_the_broken_mirror_pairs(dev, L):
    """The pixels cov3d_slice_boundary_pixels lists are the ones whose discretised slice of a
    Hermitian volume (the centred transform of a real volume) may not be the conjugate of the
    -k partner's: every pair that breaks the symmetry is listed, the list is closed under the
    mirror, ascending, and restricted by keep."""
    npl = pkg("nufft_plan")
    ds = pkg("dataset")
    ml = pkg("ml_ops")
    pf = pkg("projection_funcs")
    us, B = 2, 24
    vL = L * us
    g = torch.Generator().manual_seed(9)
    rots = ds.haar_rotations(B, g).to(dev)
    plan = npl.NufftPlanDiscretized((L,) * 3, upsample_factor=us, mode="bilinear")
    plan.setpts(ds.pts_rot_from_rotations(rots, L))
    vol = torch.randn((1, L, L, L), generator=g, dtype=torch.float64).to(dev)
    vhat = pf.centered_fft3(vol, padding_size=(vL,) * 3)
    c = plan.execute_forward(vhat).reshape(B, L * L)
    mirror = ml._mirror_index(L, dev)
    # the float32 grid makes x(-p) = vL - x(p) hold to ~1e-7 relative only; a pair that reaches the
    # Nyquist cell differs by the weight of that cell's value
    broken = (c - c[:, mirror].conj()).abs() > 1e-4 * c.abs().max()
    if L % 2 == 0:
        # the image's own Nyquist row and column pair with themselves, not with -k (the ML zeroes them)
        k = torch.arange(L * L, device=dev)
        broken &= ~((k // L == 0) | (k % L == 0)).unsqueeze(0)
    order, count = ml._boundary_order(plan, B, L, torch.ones(L * L, dtype=torch.uint8, device=dev))
    flags = torch.zeros((B, L * L), dtype=torch.bool, device=dev)
    for b in range(B):
        row = order[b, : int(count[b])].long()
        assert torch.all(row[1:] > row[:-1])  # ascending
        flags[b, row] = True
    assert broken.any(), "the points never reach the Nyquist cells; the test would be vacuous"
    assert not (broken & ~flags).any()
    assert torch.equal(flags, flags[:, mirror])
    # the cell-distance definition, restated on the host
    x = (plan._grid_flat.double().cpu() + 1) * (0.5 * (vL - 1))
    near = (((x > -1) & (x < 1)) | ((x > vL - 1) & (x < vL + 1))).any(dim=1).reshape(B, L * L)
    near = near | near[:, mirror.cpu()]
    assert torch.equal(flags.cpu(), near)
    keep = torch.zeros(L * L, dtype=torch.uint8, device=dev)
    keep[ml.half_plane_pixels(L, dev).long()] = 1
    order_h, count_h = ml._boundary_order(plan, B, L, keep)
    for b in range(B):
        row = order_h[b, : int(count_h[b])].long()
        assert torch.equal(row, torch.nonzero(flags[b] & keep.bool()).reshape(-1))
