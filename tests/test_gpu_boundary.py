"""The reference-side binding (tests/finufft_stub.py, the cufinufft.Plan stand-in a maintainer adds
to cov3d) driven with the exact arguments and call sequence of cov3d/nufft_plan.py: two plans
built as at :299-316, ``setpts(*points)`` on the wrapped (3, M) rows (:325-336), ``execute`` of
the complex-cast signal (:356-359) and of the (batch_size, M) adjoint input (:386-387), checked
against the FP64 NUDFT."""

import importlib

import numpy as np
import pytest
import torch

from tests.conftest import oracle, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("np_complex,eps,tol", [(np.complex64, 1e-6, 1e-5), (np.complex128, 1e-12, 1e-9)])
def test_reference_call_sequence(dev, np_complex, eps, tol):
    stub = importlib.import_module("tests.finufft_stub")
    o = oracle()
    L, batch_size, N = 12, 3, 4
    sz = (L, L, L)
    real = torch.float32 if np_complex is np.complex64 else torch.float64
    cplx = torch.complex64 if np_complex is np.complex64 else torch.complex128
    kw = {"gpu_method": 1, "gpu_sort": 0}
    forward_plan = stub.Plan(nufft_type=2, n_modes=sz, n_trans=batch_size, eps=eps, dtype=np_complex,
                             gpu_device_id=dev.index or 0, **kw)
    adjoint_plan = stub.Plan(nufft_type=1, n_modes=sz, n_trans=batch_size, eps=eps, dtype=np_complex,
                             gpu_device_id=dev.index or 0, **kw)
    rng = np.random.default_rng(17)
    R = np.linalg.qr(rng.standard_normal((N, 3, 3)))[0]
    pts = torch.tensor(o.pts_rot(R, L), dtype=real, device=dev)  # (N, 3, L^2)
    points = pts.transpose(0, 1).reshape((3, -1))
    points = (torch.remainder(points + torch.pi, 2 * torch.pi) - torch.pi).contiguous()
    forward_plan.setpts(*points)
    adjoint_plan.setpts(*points)
    vol = torch.randn((batch_size, *sz), dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    signal = vol.to(dev).type(cplx).contiguous()
    fwd = forward_plan.execute(signal).reshape((batch_size, -1))
    P = points.cpu().numpy().astype(np.float64)
    ref2 = o.nudft_type2(signal.cpu().numpy().astype(np.complex128), P)
    err2 = rel_l2(fwd.cpu().numpy(), ref2)
    M = P.shape[1]
    c = torch.complex(torch.randn((batch_size, M), dtype=torch.float64), torch.randn((batch_size, M), dtype=torch.float64))
    c = c.to(dev).type(cplx).contiguous()
    adj = adjoint_plan.execute(c.reshape((batch_size, -1))).reshape((batch_size, *sz))
    ref1 = o.nudft_type1(c.cpu().numpy().astype(np.complex128), P, sz)
    err1 = rel_l2(adj.cpu().numpy(), ref1)
    print(f"{np_complex.__name__}: type2 {err2:.2e}, type1 {err1:.2e}")
    assert fwd.dtype == cplx and adj.dtype == cplx
    assert err2 < tol and err1 < tol
    with pytest.raises(ValueError, match="points must be"):
        forward_plan.setpts(*points.to(torch.float64 if real == torch.float32 else torch.float32))
