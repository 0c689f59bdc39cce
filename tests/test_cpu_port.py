"""Pins the FP32 CPU port that bench.py times as its CPU baseline (oracle/cpu_nufft.py with the
C++/OpenMP spreading and interpolation of oracle/cpu_nufft.cpp): its type-2 / type-1 NUFFT against
the FP64 direct NUDFT at the FP32 bar (relative L2 <= 1e-5), and its ML cost and gradient step
against the FP64 oracle's objective (NUDFT projections, closed-form gradient, NUDFT
back-projection)."""

import numpy as np
import pytest

from tests.conftest import oracle, rel_l2


def _cpu():
    import importlib

    return importlib.import_module("oracle.cpu_nufft")


@pytest.mark.parametrize("L", [12, 15, 16])
def test_cpu_port_nufft_matches_nudft(L):
    o = oracle()
    cpu = _cpu()
    rng = np.random.default_rng(L)
    r = 3
    R = np.linalg.qr(rng.standard_normal((3, 3, 3)))[0]
    pts = np.concatenate([np.transpose(o.pts_rot(R, L), (1, 0, 2)).reshape(3, -1),
                          rng.uniform(-np.pi, np.pi, (3, 300))], axis=1).astype(np.float32)
    nu = cpu.CpuNufft(L, r, 1e-6, workers=4)
    nu.setpts(pts)
    vols = rng.standard_normal((r, L, L, L)).astype(np.float32)
    P64 = o.wrap_points(pts).astype(np.float64)
    out = nu.type2(vols).T
    ref = o.nudft_type2(vols.astype(np.complex128), P64)
    assert rel_l2(out, ref) < 1e-5
    c = (rng.standard_normal((pts.shape[1], r)) + 1j * rng.standard_normal((pts.shape[1], r))).astype(np.complex64)
    back = nu.type1_real(c)
    ref1 = o.nudft_type1(c.T.astype(np.complex128), P64, (L, L, L)).real
    assert rel_l2(back, ref1) < 1e-5


@pytest.mark.parametrize("L", [12, 13])
def test_cpu_port_ml_step_matches_fp64_oracle(L):
    o = oracle()
    cpu = _cpu()
    rng = np.random.default_rng(100 + L)
    r, B, nv = 3, 6, 0.5
    R = np.linalg.qr(rng.standard_normal((B, 3, 3)))[0]
    pts = o.pts_rot(R, L).astype(np.float32)
    vols = (rng.standard_normal((r, L, L, L)) / L).astype(np.float32)
    images = rng.standard_normal((B, L, L)).astype(np.float32)
    ctf = o.ctf(np.tile(np.array([[12000, 15000, 0.4, 300, 2.7, 0.1, 3.0]]), (B, 1)), L)
    cost, gv = cpu.sgd_step_cpu(vols, images, pts, ctf.astype(np.float32), nv, cpu.CpuNufft(L, r, 1e-6, workers=4))
    p64 = pts.astype(np.float64)
    P = o.vol_forward_nudft(vols.astype(np.float64), p64, ctf).reshape(B, r, -1)
    ref_cost, G = o.ml_cost_grad(P, images.reshape(B, -1).astype(np.float64), nv)
    ref_grad = np.stack([o.im_backward_nudft(G[:, t].reshape(B, L, L), p64, ctf) for t in range(r)])
    assert abs(cost - ref_cost) / abs(ref_cost) < 1e-4
    assert rel_l2(gv, ref_grad) < 1e-4
