"""cryoDRGN poses.pkl / ctf.pkl (formats.py; reference source.py:89-105, :476-542, poses.py:11-41)."""

import os
import pickle

import numpy as np
import pytest

from tests.conftest import pkg


def _rotations(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)


def test_pose_file_round_trip_and_conventions(tmp_path):
    fm = pkg("formats")
    L, n = 64, 17
    rots = _rotations(n, 0)
    offsets = np.random.default_rng(1).uniform(-3, 3, (n, 2))
    path = str(tmp_path / "poses.pkl")
    fm.save_poses(path, rots, offsets, L)
    # the file is cryoDRGN's layout: transposed rotations, translations as box fractions
    with open(path, "rb") as fid:
        raw = fm._ArrayUnpickler(fid).load()
    assert np.array_equal(raw[0], np.transpose(rots, (0, 2, 1)))
    assert np.allclose(raw[1], offsets / L, rtol=0, atol=1e-15)
    r2, o2 = fm.load_poses(path, L)
    assert np.array_equal(r2, rots)
    assert np.allclose(o2, offsets, rtol=0, atol=1e-12)


class _Evil:
    def __reduce__(self):
        return (os.getcwd, ())


def test_loader_refuses_non_array_globals(tmp_path):
    fm = pkg("formats")
    path = str(tmp_path / "evil.pkl")
    with open(path, "wb") as fid:
        pickle.dump((np.zeros((2, 3, 3)), _Evil()), fid)
    with pytest.raises(pickle.UnpicklingError):
        fm.load_poses(path, 8)
    # float32 and float64 arrays, tuples and lists load
    path2 = str(tmp_path / "ok.pkl")
    with open(path2, "wb") as fid:
        pickle.dump([np.arange(6, dtype=np.float32).reshape(2, 3), (1, 2.5, "a")], fid)
    obj = fm.load_array_pickle(path2)
    assert obj[0].dtype == np.float32 and obj[1] == (1, 2.5, "a")


def test_ctf_file_rescaling_and_writer(tmp_path):
    fm = pkg("formats")
    n = 5
    params7 = np.stack([np.full(n, 15000.0), np.full(n, 14000.0), np.linspace(0, 1, n), np.full(n, 300.0),
                        np.full(n, 2.7), np.full(n, 0.1), np.full(n, 1.5)], 1)
    path = str(tmp_path / "ctf.pkl")
    fm.save_ctf(path, params7, 256)
    raw = fm.load_array_pickle(path)
    assert raw.shape == (n, 9) and np.all(raw[:, 0] == 256) and np.all(raw[:, 8] == 0)
    assert np.allclose(raw[:, 4], params7[:, 2] * 180 / np.pi)
    p8 = fm.load_ctf_for_training(128, path)
    assert p8.shape == (n, 8)
    assert np.allclose(p8[:, 0], 1.5 * 256 / 128)  # the physical box is kept
    assert np.allclose(p8[:, 1:7], raw[:, 2:8])
    with pytest.raises(NotImplementedError):
        q = p8.copy()
        q[0, 7] = 10.0
        fm.ctf_params_from_cryodrgn(q)


def _cryodrgn_ctf(L, apix, dfu, dfv, dfang, volt, cs, w):
    """cryoDRGN's published CTF on the frequency lattice the reference builds for it
    (source.py:91-94, :166-169): x on image axis 0, y on axis 1, in 1/A."""
    g = np.ceil(np.arange(-L / 2, L / 2)) / L
    x = g[:, None] / apix * np.ones((1, L))
    y = g[None, :] / apix * np.ones((L, 1))
    volt = volt * 1000
    cs = cs * 1e7
    dfang = dfang * np.pi / 180
    lam = 12.2643247 / np.sqrt(volt + 0.978466e-6 * volt ** 2)
    ang = np.arctan2(y, x)
    s2 = x ** 2 + y ** 2
    df = 0.5 * (dfu + dfv + (dfu - dfv) * np.cos(2 * (ang - dfang)))
    gamma = 2 * np.pi * (-0.5 * df * lam * s2 + 0.25 * cs * lam ** 3 * s2 ** 2)
    return np.sqrt(1 - w ** 2) * np.sin(gamma) - w * np.cos(gamma)


@pytest.mark.parametrize("dfang", [0.0, 17.0, 90.0, 133.0])
def test_ctf_conversion_reproduces_cryodrgn_ctf(dfang):
    """The converted parameters through the kernel's formula (the oracle's FP64 restatement of
    cov3d_ctf_eval) give the CTF the reference computes for cryoDRGN data, astigmatism included."""
    from oracle import oracle as o

    fm = pkg("formats")
    L = 48
    p8 = np.array([[1.3, 21000.0, 16000.0, dfang, 300.0, 2.7, 0.07, 0.0]])
    p7 = fm.ctf_params_from_cryodrgn(p8).astype(np.float64)
    ours = o.ctf(p7, L)[0]
    ref = _cryodrgn_ctf(L, *p8[0, :7])
    # the parameters are float32 (the kernel's input type); the phase reaches ~350 rad at the
    # lattice corners, where a 6e-8 relative rounding of the pixel size moves the CTF by ~4e-5
    assert np.abs(ours - ref).max() < 2e-4
    # a wrong angle convention (dfang used as it is) is far off for astigmatic parameters
    bad = p7.copy()
    bad[0, 2] = dfang * np.pi / 180
    assert np.abs(o.ctf(bad, L)[0] - ref).max() > 0.1
