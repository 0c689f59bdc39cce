"""CPU checks of the C-ABI library and the host-side logic (no GPU compute calls)."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
import torch

from tests.conftest import ROOT, pkg


def _declared_symbols():
    text = (ROOT / "include" / "cov3d_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cov3d_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    native = pkg("_native")
    lib = native.load()
    declared = _declared_symbols()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in include/cov3d_hip.h but not exported"
    assert set(declared) == set(native.EXPORTED), "ctypes signature table out of sync with the header"


def test_error_path_returns_status_and_message():
    native = pkg("_native")
    lib = native.load()
    h = ctypes.c_void_p()
    rc = lib.cov3d_nufft_makeplan(7, 8, 8, 8, 0, 1, 1e-6, 0, ctypes.byref(h))
    assert rc == 1
    assert b"type" in lib.cov3d_last_error()
    with pytest.raises(RuntimeError, match="type must be"):
        native.check(rc)
    assert lib.cov3d_ml_forward(None, None, None, None, None, 0, 4, 40, 8, 1.0, 0, 1, None, None, None, None,
                                None, None) == 1
    assert b"rank" in lib.cov3d_last_error()
    assert lib.cov3d_nufft_setpts(None, 10, None, None, None, None) == 1


def test_fft3c_pad_real_sizes_and_errors():
    """Work size of the centred padded transform, and its argument checks (which run before any
    device call)."""
    native = pkg("_native")
    lib = native.load()
    assert lib.cov3d_fft3c_pad_real_work(128, 256, 20) == 20 * 128 * 129 * (128 + 256)
    assert lib.cov3d_fft3c_pad_real_work(16, 40, 2) == -1   # no pruned grid of 40 cells
    assert lib.cov3d_fft3c_pad_real_work(40, 64, 2) == -1   # 2L > nf + 2
    buf = (ctypes.c_float * 4)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    assert lib.cov3d_fft3c_pad_real(p, p, p, p, 2, 16, 40, None) == 1
    assert b"padded grid" in lib.cov3d_last_error()
    assert lib.cov3d_fft3c_pad_real_adjoint(None, p, p, p, 2, 32, 64, None) == 1
    assert b"null" in lib.cov3d_last_error()


def test_nufft_plan_refuses_cpu():
    nufft_plan = pkg("nufft_plan")
    with pytest.raises(RuntimeError, match="HIP device"):
        nufft_plan.NufftPlan((8, 8, 8), batch_size=2, device=torch.device("cpu"))


def test_discretized_plan_refuses_cpu_tensors():
    nufft_plan = pkg("nufft_plan")
    plan = nufft_plan.NufftPlanDiscretized((4, 4, 4), upsample_factor=2, mode="bilinear")
    plan.setpts(torch.zeros((1, 3, 16)))
    with pytest.raises(RuntimeError, match="HIP device"):
        plan.execute_forward(torch.zeros((1, 8, 8, 8), dtype=torch.complex64))


def test_spec_kwargs_and_half_grid():
    nufft_plan = pkg("nufft_plan")
    spec = nufft_plan.NufftSpec(nufft_plan.NufftType.DISCRETIZED, (8, 8, 8), upsample_factor=2, mode="nearest",
                                batch_size=4)
    assert spec.get_nufft_kwargs() == {"sz": (8, 8, 8), "upsample_factor": 2, "mode": "nearest"}
    assert spec.input_in_fourier
    spec2 = nufft_plan.NufftSpec(nufft_plan.NufftType.FINUFFT, (8, 8, 8), batch_size=3, eps=1e-6)
    assert spec2.get_nufft_kwargs() == {"sz": (8, 8, 8), "batch_size": 3, "eps": 1e-6}
    assert not spec2.input_in_fourier
    pts = torch.arange(2 * 6 * 6 * 3).reshape(2, 6, 6, 3)
    half = nufft_plan.get_half_fourier_grid(pts, (1, 2))
    assert half.shape == (2, 3, 5, 3)


@pytest.mark.parametrize("L,pad", [(7, 14), (8, 16), (8, 15)])
def test_centered_fft_pad_crop_identity(L, pad):
    """ifft3c(fft3c(pad(x)), crop) == x (reference test_torch_wraps.py:42-50)."""
    pf = pkg("projection_funcs")
    x = torch.randn((2, L, L, L), dtype=torch.float64)
    y = pf.centered_ifft3(pf.centered_fft3(x, padding_size=(pad,) * 3), cropping_size=(L,) * 3)
    assert torch.allclose(y.real, x, atol=1e-12)
    assert torch.allclose(pf.crop_tensor(pf.pad_tensor(x, [pad] * 3), [L] * 3), x)


@pytest.mark.parametrize("dtype,code", [(np.complex64, 0), (np.dtype("complex64"), 0), ("complex64", 0),
                                        (torch.complex64, 0), (np.float32, 0), (np.complex128, 1),
                                        (np.dtype("complex128"), 1), ("complex128", 1), (torch.complex128, 1),
                                        (torch.float64, 1)])
def test_reference_side_stub_dtype_mapping(dtype, code):
    """The cufinufft.Plan stand-in maps every spelling of the dtype the reference passes
    (np.complex64 at nufft_plan.py:299-316) to the library's dtype code, and refuses others."""
    import importlib

    stub = importlib.import_module("tests.finufft_stub")
    assert stub.dtype_code(dtype) == code


@pytest.mark.parametrize("bad", [np.int32, "complex32", torch.int64, object])
def test_reference_side_stub_refuses_unknown_dtypes(bad):
    import importlib

    stub = importlib.import_module("tests.finufft_stub")
    with pytest.raises(ValueError):
        stub.dtype_code(bad)
