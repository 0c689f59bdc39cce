"""The no-allocation Toeplitz entry point (cov3d_toeplitz_apply_noalloc) against the allocating one
(cov3d_toeplitz_apply), through the C ABI: the same pruned passes on a static device workspace must
give bit-identical results, on two calls in a row with the second one smaller, and sizes beyond the
workspace (or grids the pruned passes do not cover) are refused with a status and a message."""

import numpy as np
import pytest
import torch

from tests.conftest import oracle, pkg, rel_l2

pytestmark = pytest.mark.gpu


def _lib():
    return pkg("_native").load()


def test_noalloc_toeplitz_matches_allocating_entry(dev):
    native = pkg("_native")
    lib = _lib()
    o = oracle()
    st = native.stream_ptr(dev)
    g = torch.Generator(device=dev).manual_seed(3)
    for L, nf in ((64, 128), (31, 64)):
        x = torch.complex(torch.randn(L**3, device=dev, generator=g), torch.randn(L**3, device=dev, generator=g))
        lhs = torch.complex(torch.randn((nf,) * 3, device=dev, generator=g),
                            torch.randn((nf,) * 3, device=dev, generator=g))
        y_ref = torch.empty_like(x)
        y = torch.full_like(x, float("nan"))
        print(f"case L={L} nf={nf}", flush=True)
        native.check(lib.cov3d_toeplitz_apply(x.data_ptr(), lhs.data_ptr(), y_ref.data_ptr(), L, nf, st))
        native.check(lib.cov3d_toeplitz_apply_noalloc(x.data_ptr(), lhs.data_ptr(), y.data_ptr(), L, nf, st))
        torch.cuda.synchronize(dev)
        assert torch.equal(y, y_ref)
        ref = o.toeplitz_apply(x.cpu().numpy().astype(np.complex128), lhs.cpu().numpy().astype(np.complex128), L)
        assert rel_l2(y.cpu().numpy(), ref.ravel()) < 1e-5
    x = torch.zeros(8, dtype=torch.complex64, device=dev)
    for L, nf, msg in ((256, 512, b"exceeds the workspace"), (40, 96, b"nf must be")):
        rc = lib.cov3d_toeplitz_apply_noalloc(x.data_ptr(), x.data_ptr(), x.data_ptr(), L, nf, st)
        assert rc == 1
        assert msg in lib.cov3d_last_error()
