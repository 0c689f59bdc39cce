"""Reference-side binding of libcov3d_hip.so in place of cufinufft.Plan.

A maintainer who keeps the reference's own cov3d/nufft_plan.py and only swaps the NUFFT library
imports this ``Plan`` instead of ``cufinufft.Plan``.  It accepts the arguments the reference passes
(nufft_plan.py:299-316): ``Plan(nufft_type=2 | 1, n_modes=sz, n_trans=batch_size, eps=eps,
dtype=np.complex64 | np.complex128, gpu_device_id=device.index, gpu_method=1, gpu_sort=0)``, then
``setpts(x, y, z)`` with the three rows of the wrapped (3, M) point tensor (:335-336) and
``execute(signal)`` (:359 for type 2, :387 for type 1).  The GPU options of cufinufft
(``gpu_method``, ``gpu_sort``, ...) have no meaning for this library and are accepted and ignored.

The library is loaded from the package directory unless ``COV3D_HIP_LIB`` names another path.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np
import torch

_LIB_PATH = os.environ.get(
    "COV3D_HIP_LIB", str(Path(__file__).resolve().parent.parent / "3dcovarestimation_amd" / "libcov3d_hip.so"))
_lib = None


def _load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(_LIB_PATH)
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        lib.cov3d_nufft_makeplan.argtypes = [ctypes.c_int, i64, i64, i64, ctypes.c_int, i64, ctypes.c_double,
                                             ctypes.c_int, ctypes.POINTER(vp)]
        lib.cov3d_nufft_setpts.argtypes = [vp, i64, vp, vp, vp, vp]
        lib.cov3d_nufft_execute.argtypes = [vp, vp, vp, ctypes.c_int, i64, vp]
        lib.cov3d_nufft_destroy.argtypes = [vp]
        lib.cov3d_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(_load().cov3d_last_error().decode())


def dtype_code(dtype) -> int:
    """0 for single precision (np.complex64, np.float32, "complex64", np.dtype("complex64"),
    torch.complex64, torch.float32), 1 for double precision; anything else raises ValueError."""
    if isinstance(dtype, torch.dtype):
        table = {torch.complex64: 0, torch.float32: 0, torch.complex128: 1, torch.float64: 1}
        if dtype in table:
            return table[dtype]
        raise ValueError(f"unsupported NUFFT dtype {dtype}")
    try:
        d = np.dtype(dtype)
    except TypeError as e:
        raise ValueError(f"unsupported NUFFT dtype {dtype!r}") from e
    if d in (np.dtype(np.complex64), np.dtype(np.float32)):
        return 0
    if d in (np.dtype(np.complex128), np.dtype(np.float64)):
        return 1
    raise ValueError(f"unsupported NUFFT dtype {dtype!r}")


class Plan:
    """Drop-in for ``cufinufft.Plan`` as cov3d/nufft_plan.py uses it."""

    def __init__(self, nufft_type, n_modes, n_trans=1, eps=1e-6, isign=None, dtype="complex64",
                 gpu_device_id=None, **gpu_options):
        if nufft_type not in (1, 2):
            raise ValueError("nufft_type must be 1 or 2")
        if len(n_modes) != 3:
            raise ValueError("only 3-D plans are supported")
        self.type = int(nufft_type)
        self.n_modes = tuple(int(n) for n in n_modes)
        self.n_trans = int(n_trans)
        self.dtype_code = dtype_code(dtype)
        self.real_dtype = torch.float32 if self.dtype_code == 0 else torch.float64
        self.complex_dtype = torch.complex64 if self.dtype_code == 0 else torch.complex128
        self.device = torch.device("cuda", torch.cuda.current_device() if gpu_device_id is None else gpu_device_id)
        isign = (-1 if self.type == 2 else 1) if isign is None else int(isign)
        self._h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _check(_load().cov3d_nufft_makeplan(self.type, *self.n_modes, isign, self.n_trans, float(eps),
                                                self.dtype_code, ctypes.byref(self._h)))
        self._pts = None

    def setpts(self, x, y=None, z=None):
        pts = []
        for t in (x, y, z):
            if not isinstance(t, torch.Tensor) or t.device != self.device:
                raise ValueError(f"points must be tensors on {self.device}")
            if t.dtype != self.real_dtype:
                raise ValueError(f"points must be {self.real_dtype} for this plan, got {t.dtype}")
            pts.append(t.contiguous())
        if not (pts[0].numel() == pts[1].numel() == pts[2].numel()):
            raise ValueError("x, y and z must have the same length")
        self._pts = tuple(pts)  # kept alive until the next setpts, like cufinufft's _references
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            _check(_load().cov3d_nufft_setpts(self._h, pts[0].numel(), *(p.data_ptr() for p in pts), stream))

    def execute(self, data, out=None):
        if self._pts is None:
            raise RuntimeError("setpts must be called before execute")
        M = self._pts[0].numel()
        if not isinstance(data, torch.Tensor) or data.device != self.device:
            raise ValueError(f"data must be a tensor on {self.device}")
        if self.type == 2:
            real_in = not data.is_complex()
            want = self.real_dtype if real_in else self.complex_dtype
            if data.dtype != want or data.numel() != self.n_trans * int(np.prod(self.n_modes)):
                raise ValueError(f"type 2 input must be {self.n_trans} x {self.n_modes} of {self.complex_dtype}")
            f = data.contiguous()
            c = out if out is not None else torch.empty((self.n_trans, M), dtype=self.complex_dtype,
                                                        device=self.device)
            in_out = (c, f, int(real_in))
        else:
            if data.dtype != self.complex_dtype or data.numel() != self.n_trans * M:
                raise ValueError(f"type 1 input must be {self.n_trans} x {M} values of {self.complex_dtype}")
            c = data.contiguous()
            f = out if out is not None else torch.empty((self.n_trans, *self.n_modes), dtype=self.complex_dtype,
                                                        device=self.device)
            in_out = (c, f, 0)
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            _check(_load().cov3d_nufft_execute(self._h, in_out[0].data_ptr(), in_out[1].data_ptr(), in_out[2],
                                               self.n_trans, stream))
        return c if self.type == 2 else f

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize(self.device)
                _load().cov3d_nufft_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None
