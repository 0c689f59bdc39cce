"""ctypes binding of the gfx950 C-ABI library ``libcov3d_hip.so``.

The library is the only compute path of this package: every operator raises
``RuntimeError`` when it is missing or when a HIP device is not available;
there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch

_LIB_PATH = Path(__file__).resolve().parent / "libcov3d_hip.so"
_lib = None

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p

# name -> argtypes (restype is int unless listed in _RESTYPES)
_SIGNATURES = {
    "cov3d_last_error": [],
    "cov3d_version": [],
    "cov3d_nufft_makeplan": [c_int, c_i64, c_i64, c_i64, c_int, c_i64, c_dbl, c_int, ctypes.POINTER(c_vp)],
    "cov3d_nufft_setpts": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp],
    "cov3d_nufft_execute": [c_vp, c_vp, c_vp, c_int, c_i64, c_vp],
    "cov3d_nufft_execute_dir": [c_vp, c_int, c_vp, c_vp, c_int, c_i64, c_vp],
    "cov3d_nufft_destroy": [c_vp],
    "cov3d_nufft_sorted_input": [c_vp, c_i64, ctypes.POINTER(c_vp), ctypes.POINTER(c_i64), ctypes.POINTER(c_vp)],
    "cov3d_nufft_adjoint_presorted": [c_vp, c_vp, c_int, c_i64, c_vp],
    "cov3d_nufft_info": [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_int), ctypes.POINTER(c_dbl)],
    "cov3d_slice_forward": [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_int, c_int, c_vp],
    "cov3d_slice_adjoint": [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_int, c_int, c_i64, c_int, c_vp],
    "cov3d_slice_point_grad": [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_int, c_int, c_vp],
    "cov3d_slice_plan_create": [ctypes.POINTER(c_vp)],
    "cov3d_slice_plan_setpts": [c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_int, c_int, c_vp],
    "cov3d_slice_plan_forward": [c_vp, c_vp, c_vp, c_i64, c_vp],
    "cov3d_slice_plan_adjoint": [c_vp, c_vp, c_vp, c_i64, c_vp],
    "cov3d_slice_plan_forward_split": [c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp],
    "cov3d_slice_plan_adjoint_split": [c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_i64, c_vp],
    "cov3d_slice_grid_from_pts": [c_vp, c_vp, c_i64, c_i64, ctypes.c_double, ctypes.c_double, c_int, c_vp],
    "cov3d_slice_boundary_pixels": [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, c_vp],
    "cov3d_herm_solve_c128": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_vp],
    "cov3d_slice_plan_destroy": [c_vp],
    "cov3d_ml_grad_images_complex": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_i64,
                                     c_i64, c_dbl, c_dbl, c_vp],
    "cov3d_cg_update": [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp],
    "cov3d_cg_direction": [c_vp, c_vp, c_i64, c_vp, c_vp, c_vp],
    "cov3d_ml_forward": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_dbl, c_int, c_int, c_vp, c_vp,
                         c_vp, c_vp, c_vp, c_vp],
    "cov3d_ml_grad_images": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_i64, c_i64, c_dbl, c_dbl,
                             c_vp],
    "cov3d_ml_grad_sorted": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64,
                             c_dbl, c_dbl, c_vp],
    "cov3d_pts_rot": [c_vp, c_vp, c_i64, c_i64, c_vp],
    "cov3d_points_wrap": [c_vp, c_vp, c_i64, c_i64, c_int, c_vp],
    "cov3d_ctf_eval": [c_vp, c_vp, c_i64, c_i64, c_vp],
    "cov3d_vdot_c64_f64": [c_vp, c_vp, c_i64, c_vp, c_vp],
    "cov3d_toeplitz_apply": [c_vp, c_vp, c_vp, c_i64, c_i64, c_vp],
    "cov3d_toeplitz_apply_noalloc": [c_vp, c_vp, c_vp, c_i64, c_i64, c_vp],
    "cov3d_fft3c_pad_real_work": [c_i64, c_i64, c_i64],
    "cov3d_fft3c_pad_real": [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp],
    "cov3d_fft3c_pad_real_adjoint": [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp],
    "cov3d_mlp_forward": [c_vp, c_vp, c_vp, c_i64, c_i64, c_vp],
    "cov3d_mlp_backward": [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_vp],
    "cov3d_timing_enable": [c_int],
    "cov3d_timing_read": [c_int, ctypes.POINTER(c_dbl), ctypes.POINTER(c_i64)],
}
_RESTYPES = {"cov3d_last_error": ctypes.c_char_p, "cov3d_version": ctypes.c_char_p,
             "cov3d_fft3c_pad_real_work": c_i64}

EXPORTED = tuple(_SIGNATURES)


def lib_path() -> Path:
    return _LIB_PATH


def load(path: str | os.PathLike | None = None):
    """Loads the library (once) and declares every entry point's signature."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else _LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"cov3d HIP library not found at {p}; build it with "
            "`python 3dcovarestimation_amd/build_native.py` (there is no CPU fallback)"
        )
    lib = ctypes.CDLL(str(p))
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, ctypes.c_int)
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> None:
    """Raises RuntimeError carrying the library's message for a nonzero status."""
    if rc != 0:
        msg = load().cov3d_last_error().decode()
        raise RuntimeError(f"cov3d HIP call failed (status {rc}): {msg}")


def require_device(t: torch.Tensor, name: str = "tensor") -> None:
    if not t.is_cuda:
        raise RuntimeError(
            f"{name} must live on a HIP device: this package runs only on the gfx950 kernels "
            "(no CPU fallback)"
        )


def dense(t: torch.Tensor) -> torch.Tensor:
    """A contiguous tensor whose memory holds its values, for a raw pointer.  Torch's lazy
    conjugate and negative views (gradients of ``conj()`` arrive as such) keep the unconjugated
    values in memory; they are materialised here."""
    return t.resolve_conj().resolve_neg().contiguous()


def stream_ptr(device: torch.device | None = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


TIMING_KINDS = {"interp": 0, "spread": 1, "deconv_pad": 2, "deconv_crop": 3, "fft": 4, "sort": 5, "permute": 6}


def timing_enable(on: bool) -> None:
    check(load().cov3d_timing_enable(1 if on else 0))


def timing_read(kind: str):
    """(total ms, launch count) of one NUFFT kernel class since timing_enable(True)."""
    tot = c_dbl()
    cnt = c_i64()
    check(load().cov3d_timing_read(TIMING_KINDS[kind], ctypes.byref(tot), ctypes.byref(cnt)))
    return tot.value, cnt.value
