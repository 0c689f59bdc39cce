"""cryoDRGN pose and CTF files (cov3d/source.py:89-105, :476-542; cov3d/poses.py:11-41).

A cryoDRGN particle set ships ``poses.pkl`` = (rotations (N, 3, 3), translations (N, 2) as a
fraction of the box) and ``ctf.pkl`` = (N, 9) float rows
``[D, Apix, defocus_u A, defocus_v A, defocus_angle deg, voltage kV, Cs mm, amplitude contrast,
phase_shift deg]``.  The reference reads them with ``pickle.load`` and cryoDRGN's
``load_ctf_for_training`` and converts the poses to ASPIRE's convention (transposed rotations,
translations in pixels); it writes them back the same way when it saves a simulation.

These files hold numpy arrays and tuples only, so they are read here with an unpickler that
resolves nothing but numpy's array and dtype reconstructors: a file naming any other global is
rejected before anything from it runs.  Writing uses the standard pickle protocol.
"""

from __future__ import annotations

import io
import pickle
from typing import Tuple

import numpy as np

# the globals a pickled numpy array (any protocol, numpy 1.x or 2.x module paths) refers to
_ALLOWED = {
    ("numpy", "ndarray"),
    ("numpy", "dtype"),
    ("numpy.core.multiarray", "_reconstruct"),
    ("numpy._core.multiarray", "_reconstruct"),
    ("numpy.core.multiarray", "scalar"),
    ("numpy._core.multiarray", "scalar"),
    ("numpy.core.numeric", "_frombuffer"),
    ("numpy._core.numeric", "_frombuffer"),
}


class _ArrayUnpickler(pickle.Unpickler):
    def find_class(self, module: str, name: str):
        if (module, name) not in _ALLOWED:
            raise pickle.UnpicklingError(f"{module}.{name} is not a numpy array constructor; refusing to load")
        return super().find_class(module, name)


def load_array_pickle(path: str):
    """Unpickles a file that holds numpy arrays, numbers, strings, tuples, lists and dicts only."""
    with open(path, "rb") as fid:
        return _ArrayUnpickler(io.BytesIO(fid.read())).load()


def pose_cryodrgn_to_aspire(poses: Tuple[np.ndarray, np.ndarray], L: int) -> Tuple[np.ndarray, np.ndarray]:
    """(rotations, translations as box fractions) -> (transposed rotations, offsets in pixels)
    (poses.py:11-24)."""
    return np.transpose(np.asarray(poses[0]), (0, 2, 1)), np.asarray(poses[1]) * L


def pose_aspire_to_cryodrgn(rots: np.ndarray, offsets: np.ndarray, L: int) -> Tuple[np.ndarray, np.ndarray]:
    """Inverse of ``pose_cryodrgn_to_aspire`` (poses.py:27-41)."""
    return np.transpose(np.asarray(rots), (0, 2, 1)), np.asarray(offsets) / L


def load_poses(path: str, L: int) -> Tuple[np.ndarray, np.ndarray]:
    """``poses.pkl`` -> ASPIRE rotations (N, 3, 3) and pixel offsets (N, 2) (source.py:101-103)."""
    poses = load_array_pickle(path)
    if not (isinstance(poses, (tuple, list)) and len(poses) >= 2):
        raise ValueError(f"{path}: expected a (rotations, translations) pair")
    rots, offsets = pose_cryodrgn_to_aspire(poses, L)
    if rots.ndim != 3 or rots.shape[1:] != (3, 3) or offsets.shape != (rots.shape[0], 2):
        raise ValueError(f"{path}: rotations {rots.shape} / translations {offsets.shape} are not (N,3,3) / (N,2)")
    return rots, offsets


def save_poses(path: str, rots: np.ndarray, offsets: np.ndarray, L: int) -> None:
    """Writes ``poses.pkl`` from ASPIRE rotations and pixel offsets (source.py:533-536)."""
    with open(path, "wb") as fid:
        pickle.dump(pose_aspire_to_cryodrgn(rots, offsets, L), fid)


def load_ctf_for_training(D: int, path: str) -> np.ndarray:
    """``ctf.pkl`` (N, 9) -> (N, 8) [Apix, dfu, dfv, dfang, kV, Cs, w, phase] for images of D
    pixels: cryoDRGN's published rule replaces the stored box size by D and rescales the pixel
    size to keep the physical box (Apix = D0 * Apix0 / D), then drops the size column."""
    if D % 2:
        raise ValueError("cryoDRGN CTF files are used with even box sizes")
    ctf = np.array(load_array_pickle(path), dtype=np.float64)
    if ctf.ndim != 2 or ctf.shape[1] != 9:
        raise ValueError(f"{path}: expected (N, 9) CTF rows, got {ctf.shape}")
    apix = ctf[0, 0] * ctf[0, 1] / D
    ctf[:, 0] = D
    ctf[:, 1] = apix
    return ctf[:, 1:]


def save_ctf(path: str, ctf_params: np.ndarray, L: int) -> None:
    """Writes ``ctf.pkl`` (N, 9) from this package's (N, 7) CTF parameters
    [dfu A, dfv A, ang rad, kV, Cs mm, w, Apix] as the reference's simulation writer does
    (source.py:476-494): the angle is stored in degrees as it is, with no phase shift.  (Read back
    through ``ctf_params_from_cryodrgn`` the angle takes cryoDRGN's convention, as in the
    reference.)"""
    p = np.asarray(ctf_params, dtype=np.float64)
    out = np.zeros((p.shape[0], 9))
    out[:, 0] = L
    out[:, 1] = p[:, 6]
    out[:, 2] = p[:, 0]
    out[:, 3] = p[:, 1]
    out[:, 4] = p[:, 2] / np.pi * 180
    out[:, 5] = p[:, 3]
    out[:, 6] = p[:, 4]
    out[:, 7] = p[:, 5]
    with open(path, "wb") as fid:
        pickle.dump(out, fid)


def ctf_params_from_cryodrgn(params8: np.ndarray) -> np.ndarray:
    """(N, 8) cryoDRGN training rows -> (N, 7) [dfu, dfv, ang rad, kV, Cs, w, Apix] for
    ``dataset.ctf_from_params``, giving the CTF images the reference's cryoDRGN path computes
    (source.py:156-170 with cryoDRGN's compute_ctf).

    cryoDRGN's CTF, sqrt(1 - w^2) sin(g) - w cos(g) with g = 2 pi (-df lambda s^2 / 2 +
    Cs lambda^3 s^4 / 4) - phase, is the kernel's formula when the phase shift is zero (a phase
    plate is not supported).  The two differ in the astigmatism angle: the reference's frequency
    lattice puts x on image axis 0, so its polar angle is theta_c = atan2(f_axis1, f_axis0), while
    the kernel uses theta = atan2(f_axis0, f_axis1) = pi/2 - theta_c.  cos(2 (theta - ang)) equals
    cos(2 (theta_c - dfang)) for every theta_c when ang = -dfang - pi/2 (mod pi)."""
    p = np.asarray(params8, dtype=np.float64)
    if p.ndim != 2 or p.shape[1] != 8:
        raise ValueError(f"expected (N, 8) rows, got {p.shape}")
    if np.any(p[:, 7] != 0):
        raise NotImplementedError("CTF phase shifts (phase plates) are not supported")
    ang = -p[:, 3] * np.pi / 180 - np.pi / 2
    out = np.stack([p[:, 1], p[:, 2], ang, p[:, 4], p[:, 5], p[:, 6], p[:, 0]], axis=1)
    return out.astype(np.float32)


def covar_dataset_from_cryodrgn(images, poses_path: str, ctf_path: str, noise_var: float, device=None,
                                indices=None):
    """A ``CovarDataset`` from an image stack (N, L, L) and cryoDRGN ``poses.pkl`` / ``ctf.pkl``,
    as the reference's cryoDRGN source builds its rotations, offsets and CTF parameters
    (source.py:52-105): ASPIRE-convention rotations and pixel offsets, CTFs from the parameters
    (converted by ``ctf_params_from_cryodrgn``), ``indices`` selecting particles of all three."""
    import torch

    from .dataset import CovarDataset

    L = images.shape[-1]
    rots, offsets = load_poses(poses_path, L)
    ctf = ctf_params_from_cryodrgn(load_ctf_for_training(L, ctf_path))
    if not (len(rots) == len(ctf) == images.shape[0]):
        raise ValueError(f"{images.shape[0]} images, {len(rots)} poses, {len(ctf)} CTF rows")
    if indices is not None:
        idx = np.asarray(indices)
        rots, offsets, ctf = rots[idx], offsets[idx], ctf[idx]
        images = images[torch.as_tensor(idx, device=images.device)]
    dev = torch.device(device) if device is not None else images.device
    dtype = images.real.dtype
    return CovarDataset(images.to(dev), torch.tensor(rots, dtype=dtype, device=dev), noise_var,
                        ctf_params=torch.tensor(ctf, device=dev),
                        offsets=torch.tensor(offsets, dtype=dtype, device=dev))
