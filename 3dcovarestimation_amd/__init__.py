"""MI355X-native (gfx950) COV3D hot path.

Drop-in host modules mirroring the reference's interfaces:
``nufft_plan``, ``projection_funcs``, ``covar``, ``covar_sgd``, ``mean``,
``wiener_coords``, ``fsc_utils``, ``dataset``; the compute runs in the HIP
C-ABI library ``libcov3d_hip.so`` (see ``_native``).
"""

from . import _native  # noqa: F401

__all__ = [
    "_native",
    "nufft_plan",
    "projection_funcs",
    "covar",
    "covar_sgd",
    "mean",
    "wiener_coords",
    "fsc_utils",
    "dataset",
    "ml_ops",
    "precond",
]
