"""Shared test setup: the `gpu` marker, import helpers for the package and the oracle."""

import importlib
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X / gfx950)")


def pkg(sub: str = ""):
    """Imports the package (its directory name starts with a digit, so by importlib)."""
    name = "3dcovarestimation_amd" + (("." + sub) if sub else "")
    return importlib.import_module(name)


def oracle():
    return importlib.import_module("oracle.oracle")


def gpu_available() -> bool:
    import torch

    return torch.cuda.is_available()


@pytest.fixture(scope="session")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def rel_l2(a, b) -> float:
    import numpy as np

    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
