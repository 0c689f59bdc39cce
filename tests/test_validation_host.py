"""CPU checks of the subspace metric used by the SGD agreement test and the bench."""

import torch

from tests.conftest import pkg


def test_subspace_distance_properties():
    val = pkg("validation")
    g = torch.Generator().manual_seed(1)
    Q, _ = torch.linalg.qr(torch.randn(50, 4, generator=g, dtype=torch.float64))
    R, _ = torch.linalg.qr(torch.randn(4, 4, generator=g, dtype=torch.float64))
    assert val.subspace_distance(Q, Q @ R) < 1e-12  # basis change inside the subspace
    P, _ = torch.linalg.qr(torch.randn(50, 50, generator=g, dtype=torch.float64))
    A, B = P[:, :4], P[:, 4:8]
    assert abs(val.subspace_distance(A, B) - 1.0) < 1e-12  # orthogonal subspaces
    # one rotated direction by angle t out of 4: RMS sine = sin(t) / 2
    t = 0.3
    C = A.clone()
    C[:, 0] = torch.cos(torch.tensor(t, dtype=torch.float64)) * A[:, 0] + torch.sin(torch.tensor(t, dtype=torch.float64)) * B[:, 0]
    assert abs(val.subspace_distance(A, C) - float(torch.sin(torch.tensor(t))) / 2) < 1e-6
