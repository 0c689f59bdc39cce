"""Host-side pose pieces: PoseModule.split_module / merge_modules / update_from_modules
(poses.py:350-461) and the row-wise L-BFGS (newton_opt.py:142-276), on the CPU.

The batched L-BFGS is checked three ways: against a row-by-row implementation with per-row deques
(the reference's data structure), by the secant property (on a 1-D quadratic the second step of
L-BFGS lands on the minimiser exactly), and by convergence on separable strictly convex quadratics
in three dimensions."""

from collections import deque

import pytest
import torch

from tests.conftest import pkg


def _pose_module(n, L=8, contrast=False, seed=0):
    poses = pkg("poses")
    g = torch.Generator().manual_seed(seed)
    m = poses.PoseModule(torch.randn((n, 3), generator=g, dtype=torch.float64) * 0.5,
                         torch.randn((n, 2), generator=g, dtype=torch.float64), L, dtype=torch.float64,
                         use_contrast=contrast)
    if contrast:
        m.set_contrasts(torch.rand((n, 1), generator=g, dtype=torch.float64) + 0.5)
    return m


@pytest.mark.parametrize("contrast", [False, True])
@pytest.mark.parametrize("n", [10, 11])
def test_split_then_merge_is_identity(n, contrast):
    poses = pkg("poses")
    m = _pose_module(n, contrast=contrast)
    perm = torch.randperm(n, generator=torch.Generator().manual_seed(1))
    a, b = m.split_module(perm)
    assert a.get_rotvecs().shape[0] == n // 2 and b.get_rotvecs().shape[0] == n - n // 2
    # half k holds exactly the rows perm[k-th part], in that order
    assert torch.equal(a.get_rotvecs(), m.get_rotvecs()[perm[: n // 2]])
    assert torch.equal(b.get_offsets(), m.get_offsets()[perm[n // 2:]])
    merged = poses.PoseModule.merge_modules(a, b, perm)
    assert torch.equal(merged.get_rotvecs(), m.get_rotvecs())
    assert torch.equal(merged.get_offsets(), m.get_offsets())
    if contrast:
        assert torch.equal(a.get_contrasts(), m.get_contrasts()[perm[: n // 2]])
        assert torch.equal(merged.get_contrasts(), m.get_contrasts())
    # refined halves go back to their images
    with torch.no_grad():
        a.set_rotvecs(a.get_rotvecs() + 1.0)
        b.set_offsets(b.get_offsets() - 2.0)
    target = _pose_module(n, contrast=contrast, seed=5)
    target.update_from_modules(a, b, perm)
    exp_rot = m.get_rotvecs().clone()
    exp_rot[perm[: n // 2]] += 1.0
    exp_off = m.get_offsets().clone()
    exp_off[perm[n // 2:]] -= 2.0
    assert torch.equal(target.get_rotvecs(), exp_rot)
    assert torch.equal(target.get_offsets(), exp_off)
    # the halves are copies: changing them left the original alone
    assert not torch.equal(m.get_rotvecs()[perm[: n // 2]], a.get_rotvecs())


def test_split_without_permutation_keeps_order():
    m = _pose_module(9)
    a, b = m.split_module()
    assert torch.equal(torch.cat([a.get_rotvecs(), b.get_rotvecs()]), m.get_rotvecs())


class _RowLBFGS:
    """Row-by-row L-BFGS with a deque history per row (the structure of newton_opt.py:142-276)."""

    def __init__(self, m, lr, limit):
        self.m, self.lr, self.limit = m, lr, limit
        self.S, self.Y, self.px, self.pg = {}, {}, {}, {}

    def step(self, w, rows, grads):
        for i, g in zip(rows.tolist(), grads):
            x = w[i].clone()
            S = self.S.setdefault(i, deque(maxlen=self.m))
            Y = self.Y.setdefault(i, deque(maxlen=self.m))
            if i in self.px:
                s, y = x - self.px[i], g - self.pg[i]
                if torch.dot(s, y) > 1e-10:
                    S.append(s)
                    Y.append(y)
            self.px[i], self.pg[i] = x, g.clone()
            q = g.clone()
            al, rh = [], []
            for s, y in reversed(list(zip(S, Y))):
                rho = 1.0 / torch.dot(y, s)
                a = rho * torch.dot(s, q)
                q = q - a * y
                al.append(a)
                rh.append(rho)
            h0 = torch.dot(S[-1], Y[-1]) / torch.dot(Y[-1], Y[-1]) if S else 1.0
            r = h0 * q
            for s, y, a, rho in zip(S, Y, reversed(al), reversed(rh)):
                b = rho * torch.dot(y, r)
                r = r + s * (a - b)
            st = self.lr * r
            if self.limit is not None:
                st = torch.clamp(st, -self.limit, self.limit)
            w[i] = x - st


@pytest.mark.parametrize("limit", [None, 0.05])
@pytest.mark.parametrize("history", [3, 10])
def test_lbfgs_matches_row_by_row(history, limit):
    newton_opt = pkg("newton_opt")
    g = torch.Generator().manual_seed(2)
    n, d = 12, 3
    A = torch.randn((n, d, d), generator=g, dtype=torch.float64)
    A = A @ A.transpose(1, 2) + 0.5 * torch.eye(d, dtype=torch.float64)
    A[3] = -0.05 * torch.eye(d, dtype=torch.float64)  # a concave row: its curvature pairs are rejected
    b = torch.randn((n, d), generator=g, dtype=torch.float64)
    emb = torch.nn.Embedding(n, d, sparse=True, _weight=torch.randn((n, d), generator=g, dtype=torch.float64))
    opt = newton_opt.BlockwiseLBFGS([{"params": [emb.weight], "lr": 0.7, "step_size_limit": limit}],
                                    history_size=history)
    ref = _RowLBFGS(history, 0.7, limit)
    w_ref = emb.weight.detach().clone()
    for it in range(3 * history + 4):  # past a full history
        rows = torch.randperm(n, generator=g)[: 7]
        rows = torch.cat([rows, rows[:2]])  # duplicate indices: their gradients add
        emb.weight.grad = None
        x = emb(rows)
        loss = (0.5 * torch.einsum("ki,kij,kj->k", x, A[rows], x) - (b[rows] * x).sum(1)).sum()
        loss.backward()
        # the reference sees one summed gradient per row
        ur = torch.unique(rows)
        gr = (torch.einsum("kij,kj->ki", A[ur], w_ref[ur]) - b[ur]) * torch.stack(
            [(rows == i).sum() for i in ur]).to(torch.float64)[:, None]
        opt.step()
        ref.step(w_ref, ur, gr)
        torch.testing.assert_close(emb.weight.detach(), w_ref, rtol=1e-8, atol=1e-10)


def test_lbfgs_secant_step_is_exact_in_one_dimension():
    """On f(x) = a/2 (x - c)^2 the second step is H g = g / a (H0 = s.y / y.y = 1/a, and the
    recursion leaves H0's direction unchanged in one dimension): at lr 1 it lands on c."""
    newton_opt = pkg("newton_opt")
    n = 6
    a = torch.linspace(0.5, 3.0, n, dtype=torch.float64)
    c = torch.linspace(-1.0, 2.0, n, dtype=torch.float64)
    x = torch.nn.Parameter(torch.zeros((n, 1), dtype=torch.float64))
    opt = newton_opt.BlockwiseLBFGS([x], lr=1.0)
    for _ in range(2):
        opt.zero_grad()
        (0.5 * a * (x[:, 0] - c) ** 2).sum().backward()
        opt.step()
    torch.testing.assert_close(x.detach()[:, 0], c, rtol=0, atol=1e-12)


def test_lbfgs_converges_on_convex_quadratics():
    newton_opt = pkg("newton_opt")
    g = torch.Generator().manual_seed(3)
    n, d = 32, 3
    Q = torch.linalg.qr(torch.randn((n, d, d), generator=g, dtype=torch.float64))[0]
    lam = torch.rand((n, d), generator=g, dtype=torch.float64) * 9 + 1  # condition <= 10
    A = Q @ torch.diag_embed(lam) @ Q.transpose(1, 2)
    b = torch.randn((n, d), generator=g, dtype=torch.float64)
    xs = torch.linalg.solve(A, b)
    x = torch.nn.Parameter(torch.zeros((n, d), dtype=torch.float64))
    opt = newton_opt.BlockwiseLBFGS([x], lr=1.0, step_size_limit=1.0)
    for _ in range(60):
        opt.zero_grad()
        (0.5 * torch.einsum("ki,kij,kj->k", x, A, x) - (b * x).sum(1)).sum().backward()
        opt.step()
    assert float((x.detach() - xs).norm() / xs.norm()) < 1e-8


def test_lbfgs_rejects_bad_arguments():
    newton_opt = pkg("newton_opt")
    p = torch.nn.Parameter(torch.zeros((2, 3)))
    with pytest.raises(ValueError):
        newton_opt.BlockwiseLBFGS([p], lr=0.0)
    with pytest.raises(ValueError):
        newton_opt.BlockwiseLBFGS([p], history_size=0)
