"""Row-wise L-BFGS for embedding parameters (newton_opt.py:142-276).

Every row of a parameter (one image's rotation vector, say) is its own small problem with its own
curvature history, as in the reference's ``BlockwiseLBFGS``.  The reference walks the rows of a
step in a Python loop (one ``.item()`` and a dozen small tensor operations per row); here the
histories of all rows live in device tensors, ``S`` and ``Y`` of shape (n, m, d) ordered oldest to
newest with a count per row, and one step runs the two-loop recursion for every row of the step at
once: m iterations of (k, d) tensor operations and no host synchronisation for a sparse gradient.
"""

from typing import Any, Callable, Dict, Optional

import torch
from torch.optim.optimizer import Optimizer


class BlockwiseLBFGS(Optimizer):
    """L-BFGS applied independently to every row of each parameter (newton_opt.py:142-276).

    For each row i that has a gradient in this step:
      * s = x_i - x_i(prev), y = g_i - g_i(prev) enter the row's history (the last ``history_size``
        pairs) when s.y > 1e-10 and the row was seen before;
      * the two-loop recursion over the history, newest pair first, with the initial Hessian
        s.y / y.y of the newest pair (1 without history), gives r = H g_i;
      * x_i -= lr * r, each component clamped to +-step_size_limit when one is set.
    The rows of a step are the indices of a sparse gradient (duplicates summed) or the nonzero rows
    of a dense one.
    """

    def __init__(self, params: Any, lr: float = 1.0, history_size: int = 10,
                 step_size_limit: Optional[float] = None) -> None:
        if lr <= 0:
            raise ValueError(f"Invalid learning rate: {lr}")
        if history_size < 1:
            raise ValueError(f"Invalid history size: {history_size}")
        super().__init__(params, dict(lr=lr, step_size_limit=step_size_limit))
        self.history_size = history_size

    def _row_state(self, p: torch.Tensor) -> Dict[str, torch.Tensor]:
        st = self.state[p]
        if not st:
            n = p.shape[0]
            d = p[0].numel()
            m = self.history_size
            st["S"] = torch.zeros((n, m, d), dtype=p.dtype, device=p.device)
            st["Y"] = torch.zeros((n, m, d), dtype=p.dtype, device=p.device)
            st["cnt"] = torch.zeros(n, dtype=torch.long, device=p.device)
            st["prev_x"] = torch.zeros((n, d), dtype=p.dtype, device=p.device)
            st["prev_g"] = torch.zeros((n, d), dtype=p.dtype, device=p.device)
            st["seen"] = torch.zeros(n, dtype=torch.bool, device=p.device)
        return st

    @torch.no_grad()
    def step(self, closure: Optional[Callable[[], torch.Tensor]] = None) -> Optional[torch.Tensor]:
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        for group in self.param_groups:
            lr = group["lr"]
            limit = group["step_size_limit"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                self._step_param(p, lr, limit)
        return loss

    def _step_param(self, p: torch.Tensor, lr: float, limit: Optional[float]) -> None:
        grad = p.grad
        n = p.shape[0]
        if grad.is_sparse:
            grad = grad.coalesce()
            rows = grad.indices()[0]
            g = grad.values().reshape(rows.shape[0], -1)
        else:
            g_all = grad.reshape(n, -1)
            rows = torch.nonzero(g_all.norm(dim=1), as_tuple=True)[0]
            g = g_all[rows]
        if rows.numel() == 0:
            return
        st = self._row_state(p)
        m = self.history_size
        x = p.data.reshape(n, -1)[rows]
        S, Y, cnt = st["S"][rows], st["Y"][rows], st["cnt"][rows]
        # the new pair of every row seen before that satisfies the curvature condition
        s_new = x - st["prev_x"][rows]
        y_new = g - st["prev_g"][rows]
        take = st["seen"][rows] & ((s_new * y_new).sum(dim=1) > 1e-10)
        full = take & (cnt == m)
        # a full history drops its oldest pair: shift left by one
        S = torch.where(full[:, None, None], torch.roll(S, -1, dims=1), S)
        Y = torch.where(full[:, None, None], torch.roll(Y, -1, dims=1), Y)
        pos = torch.clamp(torch.where(full, cnt - 1, cnt), max=m - 1)
        slot = torch.nn.functional.one_hot(pos, m).to(torch.bool) & take[:, None]
        S = torch.where(slot[:, :, None], s_new[:, None, :], S)
        Y = torch.where(slot[:, :, None], y_new[:, None, :], Y)
        cnt = torch.where(take & (cnt < m), cnt + 1, cnt)
        st["S"][rows], st["Y"][rows], st["cnt"][rows] = S, Y, cnt
        st["prev_x"][rows] = x
        st["prev_g"][rows] = g
        st["seen"][rows] = True
        # two-loop recursion, every row at once (entries j >= cnt are inactive)
        valid = torch.arange(m, device=p.device)[None, :] < cnt[:, None]  # (k, m)
        sy = (S * Y).sum(dim=2)
        rho = torch.where(valid, 1.0 / torch.where(valid, sy, torch.ones_like(sy)), torch.zeros_like(sy))
        q = g.clone()
        alpha = torch.zeros_like(sy)
        for j in range(m - 1, -1, -1):
            a = rho[:, j] * (S[:, j] * q).sum(dim=1)
            alpha[:, j] = a
            q = q - a[:, None] * Y[:, j]
        last = torch.clamp(cnt - 1, min=0)
        ar = torch.arange(rows.shape[0], device=p.device)
        s_l, y_l = S[ar, last], Y[ar, last]
        yy = (y_l * y_l).sum(dim=1)
        h0 = torch.where(cnt > 0, (s_l * y_l).sum(dim=1) / torch.where(cnt > 0, yy, torch.ones_like(yy)),
                         torch.ones_like(yy))
        r = h0[:, None] * q
        for j in range(m):
            b = rho[:, j] * (Y[:, j] * r).sum(dim=1)
            r = r + S[:, j] * (alpha[:, j] - b)[:, None]
        step = lr * r
        if limit is not None:
            step = torch.clamp(step, -limit, limit)
        p.data.view(n, -1)[rows] = x - step
