"""Learned diagonal preconditioner for the mean-volume CG.

The CG operator of ``mean.correct_upsampling_padding`` is the Toeplitz
normal operator A = crop_L F^-1 diag(lhs) F pad_nf.  In the L^3 Fourier basis
its diagonal is exactly

    T_k = (lhs (*) |D|^2)(us*k) / nf^3,

with D the Fourier transform of the L-box on the nf grid; by the convolution
theorem this is one probe of the Toeplitz kernel ifft3c(lhs) against the
box autocorrelation (a separable triangle), i.e. two nf^3 FFTs
(``probe_diagonal``).  The Jacobi weights 1/T_k are the training targets.

``DiagonalMLP`` (in -> 64 -> 64 -> 64 -> 1, tanh-GELU, gfx950 forward and
backward kernels ``cov3d_mlp_forward`` / ``cov3d_mlp_backward``) predicts
log T_k from cheap features (``diagonal_features``):

    [log(sampling density x CTF^2 at us*k), log of its box means over
     3^3 / 7^3 / 15^3 us-grid voxels, log(shell mean of the same),
     log(noise variance), |k| / (L/2), 1 / (1 + |k|)]

so the preconditioner needs neither the probe nor extra operator
applications.  The default network (``FEATURE_SETS["shell"]``) reads the
per-shell inputs only -- the shell mean of density x CTF^2 (T_k depends on the
two only through their product), the noise variance and the radius -- so its
weights depend on |k| alone; the shell average of the probed log T_k
(``shell_average``) is the best shell-constant reference.  The "voxel" network
reads all eight.
On the padded mean systems it was measured on, the per-shell weights converge
in fewer iterations than the per-voxel ones, although the latter fit log T_k
more closely: the voxel-level variation of T_k inside a shell is not what
slows CG down.  ``train_default`` fits it on seeded acquisitions (Haar poses,
radial CTFs, L from 32 to 128) whose probed diagonals are the targets;
``save`` / ``load`` keep the weights in a plain ``.npz``.

The preconditioner M^-1 r = ifft3c(w * fft3c(r)) with w > 0 is linear and
Hermitian positive-definite, so PCG stays valid; plain CG is
the fallback (``mean.complex_conjugate_gradient``) and judges the final
residual.
"""

from __future__ import annotations

import math
from pathlib import Path
from typing import Callable, Optional, Sequence

import numpy as np
import torch

from . import _native
from .fsc_utils import average_fourier_shell, expand_fourier_shell
from .projection_funcs import centered_fft3, centered_ifft3

HID = 64


def n_params(n_in: int) -> int:
    return HID * n_in + HID + 2 * (HID * HID + HID) + HID + 1


def mlp_torch_reference(params: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """Differentiable torch restatement of the kernel's MLP (used as a test reference)."""
    n_in = x.shape[1]
    o = 0
    W1 = params[o : o + HID * n_in].reshape(HID, n_in); o += HID * n_in
    b1 = params[o : o + HID]; o += HID
    W2 = params[o : o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = params[o : o + HID]; o += HID
    W3 = params[o : o + HID * HID].reshape(HID, HID); o += HID * HID
    b3 = params[o : o + HID]; o += HID
    W4 = params[o : o + HID]; o += HID
    b4 = params[o]
    g = lambda a: torch.nn.functional.gelu(a, approximate="tanh")
    return g(g(g(x @ W1.T + b1) @ W2.T + b2) @ W3.T + b3) @ W4 + b4


class DiagonalMLP:
    """Parameters on the device; forward/backward run on the gfx950 kernels."""

    def __init__(self, n_in: int = 8, device="cuda", seed: int = 0, params: Optional[torch.Tensor] = None,
                 columns: Optional[Sequence[int]] = None):
        self.n_in = n_in
        # the columns of ``diagonal_features`` the network reads (all of them by default)
        self.columns = tuple(range(n_in)) if columns is None else tuple(int(c) for c in columns)
        assert len(self.columns) == n_in
        self.device = torch.device(device)
        if params is None:
            g = torch.Generator().manual_seed(seed)
            p = []
            for fan_in, fan_out in ((n_in, HID), (HID, HID), (HID, HID)):
                p.append(torch.randn((fan_out, fan_in), generator=g).reshape(-1) * math.sqrt(2.0 / fan_in))
                p.append(torch.zeros(fan_out))
            p.append(torch.randn(HID, generator=g) * math.sqrt(1.0 / HID))
            p.append(torch.zeros(1))
            params = torch.cat(p)
        self.params = params.to(device=self.device, dtype=torch.float32).contiguous()
        assert self.params.numel() == n_params(n_in)
        # feature normalisation (set by fit)
        self.x_mean = torch.zeros(n_in, device=self.device)
        self.x_std = torch.ones(n_in, device=self.device)
        self.y_mean = 0.0
        self.y_std = 1.0

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        _native.require_device(x, "features")
        x = x.to(torch.float32).contiguous()
        y = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        _native.check(_native.load().cov3d_mlp_forward(self.params.data_ptr(), x.data_ptr(), y.data_ptr(),
                                                        x.shape[0], self.n_in, _native.stream_ptr(x.device)))
        return y

    def backward(self, x: torch.Tensor, gy: torch.Tensor) -> torch.Tensor:
        x = x.to(torch.float32).contiguous()
        gy = gy.to(torch.float32).contiguous()
        grad = torch.zeros_like(self.params)
        _native.check(_native.load().cov3d_mlp_backward(self.params.data_ptr(), x.data_ptr(), gy.data_ptr(),
                                                         grad.data_ptr(), x.shape[0], self.n_in,
                                                         _native.stream_ptr(x.device)))
        return grad

    # --- normalised prediction of log T_k
    def predict_log_diag(self, feats: torch.Tensor) -> torch.Tensor:
        xn = (feats - self.x_mean) / self.x_std
        return self.forward(xn) * self.y_std + self.y_mean

    def fit(self, feats: torch.Tensor, targets: torch.Tensor, steps: int = 2000, lr: float = 2e-3,
            batch: int = 8192, seed: int = 0, log: Optional[list] = None) -> float:
        """Adam (cosine-decayed step) on the mean squared error of log T_k; gradients from the native backward."""
        self.x_mean = feats.mean(0)
        self.x_std = feats.std(0).clamp(min=1e-6)
        self.y_mean = float(targets.mean())
        self.y_std = float(targets.std().clamp(min=1e-6))
        xn = ((feats - self.x_mean) / self.x_std).contiguous()
        yn = ((targets - self.y_mean) / self.y_std).contiguous()
        m = torch.zeros_like(self.params)
        v = torch.zeros_like(self.params)
        b1, b2, eps = 0.9, 0.999, 1e-8
        g = torch.Generator(device=feats.device).manual_seed(seed)
        n = xn.shape[0]
        loss = float("nan")
        if log is not None:
            probe = torch.randint(0, n, (min(batch, n),), device=feats.device, generator=g)
            log.append(float((self.forward(xn[probe]) - yn[probe]).pow(2).mean()))
        for t in range(1, steps + 1):
            idx = torch.randint(0, n, (min(batch, n),), device=feats.device, generator=g)
            xb, yb = xn[idx], yn[idx]
            pred = self.forward(xb)
            err = pred - yb
            grad = self.backward(xb, 2.0 * err / xb.shape[0])
            m.mul_(b1).add_(grad, alpha=1 - b1)
            v.mul_(b2).addcmul_(grad, grad, value=1 - b2)
            mh = m / (1 - b1**t)
            vh = v / (1 - b2**t)
            lr_t = lr * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * t / steps)))
            self.params.sub_(lr_t * mh / (vh.sqrt() + eps))
            if log is not None and (t % 100 == 0 or t == steps):
                loss = float((err**2).mean())
                log.append(loss)
        return loss

    def state_dict(self) -> dict:
        return {"n_in": self.n_in, "params": self.params.cpu(), "x_mean": self.x_mean.cpu(),
                "x_std": self.x_std.cpu(), "y_mean": self.y_mean, "y_std": self.y_std, "columns": self.columns}

    @classmethod
    def from_state_dict(cls, sd: dict, device="cuda") -> "DiagonalMLP":
        m = cls(sd["n_in"], device=device, params=sd["params"], columns=sd.get("columns"))
        m.x_mean = sd["x_mean"].to(m.device)
        m.x_std = sd["x_std"].to(m.device)
        m.y_mean = float(sd["y_mean"])
        m.y_std = float(sd["y_std"])
        return m

    def make_preconditioner(self, lhs: torch.Tensor, L: int, noise_var: Optional[float]) -> Optional[Callable]:
        """Returns r -> ifft3c(w * fft3c(r)) with w = exp(-predicted log T), or None if w is invalid."""
        feats = diagonal_features(lhs, L, 1.0 if noise_var is None else noise_var)
        logd = self.predict_log_diag(feats[:, list(self.columns)])
        return preconditioner_from_log_diag(logd, L)


def preconditioner_from_log_diag(logd: torch.Tensor, L: int) -> Optional[Callable]:
    """r -> ifft3c(w * fft3c(r)) with w = exp(-logd) on the centred L^3 frequency grid.

    With S the half-length roll, fft3c = S F S, and S F^-1 S (w o S F S r) = F^-1 ((S w) o F r)
    (the shift theorem's phases (-1)^k cancel in pairs), so the operator runs as two plain FFTs
    and one multiply by the pre-rolled real weights.  On the GPU at L = 64, 128, 256 or 512 it
    runs as the Toeplitz operator with nf = L instead."""
    w = torch.exp(-logd).reshape(L, L, L)
    if not bool(torch.isfinite(w).all()) or bool((w <= 0).any()):
        return None
    if w.is_cuda and L in (64, 128, 256, 512):
        # the Toeplitz operator crop_L(ifft3c(lhs * fft3c(pad_nf(x)))) at nf = L (pad and crop are
        # the identity) is this operator with lhs = w: its LDS passes fuse the multiply into the x
        # transforms, where the plain form below runs three library kernels
        wc = w.to(torch.complex64).contiguous()
        lib, st = _native.load(), _native.stream_ptr(w.device)

        def apply_native(r: torch.Tensor) -> torch.Tensor:
            x = _native.dense(r.reshape(-1).to(torch.complex64))
            y = torch.empty_like(x)
            _native.check(lib.cov3d_toeplitz_apply(x.data_ptr(), wc.data_ptr(), y.data_ptr(), L, L, st))
            return y

        return apply_native
    ws = torch.fft.ifftshift(w).to(torch.float32).contiguous()

    def apply(r: torch.Tensor) -> torch.Tensor:
        return torch.fft.ifftn(ws * torch.fft.fftn(r.reshape(L, L, L).to(torch.complex64))).reshape(-1)

    return apply


def _box_autocorr(L: int, nf: int, device) -> torch.Tensor:
    """Separable triangle (L - |n|)_+ on the centred nf grid: autocorrelation of the L-box."""
    n = torch.arange(nf, device=device, dtype=torch.float64) - nf // 2
    t = torch.clamp(L - n.abs(), min=0)
    return t[:, None, None] * t[None, :, None] * t[None, None, :]


def probe_diagonal(lhs: torch.Tensor, L: int) -> torch.Tensor:
    """Exact diagonal T_k of the Toeplitz CG operator in the L^3 Fourier basis (flattened L^3, real)."""
    nf = lhs.shape[-1]
    us = nf // L
    kern = centered_ifft3(lhs.to(torch.complex128))  # Toeplitz kernel
    conv = centered_fft3(kern * _box_autocorr(L, nf, lhs.device))
    # frequency k on the L grid sits at index us*k + nf//2 of the centred nf grid
    k = torch.arange(L, device=lhs.device) - L // 2
    idx = us * k + nf // 2
    T = conv[idx][:, idx][:, :, idx].real / (L**3)
    return T.reshape(-1).to(torch.float32)


N_FEATURES = 8
SMOOTH_WIDTHS = (3, 7, 15)
# columns of ``diagonal_features``: "voxel" reads all eight; "shell" reads the per-shell ones only
# (shell mean of density x CTF^2, noise variance, |k| / (L/2), 1 / (1 + |k|))
FEATURE_SETS = {"voxel": tuple(range(N_FEATURES)), "shell": (4, 5, 6, 7)}


def shell_average(logd: torch.Tensor, L: int) -> torch.Tensor:
    """The shell average of a log diagonal (flattened L^3), expanded back to L^3: the
    least-squares best shell-constant prediction of log T_k."""
    return expand_fourier_shell(average_fourier_shell(logd.reshape(L, L, L)), L, 3).reshape(-1)


def _box_smooth(v: torch.Tensor, width: int) -> torch.Tensor:
    """Separable mean over a width^3 box on the us-grid (edge windows shrink)."""
    x = v.reshape(1, 1, *v.shape)
    for k in ((width, 1, 1), (1, width, 1), (1, 1, width)):
        pad = tuple(q // 2 for q in k)
        x = torch.nn.functional.avg_pool3d(x, k, stride=1, padding=pad, count_include_pad=False)
    return x.reshape(v.shape)


def diagonal_features(lhs: torch.Tensor, L: int, noise_var: float) -> torch.Tensor:
    """(L^3, 8) features of each L-grid frequency k, read off the kernel lhs at us*k.

    [log density*CTF^2, log of its box means over 3^3, 7^3 and 15^3 us-grid
     voxels, log shell mean, log noise variance, |k| / (L/2), 1 / (1 + |k|)].
    The diagonal is lhs averaged against the box's |Dirichlet|^2 kernel, whose
    core spans a few us-grid voxels and whose sidelobes reach further (they
    carry all of the diagonal in the unsampled corners), so the multi-scale
    means and the absolute radius carry what changes with L.
    """
    nf = lhs.shape[-1]
    us = nf // L
    k = torch.arange(L, device=lhs.device) - L // 2
    idx = us * k + nf // 2
    a = lhs.abs().float()
    at = lambda v: v[idx][:, idx][:, :, idx].clamp(min=1e-12)
    dens = at(a)
    smooth = [at(_box_smooth(a, w)).log() for w in SMOOTH_WIDTHS]
    shell = expand_fourier_shell(average_fourier_shell(dens), L, 3).clamp(min=1e-12)
    ax = torch.ceil(torch.arange(-L / 2, L / 2, device=lhs.device, dtype=torch.float32))[:L]
    zz, yy, xx = torch.meshgrid(ax, ax, ax, indexing="ij")
    kabs = torch.sqrt(zz**2 + yy**2 + xx**2)
    nv = torch.full_like(dens, math.log(max(noise_var, 1e-30)))
    return torch.stack([dens.log(), *smooth, shell.log(), nv, kabs / (L / 2), 1.0 / (1.0 + kabs)],
                       dim=-1).reshape(-1, N_FEATURES)


# ---------------------------------------------------------------------------
# Training data: Toeplitz kernels of seeded acquisition geometries
# ---------------------------------------------------------------------------


def acquisition_lhs(L: int, n_images: int, seed: int, device, defocus_um=(0.5, 3.0), pixel_size: float = 1.5,
                    upsampling_factor: int = 2, batch: int = 4096) -> torch.Tensor:
    """Regularised CG kernel lhs of a random acquisition (Haar poses, radial CTFs), (us*L)^3 complex.

    lhs depends only on poses and CTFs (it is the back-projection of CTF^2,
    mean.py:206-216), so no images are needed.
    """
    from .dataset import ctf_from_params, haar_rotations, pts_rot_from_rotations
    from .mean import regularize_lhs
    from .nufft_plan import NufftPlanDiscretized
    from .projection_funcs import im_backward

    gen = torch.Generator().manual_seed(seed)
    dev = torch.device(device)
    plan = NufftPlanDiscretized((L,) * 3, upsample_factor=upsampling_factor, mode="nearest")
    vL = L * upsampling_factor
    lhs = torch.zeros((vL,) * 3, dtype=torch.complex64, device=dev)
    for i in range(0, n_images, batch):
        n = min(batch, n_images - i)
        rots = haar_rotations(n, gen).to(dev)
        du = torch.empty(n, dtype=torch.float64).uniform_(defocus_um[0], defocus_um[1], generator=gen) * 1e4
        prm = torch.stack([du, du, torch.zeros(n, dtype=torch.float64), torch.full((n,), 300.0, dtype=torch.float64),
                           torch.full((n,), 2.7, dtype=torch.float64), torch.full((n,), 0.1, dtype=torch.float64),
                           torch.full((n,), pixel_size, dtype=torch.float64)], 1).to(torch.float32).to(dev)
        f = ctf_from_params(prm, L)
        plan.setpts(pts_rot_from_rotations(rots, L))
        lhs += im_backward(torch.complex(f, torch.zeros_like(f)), plan, f, fourier_domain=True)[0]
    lhs /= L
    return regularize_lhs(lhs)


def training_set(n_datasets: int, device, L_choices=(32, 48, 64, 96, 128), rows_per_dataset: int = 4096,
                 seed0: int = 5000, progress: Optional[Callable[[str], None]] = None):
    """Features and log-diagonal targets of ``n_datasets`` seeded acquisitions.

    Dataset k draws L from ``L_choices``, N = L^2 * loguniform(0.5, 20) images,
    pixel size U[1, 3] A, defocus range inside [0.4, 4] um and a noise variance
    10^U[-2, 2]; the targets are the probed diagonals.  Rows: 3/4 uniform over
    the L^3 frequencies, 1/4 from the ball |k| <= 8 where the kernel bends most.
    """
    rng = np.random.default_rng(seed0)
    X, Y = [], []
    for k in range(n_datasets):
        L = int(rng.choice(L_choices))
        n_img = int(L * L * np.exp(rng.uniform(np.log(0.5), np.log(20.0))))
        lo = rng.uniform(0.4, 1.5)
        hi = lo + rng.uniform(0.5, 2.5)
        px = rng.uniform(1.0, 3.0)
        noise_var = float(10 ** rng.uniform(-2, 2))
        lhs = acquisition_lhs(L, n_img, dataset_seed(seed0, k), device, (lo, hi), px)
        feats = diagonal_features(lhs, L, noise_var)
        target = torch.log(probe_diagonal(lhs, L).clamp(min=1e-30))
        g = torch.Generator().manual_seed(seed0 + k)
        n_u = 3 * rows_per_dataset // 4
        central = torch.nonzero(feats[:, 6].cpu() * (L / 2) <= 8).reshape(-1)
        pick = central[torch.randint(0, central.numel(), (rows_per_dataset - n_u,), generator=g)]
        idx = torch.cat([torch.randint(0, L**3, (n_u,), generator=g), pick]).to(feats.device)
        X.append(feats[idx])
        Y.append(target[idx])
        if progress is not None and (k + 1) % 20 == 0:
            progress(f"{k + 1}/{n_datasets} acquisitions probed")
    return torch.cat(X), torch.cat(Y)


def dataset_seed(seed0: int, k: int) -> int:
    """Pose/CTF seed of training acquisition k (held-out checks use seeds outside this set)."""
    return seed0 + 7 * k + 1


def train_default(device, n_datasets: int = 200, steps: int = 4000, seed: int = 0,
                  log: Optional[list] = None, features: str = "shell") -> DiagonalMLP:
    """Trains the preconditioner MLP (features -> log T_k) on ``n_datasets`` acquisitions."""
    cols = FEATURE_SETS[features]
    X, Y = training_set(n_datasets, device)
    model = DiagonalMLP(len(cols), device=device, seed=seed, columns=cols)
    X = X[:, list(cols)].contiguous()
    model.fit(X, Y, steps=steps, log=log)
    return model


def save(model: DiagonalMLP, path) -> None:
    sd = model.state_dict()
    np.savez(path, params=sd["params"].numpy(), x_mean=sd["x_mean"].numpy(), x_std=sd["x_std"].numpy(),
             y_mean=np.float64(sd["y_mean"]), y_std=np.float64(sd["y_std"]), n_in=np.int64(sd["n_in"]),
             columns=np.asarray(sd["columns"], dtype=np.int64))


def load(path, device) -> DiagonalMLP:
    z = np.load(path, allow_pickle=False)
    sd = {"n_in": int(z["n_in"]), "params": torch.tensor(z["params"]), "x_mean": torch.tensor(z["x_mean"]),
          "x_std": torch.tensor(z["x_std"]), "y_mean": float(z["y_mean"]), "y_std": float(z["y_std"]),
          "columns": tuple(int(c) for c in z["columns"]) if "columns" in z.files else None}
    return DiagonalMLP.from_state_dict(sd, device=device)


WEIGHTS = {"shell": Path(__file__).with_name("precond_mlp.npz"),
           "voxel": Path(__file__).with_name("precond_mlp_voxel.npz")}


def load_default(device, features: str = "shell") -> DiagonalMLP:
    """A committed network, trained by ``python -m 3dcovarestimation_amd.precond --train OUT
    --features {shell,voxel}``: the per-shell one (default) or the per-voxel one."""
    path = WEIGHTS[features]
    if not path.exists():
        raise FileNotFoundError(f"no trained preconditioner at {path}")
    return load(path, device)


def cg_comparison(reg_lhs: torch.Tensor, rhs: torch.Tensor, L: int, model: Optional[DiagonalMLP],
                  noise_var: Optional[float], rtol: float = 1e-6, maxiter: int = 3000,
                  with_probe: bool = True, repeats: int = 1, extra_models: Optional[dict] = None,
                  with_shell_probe: bool = False) -> dict:
    """Plain CG vs learned-diagonal PCG (vs probed-diagonal PCG) on one padded mean system.

    Each solve starts from the direct estimate crop(ifft3c(rhs / lhs)), as
    ``reconstruct_mean`` does.  Seconds are wall clock between device
    synchronisations and include building the preconditioner (features and
    MLP, or the probe).  The best of ``repeats`` timed solves is kept.  ``extra_models``
    ({name: network}) adds a solve "learned_<name>" per network; ``with_shell_probe`` adds the
    shell average of the probed diagonal (the best shell-constant preconditioner).
    """
    import time

    from .mean import correct_upsampling_padding

    x0 = centered_ifft3(rhs / reg_lhs, cropping_size=(L,) * 3).real
    us = reg_lhs.shape[-1] // L

    def solve(make):
        best = None
        for _ in range(max(1, repeats)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pre = make()
            vol, info = correct_upsampling_padding(reg_lhs, rhs, x0, L, us, precond=pre, rtol=rtol,
                                                   return_info=True, maxiter=maxiter)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if best is None or dt < best[0]:
                best = (dt, info, vol, pre is not None)
        dt, info, vol, used = best
        ms_it = 1e3 * dt / max(info.iterations_run, 1)
        reached = {f"{t:.0e}": next((i for i, rr in enumerate(info.residuals) if rr < t), None)
                   for t in (1e-3, 1e-4, 1e-5, 1e-6)}
        return {"iterations": info.iterations, "iterations_run": info.iterations_run, "converged": min(info.residuals) < rtol, "seconds": round(dt, 4),
                "ms_per_iteration": round(ms_it, 4), "final_residual": info.final_residual,
                "iterations_to": reached, "preconditioned": used, "fallback": info.fallback}, vol

    out = {}
    out["plain"], v_plain = solve(lambda: None)
    if model is not None:
        out["learned"], v_learn = solve(lambda: model.make_preconditioner(reg_lhs, L, noise_var))
        out["learned"]["rel_diff_vs_plain"] = float((v_learn - v_plain).norm() / v_plain.norm())
    for name, m in (extra_models or {}).items():
        out[f"learned_{name}"], _ = solve(lambda: m.make_preconditioner(reg_lhs, L, noise_var))
    if with_probe:
        out["probe"], _ = solve(lambda: preconditioner_from_log_diag(
            torch.log(probe_diagonal(reg_lhs, L).clamp(min=1e-30)), L))
    if with_shell_probe:
        out["probe_shell"], _ = solve(lambda: preconditioner_from_log_diag(
            shell_average(torch.log(probe_diagonal(reg_lhs, L).clamp(min=1e-30)), L), L))
    return out


def _main(argv=None) -> int:
    import argparse
    import time

    ap = argparse.ArgumentParser(description="train the diagonal preconditioner network on the GPU")
    ap.add_argument("--train", metavar="OUT.npz", help="train and write the weights")
    ap.add_argument("--datasets", type=int, default=200)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--features", choices=sorted(FEATURE_SETS), default="shell",
                    help="shell: the per-shell inputs only; voxel: all eight")
    args = ap.parse_args(argv)
    if not args.train:
        ap.print_help()
        return 1
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    X, Y = training_set(args.datasets, dev, progress=lambda m: print(m, flush=True))
    torch.cuda.synchronize()
    print(f"training set: {X.shape[0]} rows from {args.datasets} acquisitions in {time.perf_counter() - t0:.1f} s",
          flush=True)
    cols = FEATURE_SETS[args.features]
    X = X[:, list(cols)].contiguous()
    model = DiagonalMLP(len(cols), device=dev, seed=0, columns=cols)
    log: list = []
    t0 = time.perf_counter()
    model.fit(X, Y, steps=args.steps, log=log)
    torch.cuda.synchronize()
    print(f"fit: {args.steps} Adam steps in {time.perf_counter() - t0:.1f} s; normalised MSE "
          f"first {log[0]:.4f} last {log[-1]:.5f}", flush=True)
    resid = (model.predict_log_diag(X) - Y)
    print(f"log-diagonal residual: rms {float(resid.pow(2).mean().sqrt()):.4f}, max {float(resid.abs().max()):.4f}",
          flush=True)
    save(model, args.train)
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
