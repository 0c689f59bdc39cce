"""Fused maximum-likelihood contractions and their hand-written adjoint.

``ml_forward`` / ``ml_grad_images`` wrap the gfx950 kernels
``cov3d_ml_forward`` / ``cov3d_ml_grad_images``.  ``MLCostNufft`` is the
chosen form of the spatial ML objective of cov3d/covar_sgd.py:1238-1286 on the
NUFFT path: its forward runs the type-2 NUFFT and the per-image Gram + FP64
Cholesky; its backward forms the CTF-weighted residual-times-coefficient
gradient images and back-projects them with one type-1 NUFFT, replacing
autograd through vol_forward, matmul, solve and logdet.

Equivalence with the reference's spatial-domain cost: with
X_t = mask*CTF*c_t/L (the Fourier-domain projection) the spatial projection
is Re(ifft2c(X_t)); by Parseval its inner products equal the real inner
products of Q_t = X_t / L and y~ = fft2c(y) / L on the Hermitian part of X_t.
The anti-Hermitian part of X_t is the NUFFT's own O(eps) asymmetry for a real
volume; it enters the Gram only at O(eps^2) and is removed from the
gradient by the final real part.
"""

from __future__ import annotations

from typing import Optional, Tuple

import torch

from . import _native

PRECISION_FP32 = 0
PRECISION_BF16X3 = 1

_HALF_CACHE = {}


def half_plane_pixels(L: int, device) -> torch.Tensor:
    """Row-major indices (row * L + col) of one pixel of every +-k pair, plus the centre.

    Even L: rows/columns 0 are the zeroed Nyquist lines, pixel (i, j) pairs with
    (L - i, L - j); kept: rows > L/2 and row L/2 from column L/2 on.
    Odd L: (i, j) pairs with (L-1-i, L-1-j); kept: rows > c and row c from c on.
    """
    key = (L, str(device))
    if key not in _HALF_CACHE:
        c = L // 2
        rows = []
        lo = 1 if L % 2 == 0 else 0
        for i in range(c, L):
            j0 = c if i == c else lo
            rows.append(torch.arange(j0, L) + i * L)
        _HALF_CACHE[key] = torch.cat(rows).to(torch.int32).to(device)
    return _HALF_CACHE[key]


def ml_forward(
    c: torch.Tensor,
    yhat: torch.Tensor,
    L: int,
    noise_var: float,
    ctf: Optional[torch.Tensor] = None,
    lam_inv: Optional[torch.Tensor] = None,
    mean_const: bool = False,
    precision: int = PRECISION_BF16X3,
    want_m: bool = False,
    want_minv: bool = True,
    want_b: bool = False,
    pix_map: Optional[torch.Tensor] = None,
):
    """Per-image ML terms from raw type-2 values.

    c: (r, B, L*L) complex64 raw NUFFT values (or (r, B, npix) on the pixels of
    ``pix_map``); yhat: (B, L, L) complex64 centred FFT of the images; ctf:
    (B, L, L) float32 or None.
    Returns dict with cost (B,), z (B, r) and optionally m (B, r, r),
    minv (B, r, r), b (B, r).
    """
    r, B = c.shape[0], c.shape[1]
    dev = c.device
    _native.require_device(c, "c")
    c = _native.dense(c)
    yhat = _native.dense(yhat.to(torch.complex64))
    ctf_t = ctf.to(torch.float32).contiguous() if ctf is not None else None
    lam_t = lam_inv.to(device=dev, dtype=torch.float32).contiguous() if lam_inv is not None else None
    out = {
        "cost": torch.empty(B, dtype=torch.float32, device=dev),
        "z": torch.empty((B, r), dtype=torch.float32, device=dev),
    }
    if want_m:
        out["m"] = torch.empty((B, r, r), dtype=torch.float32, device=dev)
    if want_minv:
        out["minv"] = torch.empty((B, r, r), dtype=torch.float32, device=dev)
    if want_b:
        out["b"] = torch.empty((B, r), dtype=torch.float32, device=dev)
    lib = _native.load()
    _native.check(
        lib.cov3d_ml_forward(
            c.data_ptr(), _native.ptr(ctf_t), yhat.data_ptr(), _native.ptr(lam_t), _native.ptr(pix_map),
            0 if pix_map is None else pix_map.numel(), B, r, L,
            float(noise_var), 1 if mean_const else 0, int(precision),
            out["cost"].data_ptr(), out["z"].data_ptr(), _native.ptr(out.get("m")),
            _native.ptr(out.get("minv")), _native.ptr(out.get("b")), _native.stream_ptr(dev),
        )
    )
    return out


def ml_grad_images(
    c: torch.Tensor,
    yhat: torch.Tensor,
    z: torch.Tensor,
    minv: torch.Tensor,
    L: int,
    noise_var: float,
    grad_scale: float,
    ctf: Optional[torch.Tensor] = None,
    out: Optional[torch.Tensor] = None,
    pix_map: Optional[torch.Tensor] = None,
) -> torch.Tensor:
    """Gradient of the summed ML cost with respect to the raw type-2 values, times grad_scale."""
    r, B = c.shape[0], c.shape[1]
    g = out if out is not None else torch.empty_like(c)
    ctf_t = ctf.to(torch.float32).contiguous() if ctf is not None else None
    lib = _native.load()
    # the converted inputs are bound to names: a temporary freed inside the argument list could
    # hand its memory to the next one before the call
    yhat, z, minv = _native.dense(yhat), _native.dense(z), _native.dense(minv)
    _native.check(
        lib.cov3d_ml_grad_images(
            c.data_ptr(), _native.ptr(ctf_t), yhat.data_ptr(), z.data_ptr(),
            minv.data_ptr(), _native.ptr(pix_map), 0 if pix_map is None else pix_map.numel(),
            g.data_ptr(), B, r, L, float(noise_var), float(grad_scale),
            _native.stream_ptr(c.device),
        )
    )
    return g


def ml_grad_sorted(
    c: torch.Tensor,
    yhat: torch.Tensor,
    z: torch.Tensor,
    minv: torch.Tensor,
    L: int,
    noise_var: float,
    grad_scale: float,
    values: int,
    spos: int,
    ldc: int,
    ctf: Optional[torch.Tensor] = None,
    pix_map: Optional[torch.Tensor] = None,
) -> None:
    """ml_grad_images written into a plan's sorted type-1 input (``NufftPlan.sorted_input``)."""
    r, B = c.shape[0], c.shape[1]
    ctf_t = ctf.to(torch.float32).contiguous() if ctf is not None else None
    lib = _native.load()
    yhat, z, minv = _native.dense(yhat), _native.dense(z), _native.dense(minv)
    _native.check(
        lib.cov3d_ml_grad_sorted(
            c.data_ptr(), _native.ptr(ctf_t), yhat.data_ptr(), z.data_ptr(),
            minv.data_ptr(), _native.ptr(pix_map), 0 if pix_map is None else pix_map.numel(),
            values, spos, ldc, B, r, L, float(noise_var), float(grad_scale), _native.stream_ptr(c.device),
        )
    )


def _half_plan(plan):
    """A companion plan of ``plan`` for the half-plane points (created once, kept on the plan)."""
    hp = getattr(plan, "_half_plan", None)
    if hp is None:
        from .nufft_plan import NufftPlan

        hp = NufftPlan(plan.sz, batch_size=plan.batch_size, eps=plan.eps, dtype=plan.dtype, device=plan.device)
        plan._half_plan = hp
    return hp


class MLCostNufft(torch.autograd.Function):
    """Batch-mean spatial ML cost of rank-r real eigenvolumes on a NufftPlan.

    forward(vols (r, L, L, L) f32, yhat (B, L, L) c64, plan, ctf (B, L, L) or
    None, noise_var, mean_const, precision, half) -> scalar cost
    = 0.5 * mean_i(-b_i.z_i + logdet M_i) [+ |y_i|^2/noise_var].
    With ``half`` the NUFFTs run on one pixel of every +-k pair (the
    projections of real volumes are Hermitian), through a companion plan
    whose points are taken from ``plan.points``.  The plan's points must stay
    set until backward has run.
    """

    @staticmethod
    def forward(ctx, vols, yhat, plan, ctf, noise_var, mean_const, precision, half=True):
        r = vols.shape[0]
        L = vols.shape[-1]
        B = yhat.shape[0]
        if half:
            pix = half_plane_pixels(L, vols.device)
            hp = _half_plan(plan)
            pts = plan.points.reshape(3, B, L * L)[:, :, pix.long()].reshape(3, -1)
            hp.set_points_flat(pts, B)
            run_plan = hp
            c = hp.execute_forward_flat(vols.detach())[:r].reshape(r, B, pix.numel())
        else:
            pix = None
            run_plan = plan
            c = plan.execute_forward(vols.detach())[:r].reshape(r, B, L * L)
        res = ml_forward(c, yhat, L, noise_var, ctf=ctf, mean_const=False, precision=precision, pix_map=pix)
        cost = res["cost"].mean()
        if mean_const:
            cost = cost + 0.5 * (yhat.abs() ** 2).sum(dim=(1, 2)).mean() / (L * L) / noise_var
        ctx.plan = run_plan
        ctx.pix = pix
        ctx.meta = (L, B, float(noise_var))
        ctx.save_for_backward(c, yhat, res["z"], res["minv"], ctf if ctf is not None else torch.empty(0))
        ctx.has_ctf = ctf is not None
        return cost

    @staticmethod
    def backward(ctx, grad_out):
        c, yhat, z, minv, ctf = ctx.saved_tensors
        L, B, noise_var = ctx.meta
        r = c.shape[0]
        plan = ctx.plan
        ctf_arg = ctf if ctx.has_ctf else None
        vals, ldc, spos = plan.sorted_input(r) if plan.dtype == torch.float32 else (None, 0, None)
        if vals is not None and ldc % 2 == 0 and r <= ldc <= 24:
            # gradient rows go straight to where the type-1 spreading reads them
            ml_grad_sorted(c, yhat, z, minv, L, noise_var, 1.0 / B, vals, spos, ldc, ctf=ctf_arg,
                           pix_map=ctx.pix)
            grad_v = plan.execute_adjoint_presorted(r, real_output=True)
        else:
            g = ml_grad_images(c, yhat, z, minv, L, noise_var, 1.0 / B, ctf=ctf_arg, out=torch.empty_like(c),
                               pix_map=ctx.pix)
            grad_v = plan.execute_adjoint(g.reshape(r, -1), real_output=True)
        grad_v = grad_v[:r] * grad_out
        return grad_v, None, None, None, None, None, None, None


def _boundary_order(plan, B: int, L: int, keep: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """The pixels whose discretised slice is not the mirror image of its -k partner, per image in
    ascending order (``cov3d_slice_boundary_pixels``): (order (B, L*L) int32, valid in the first
    count[b] entries; count (B,) int32).

    The slice of pixel k reads the centred volume at index x = p vL / (2 pi) + vL / 2 per axis,
    and its -k partner at vL - x; cell j mirrors to cell vL - j, so everything is mirrored except
    the Nyquist cell 0, whose mirror vL lies outside the volume.  A pixel is flagged when its
    stencil can reach cell 0 or cell vL on any axis (x within one cell of 0 or of vL; a superset of
    both interpolation modes), and the flag is OR-ed with its partner's, so the set is closed under
    k -> -k; only pixels with ``keep`` (uint8, L*L) set are listed.
    """
    vL = plan.sz[0] * plan.upsample_factor
    grid = _native.dense(plan._grid_flat.detach().to(torch.float32))
    dev = grid.device
    order = torch.empty((B, L * L), dtype=torch.int32, device=dev)
    count = torch.empty(B, dtype=torch.int32, device=dev)
    flags = torch.empty(B * L * L, dtype=torch.uint8, device=dev)
    keep = _native.dense(keep.to(device=dev, dtype=torch.uint8))
    _native.check(_native.load().cov3d_slice_boundary_pixels(
        grid.data_ptr(), keep.data_ptr(), B, L, vL, flags.data_ptr(), order.data_ptr(), count.data_ptr(), L * L,
        _native.stream_ptr(dev)))
    return order, count


def herm_solve(M: torch.Tensor, b: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """M^-1, M^-1 b and log det M for a batch of Hermitian positive definite complex128 matrices M
    (n, r, r), r <= 32, and right-hand sides b (n, r), from one FP64 Cholesky per matrix
    (``cov3d_herm_solve_c128``; the lower triangle of M is read)."""
    n, r = M.shape[0], M.shape[-1]
    Md = _native.dense(M.to(torch.complex128))
    bd = _native.dense(b.to(torch.complex128))
    minv = torch.empty_like(Md)
    z = torch.empty((n, r), dtype=torch.complex128, device=M.device)
    logdet = torch.empty(n, dtype=torch.float64, device=M.device)
    _native.check(_native.load().cov3d_herm_solve_c128(Md.data_ptr(), bd.data_ptr(), minv.data_ptr(), z.data_ptr(),
                                                        logdet.data_ptr(), n, r, _native.stream_ptr(M.device)))
    return minv, z, logdet


def _first_entries(order: torch.Tensor, count: torch.Tensor, n: int) -> Tuple[torch.Tensor, torch.Tensor]:
    """The first n entries of every row of a compacted list as int64 indices (0 past a row's
    count) and the mask of the valid ones."""
    valid = torch.arange(n, device=order.device).unsqueeze(0) < count.unsqueeze(1)
    idx = torch.where(valid, order[:, :n].long(), torch.zeros((), dtype=torch.long, device=order.device))
    return idx, valid


class MLCostDiscretized(torch.autograd.Function):
    """Batch-mean Fourier-domain ML cost on the discretised projection (cov3d/covar_sgd.py:1350-1399
    with nufft_disc "bilinear" / "nearest", the reference's default trainer path).

    forward(vhat (r, vL, vL, vL) complex64, yhat (B, L, L) complex64 Fourier-domain images, plan
    (NufftPlanDiscretized with its points set), ctf (B, L, L) or None, noise_var (the Fourier-domain
    noise variance, L^2 times the spatial one), mean_const, precision, half) -> scalar cost.

    With P = mask * CTF * S(vhat) / L (the reference's Fourier-domain projection) the objective is
    0.5 mean_i(-Re(b_i^H z_i) + logdet M_i), M = I + conj(P) P^T / noise_var, b = conj(P) y / noise_var,
    z = M^-1 b (complex Hermitian M).  The slices come from the native gather (cov3d_slice_forward);
    the fused kernel (split-BF16 MFMA Gram, raw values c = S(vhat), images yhat, spatial noise
    variance noise_var / L^2) gives Re M and Re b over all L^2 pixels.  Their imaginary parts cancel
    between every pixel and its -k partner except where the slice touches the padded volume's
    Nyquist cell (``_boundary_order``), so they are summed over those pixels only.  The r x r
    complex solve and log-determinant run in FP64.  Backward: the gradient with respect to the raw
    values, conj(M^-1) Q + conj(z) (z^T Q - y~) in the kernel's units, splits into two passes of the
    fused gradient kernel with real coefficients (conj(M^-1) = R - iJ, z = a + ib):
        real part:      (R + b b^T) Q + a (a^T Q - y~)
        imaginary part: (-J + a b^T - b a^T - b b^T) Q + (-b) ((-b)^T Q - y~)
    both computed in one pass by ``cov3d_ml_grad_images_complex`` (g = g1 + i g2 written once), and the
    native exact adjoint of the slice (the slice plan's gather) scatters it, replacing autograd through
    grid_sample, solve and logdet.

    With ``half`` (the default) the slices are taken on one pixel of every +-k pair only (the
    ``half_plane_pixels`` the NUFFT path uses, weight 2, the centre 1): for a volume that is the
    transform of a real one the slice at -k is the conjugate of the slice at k, so the pair sums
    conj(Q) Q^T + Q conj(Q)^T and conj(Q) y + Q conj(y) are twice the real parts, which the fused
    kernel forms.  Only the boundary pairs break the mirror (the -k member reads the Nyquist cell):
    for each of them the other member is sliced as well (a second, small point set) and the cost
    takes the exact terms conj(Q_o) Q_o^T - Q_h conj(Q_h)^T (and likewise for b) as a correction.
    Backward: a mirrored pair contributes g(k) + conj(g(-k)) = 2 g1(k) once the volume gradient
    reaches the real eigenvolumes (the imaginary pass cancels within the pair), so the half plane
    runs the real-coefficient pass only, with the pair weight; a boundary pair takes both members'
    exact complex gradients with weight 1 (the h value replaced, the o member scattered through the
    second point set).  The resulting volume gradient differs from the full-plane one only by a
    part that is anti-Hermitian in the volume, which vanishes through the Fourier transform of the
    real eigenvolumes.
    """

    @staticmethod
    def forward(ctx, vhat, yhat, plan, ctf, noise_var, mean_const, precision, half=True):
        r = vhat.shape[0]
        L = yhat.shape[-1]
        B = yhat.shape[0]
        dev = vhat.device
        nv_s = float(noise_var) / (L * L)
        from .projection_funcs import nyquist_mask

        scale = nyquist_mask(L, dev).reshape(1, -1).expand(B, -1)
        if ctf is not None:
            scale = scale * ctf.reshape(B, -1).to(torch.float32)
        ctx.half = bool(half)
        if half:
            pix = half_plane_pixels(L, dev)
            npix = pix.numel()
            gfull = plan._grid_flat.detach().reshape(B, L * L, 3)
            # boundary pairs: h in the half plane, o = its mirror pixel
            in_h = torch.zeros(L * L, dtype=torch.uint8, device=dev)
            in_h[pix.long()] = 1
            border, bcount = _boundary_order(plan, B, L, in_h)
            nb = int(bcount.max())
            gh = gfull[:, pix.long()].reshape(-1, 3)
            if nb > 0:
                order, valid = _first_entries(border, bcount, nb)  # (B, nb)
                oidx = _mirror_index(L, dev)[order]  # mirror pixels (full-plane indices)
                go = torch.gather(gfull, 1, oidx.unsqueeze(-1).expand(-1, -1, 3)).clone()
                go[~valid] = 4.0  # outside the volume: padded entries slice to zero
                gh = torch.cat([gh, go.reshape(-1, 3)])
            # one plan over the half-plane points followed by the mirror points of the boundary
            # pairs: one sort, one forward pass, and one adjoint that sums both sets
            hp = _disc_companion(plan, "_disc_hb")
            hp.set_grid_flat(gh)
            c, co = hp.execute_forward_split(vhat.detach(), B * npix)
            c = c.reshape(r, B, npix)
            res = ml_forward(c, yhat, L, nv_s, ctf=ctf, mean_const=False, precision=precision, want_m=True,
                             want_minv=False, want_b=True, pix_map=pix)
            M = res["m"].double()
            bvec = res["b"].double()
            Mc = torch.complex(M, torch.zeros_like(M))
            bc = torch.complex(bvec, torch.zeros_like(bvec))
            bnd = None
            if nb > 0:
                hpos = torch.full((L * L,), -1, dtype=torch.long, device=dev)
                hpos[pix.long()] = torch.arange(npix, device=dev)
                hidx = hpos[order].clamp(min=0)  # positions in the half set
                co = co.reshape(r, B, nb)
                vf = valid.to(torch.float32)
                wh = torch.gather(scale, 1, order) * vf / float(L * L)
                wo = torch.gather(scale, 1, oidx) * vf / float(L * L)
                qh = (torch.gather(c, 2, hidx.unsqueeze(0).expand(r, -1, -1)) * wh).permute(1, 0, 2).to(torch.complex128)
                qo = (co * wo).permute(1, 0, 2).to(torch.complex128)  # (B, r, nb)
                yf = yhat.reshape(B, -1)
                yh = (torch.gather(yf, 1, order) * (vf / L)).to(torch.complex128)
                yo = (torch.gather(yf, 1, oidx) * (vf / L)).to(torch.complex128)
                Mc = Mc + ((qo.conj() @ qo.transpose(1, 2)) - (qh @ qh.conj().transpose(1, 2))) / nv_s
                bc = bc + ((qo.conj() @ yo.unsqueeze(-1)).squeeze(-1) - (qh @ yh.conj().unsqueeze(-1)).squeeze(-1)) / nv_s
                bnd = (hidx, valid, co, wh, wo, order, oidx)
        else:
            pix = None
            c = plan.execute_forward(vhat.detach()).reshape(r, B, L * L)
            res = ml_forward(c, yhat, L, nv_s, ctf=ctf, mean_const=False, precision=precision, want_m=True,
                             want_minv=False, want_b=True)
            M = res["m"].double()
            bvec = res["b"].double()
            # imaginary parts from the pixels whose -k partner is not their mirror image
            border, bcount = _boundary_order(plan, B, L, torch.ones(L * L, dtype=torch.uint8, device=dev))
            nbd = int(bcount.max())
            Mi = torch.zeros_like(M)
            bi = torch.zeros_like(bvec)
            if nbd > 0:
                order, valid = _first_entries(border, bcount, nbd)  # (B, nbd)
                w = torch.gather(scale, 1, order) * valid / float(L * L)  # Q = mask ctf c / L^2
                q = torch.gather(c, 2, order.unsqueeze(0).expand(r, -1, -1)) * w.unsqueeze(0)  # (r, B, nbd)
                q = q.permute(1, 0, 2).to(torch.complex128)  # (B, r, nbd)
                yt = torch.gather(yhat.reshape(B, -1), 1, order).to(torch.complex128) * (valid / float(L))
                Mi = (q.conj() @ q.transpose(1, 2)).imag / nv_s
                bi = (q.conj() @ yt.unsqueeze(-1)).squeeze(-1).imag / nv_s
            Mc = torch.complex(M, Mi)
            bc = torch.complex(bvec, bi)
            bnd = None
        Minv, z, logdet = herm_solve(Mc, bc)
        cost_i = 0.5 * (-(bc.conj() * z).sum(dim=1).real + logdet)
        cost = cost_i.mean().to(torch.float32)
        if mean_const:
            cost = cost + 0.5 * (yhat.abs() ** 2).sum(dim=(1, 2)).mean() / float(noise_var)
        ctx.plan = plan
        ctx.pix = pix
        ctx.bnd = bnd
        ctx.meta = (L, B, nv_s, tuple(vhat.shape))
        ctx.save_for_backward(c, yhat, z, Minv, ctf if ctf is not None else torch.empty(0))
        ctx.has_ctf = ctf is not None
        return cost

    @staticmethod
    def backward(ctx, grad_out):
        c, yhat, z, Minv, ctf = ctx.saved_tensors
        L, B, nv_s, shape = ctx.meta
        r = c.shape[0]
        ctf_arg = ctf if ctx.has_ctf else None
        Rm, Jm = Minv.real, Minv.imag
        a, bz = z.real, z.imag
        A1 = (Rm + bz.unsqueeze(2) * bz.unsqueeze(1)).to(torch.float32)
        A2 = (-Jm + a.unsqueeze(2) * bz.unsqueeze(1) - bz.unsqueeze(2) * a.unsqueeze(1)
              - bz.unsqueeze(2) * bz.unsqueeze(1)).to(torch.float32)
        g = torch.empty_like(c)
        ctf_t = ctf_arg.to(torch.float32).contiguous() if ctf_arg is not None else None
        pix = ctx.pix
        if not ctx.half:
            # bound to names, so that no converted input is freed before the call
            yh = _native.dense(yhat)
            z1, m1 = _native.dense(a.to(torch.float32)), _native.dense(A1)
            z2, m2 = _native.dense((-bz).to(torch.float32)), _native.dense(A2)
            _native.check(_native.load().cov3d_ml_grad_images_complex(
                c.data_ptr(), _native.ptr(ctf_t), yh.data_ptr(), z1.data_ptr(), m1.data_ptr(), z2.data_ptr(),
                m2.data_ptr(), None, 0, g.data_ptr(), B, r, L, float(nv_s), 1.0 / B, _native.stream_ptr(c.device)))
            grad_v = ctx.plan.adjoint_split(g.reshape(r, -1), None, scale=grad_out)
        else:
            # A +-k pair of mirrored slices contributes g(k) + conj(g(-k)) = 2 g1(k) once the volume
            # gradient reaches the real eigenvolumes: the imaginary pass cancels within the pair, so
            # the half plane needs the real-coefficient pass only, with the pair weight.
            ml_grad_images(c, yhat, a.to(torch.float32), A1, L, nv_s, 1.0 / B, ctf=ctf_arg, out=g, pix_map=pix)
            bnd = ctx.bnd
            gov = None
            if bnd is not None:
                hidx, valid, co, wh, wo, order, oidx = bnd
                yf = yhat.reshape(B, -1)
                vf = valid.to(torch.float32)

                def exact(q, y, w):
                    # conj(Minv) Q + conj(z) (z^T Q - y~) in the kernel's units (weight 1)
                    e = (z.unsqueeze(1) @ q).squeeze(1) - y
                    gg = Minv.conj() @ q + z.conj().unsqueeze(2) * e.unsqueeze(1)
                    return (gg * (w / (B * nv_s)).unsqueeze(1)).permute(1, 0, 2).to(torch.complex64)

                qh = (torch.gather(c, 2, hidx.unsqueeze(0).expand(r, -1, -1)) * wh).permute(1, 0, 2).to(torch.complex128)
                qo = (co * wo).permute(1, 0, 2).to(torch.complex128)
                yh = (torch.gather(yf, 1, order) * (vf / L)).to(torch.complex128)
                yo = (torch.gather(yf, 1, oidx) * (vf / L)).to(torch.complex128)
                # the boundary pairs are exact: both members with weight 1 and their complex gradients
                gh = exact(qh, yh, wh)  # (r, B, nb)
                npix = g.shape[2]
                flat = (torch.arange(B, device=g.device).unsqueeze(1) * npix + hidx)[valid]
                gv = g.view(r, B * npix)
                gv[:, flat] = gh[:, valid]
                gov = exact(qo, yo, wo).contiguous().reshape(r, -1)
            # the half-plane values and the mirror values of the boundary pairs into one volume,
            # both scaled by the upstream gradient on the device
            grad_v = ctx.plan._disc_hb.adjoint_split(g.reshape(r, -1), gov, scale=grad_out)
        return grad_v.reshape(shape), None, None, None, None, None, None, None


def _disc_companion(plan, name: str):
    """A discretised plan kept on ``plan`` for a derived point set (same volume geometry)."""
    from .nufft_plan import NufftPlanDiscretized

    cp = getattr(plan, name, None)
    if cp is None:
        cp = NufftPlanDiscretized(plan.sz, upsample_factor=plan.upsample_factor, mode=plan.mode)
        setattr(plan, name, cp)
    return cp


_MIRROR_CACHE: dict = {}


def _mirror_index(L: int, device) -> torch.Tensor:
    """Row-major index of the -k partner of every pixel: (L - i, L - j) mod L for even L (the
    Nyquist lines pair with themselves), (L - 1 - i, L - 1 - j) for odd L."""
    key = (L, str(device))
    if key not in _MIRROR_CACHE:
        i = torch.arange(L)
        m = (L - i) % L if L % 2 == 0 else L - 1 - i
        _MIRROR_CACHE[key] = (m[:, None] * L + m[None, :]).reshape(-1).to(device)
    return _MIRROR_CACHE[key]

