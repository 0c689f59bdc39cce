"""Benchmark of the COV3D covariance-SGD hot path on one MI355X.

One step = one minibatch SGD step of the rank-r ML objective on the NUFFT
path at L=128, r=20, B=1024 (100k-image workload, synthetic data resident in
HBM): pts_rot from rotations on the device, type-2 NUFFT of the r
eigenvolumes, CTF, per-image Gram + FP64 Cholesky, gradient, type-1 NUFFT
back-projection, Adam.

``value`` is the chosen form (split-BF16 MFMA contractions + hand-written
adjoint); the plain FP32 autograd form on the same GPU is reported beside it,
with per-kernel HIP-event times, the roofline of the dominant kernel and a
CPU baseline (the CPU port under oracle/, on a bounded sample).

The second half of the metric is the mean-volume CG at L=128 (100k images,
r=0, 5% amplitude jitter): time and iterations to relative residual 1e-6 for
plain CG and for PCG with the learned per-shell diagonal (``learned``); the
per-voxel network, the probed exact diagonal and its shell average are
reference points, all reported under ``mean_cg``.

``pipeline_L256`` is the full chosen-form pipeline at L=256 with the images
streamed from pinned host memory (a bounded sample of the 300k-image workload):
SGD images/s with the NUFFT kernels' rates, the mean system's plain CG and
learned PCG per iteration, and the Wiener coordinates' images/s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cg]
"""

from __future__ import annotations

import argparse
import importlib
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# LDS: a conflict-free ds_read_b128 stream measured on the MI355X (16 waves per CU, all CUs);
# AMD publishes no chip-wide LDS bandwidth figure
LDS_STREAM_TBS = 97.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--L", type=int, default=128)
    ap.add_argument("--rank", type=int, default=20)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--plain-steps", type=int, default=None)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-images", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=3, help="timed CPU steps (the median is reported)")
    ap.add_argument("--no-nudft", action="store_true", help="skip the FP64 NUDFT check of the NUFFT outputs")
    ap.add_argument("--bilinear-steps", type=int, default=6,
                    help="timed steps of each form on the discretised bilinear path (0 skips that leg)")
    ap.add_argument("--no-cg", action="store_true", help="skip the mean-CG leg")
    ap.add_argument("--subspace-batches", type=int, default=98,
                    help="SGD steps of the FP64-trajectory subspace comparison (98 = one epoch of the "
                         "100k-image workload at batch 1024; 0 skips it)")
    ap.add_argument("--fp64-eps", type=float, default=1e-9, help="NUFFT tolerance of the FP64 trajectory")
    ap.add_argument("--cg-images", type=int, default=100000)
    ap.add_argument("--cg-cap", type=int, default=20000, help="iteration cap of each CG solve")
    ap.add_argument("--pipeline-images", type=int, default=8192,
                    help="images of the L=256 streamed pipeline leg (0 skips it)")
    ap.add_argument("--pipeline-cg-iters", type=int, default=300,
                    help="iterations of each mean-CG solve in the L=256 pipeline leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 1 GiB device copy (calibrates HBM byte counters)")
    return ap.parse_args()


def sgd_run(pkg, data, L, r, B, steps, warmup, form, timing=False, precision=None, nufft_disc="nufft"):
    import torch

    covar_mod = importlib.import_module(pkg + ".covar")
    sgd = importlib.import_module(pkg + ".covar_sgd")
    dsm = importlib.import_module(pkg + ".dataset")
    native = importlib.import_module(pkg + "._native")
    dev = data.dataset.device
    torch.manual_seed(0)
    covar = covar_mod.Covar(L, r, pixel_var_estimate=1.0 / L).to(dev)
    loader = dsm.create_dataloader(data.dataset, batch_size=B)
    trainer = sgd.CovarTrainer(covar, loader, dev)
    kw = {} if precision is None else {"precision": precision}
    trainer.setup_training(nufft_disc=nufft_disc, objective_func="ml", form=form, **kw)
    batches = iter(loader)
    for _ in range(warmup):
        images, pts, filt = trainer.prepare_batch(next(batches))
        trainer.run_batch(images, pts, filt)
    torch.cuda.synchronize()
    if timing:
        native.timing_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        images, pts, filt = trainer.prepare_batch(next(batches))
        trainer.run_batch(images, pts, filt)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kern = {}
    if timing:
        for k in native.TIMING_KINDS:
            tot, cnt = native.timing_read(k)
            kern[k] = {"total_ms": tot, "launches": cnt, "avg_ms": tot / cnt if cnt else None}
        native.timing_enable(False)
    trainer.complete_training()  # the discretised path moved the images to the Fourier domain
    del trainer
    return dt, kern


def contraction_run(pkg, data, L, r, B, reps=20):
    """ml_forward (split-BF16 MFMA Gram + FP64 Cholesky) on one batch's half-plane type-2 values."""
    import torch

    npl = importlib.import_module(pkg + ".nufft_plan")
    ml = importlib.import_module(pkg + ".ml_ops")
    pf = importlib.import_module(pkg + ".projection_funcs")
    dev = data.dataset.images.device
    images, pts, filt, _ = data.dataset[0:B]
    pix = ml.half_plane_pixels(L, dev)
    plan = npl.NufftPlan((L,) * 3, batch_size=r, device=dev)
    plan.set_points_flat(pts.transpose(0, 1).reshape(3, B, L * L)[:, :, pix.long()].reshape(3, -1), B)
    vols = torch.randn((r, L, L, L), device=dev) / L
    c = plan.execute_forward_flat(vols)[:r].reshape(r, B, pix.numel())
    yhat = pf.centered_fft2(images.float()) / L
    out = {}
    for name, prec in (("bf16x3", ml.PRECISION_BF16X3), ("fp32", ml.PRECISION_FP32)):
        ml.ml_forward(c, yhat, L, 1.0, ctf=filt, precision=prec, pix_map=pix)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            ml.ml_forward(c, yhat, L, 1.0, ctf=filt, precision=prec, pix_map=pix)
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        npx = pix.numel()
        # useful Gram flops: (r+1)^2 products over 2*npix reals, x3 for the split-BF16 products
        useful = B * (r + 1) ** 2 * 2 * (2 * npx) * (3 if prec == ml.PRECISION_BF16X3 else 1)
        read = B * (r * npx * 8 + npx * 8 + npx * 4)
        out[name] = {"ms": round(ms, 3), "tflops": round(useful / (ms * 1e-3) / 1e12, 2),
                     "mfma_frac_bf16_peak": round(useful / (ms * 1e-3) / 1e12 / 2500.0, 4) if prec else None,
                     "GBs": round(read / (ms * 1e-3) / 1e9, 1)}
    del plan
    return out


def subspace_run(pkg, data, L, r, B, n_batches, fp64_eps):
    """Subspaces of the SGD forms vs an FP64 trajectory (same start, same batches, lr 1e-6 per image).

    chosen = split-BF16 Gram on the half plane; fused_fp32 = FP32 Gram on the half plane;
    chosen_full = split-BF16 Gram on the full plane; plain_fp32 = the reference's FP32 autograd
    objective.  Comparing them says which of the two changes (BF16 contractions, half plane)
    moves the learned subspace."""
    import torch

    val = importlib.import_module(pkg + ".validation")
    ds = data.dataset
    signal_var = max(float(ds.images[: n_batches * B].var()) - ds.noise_var, 1e-6)
    v0 = torch.randn((r,) + (L,) * 3, generator=torch.Generator().manual_seed(0)) * (signal_var / L) ** 0.5
    forms = ("fp64", "chosen", "fused_fp32", "chosen_full", "plain_fp32")
    # lr 1e-6 per image: the reference workflow's default for the eigenvector fit (workflow.py:377)
    bases, secs = val.sgd_subspaces(ds, r, v0, batch_size=B, epochs=1, n_images=min(n_batches * B, len(ds)),
                                    lr=1e-6, forms=forms, fp64_eps=fp64_eps)
    err = val.subspace_errors(bases)
    q0, _ = torch.linalg.qr(v0.reshape(r, -1).double().T.to(bases["fp64"].device))
    out = {f"{k}_vs_fp64": err[k] for k in forms[1:]}
    out.update({
        "chosen_vs_plain_fp32": val.subspace_distance(bases["plain_fp32"], bases["chosen"]),
        "fp64_moved_from_start": val.subspace_distance(q0, bases["fp64"]),
        "steps": n_batches,
        "images": min(n_batches * B, len(ds)),
        "seconds": {k: round(v, 2) for k, v in secs.items()},
        "definition": (f"RMS sine of principal angles between r={r} eigen-subspaces after {n_batches} Adam steps "
                       f"of {B} images (lr 1e-6 per image) from the same start; reference trajectory in FP64 "
                       f"(parameters, images, NUFFT eps {fp64_eps:g}, plain autograd objective)"),
    })
    return out


def mean_cg_run(pkg, L, n_images, dev, cap=20000):
    """Mean reconstruction system at L (100k images, r=0, 5% amplitude jitter): plain CG vs PCG to 1e-6."""
    import torch

    dsm = importlib.import_module(pkg + ".dataset")
    mean = importlib.import_module(pkg + ".mean")
    pc = importlib.import_module(pkg + ".precond")
    cfg = dsm.config_by_name("L128-mean") if L == 128 else dsm.SyntheticConfig(L, n_images, 0, [], 1.5, 0.02,
                                                                                (0.5, 3.0), amp_jitter=0.05)
    t0 = time.perf_counter()
    data = dsm.make_synthetic(cfg, dev, n_images=n_images, subtract_mean=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lhs, rhs = mean.backproject_lhs_rhs(data.dataset, batch_size=1024)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    log(f"mean CG: {n_images} images generated in {t1 - t0:.1f} s, back-projected in {t2 - t1:.2f} s")
    reg_lhs = mean.regularize_lhs(lhs)
    try:
        model = pc.load_default(dev)
    except FileNotFoundError as e:
        log(f"mean CG: {e}; learned leg skipped")
        model = None
    try:
        extra = {"voxel": pc.load_default(dev, "voxel")}
    except FileNotFoundError as e:
        log(f"mean CG: {e}; per-voxel leg skipped")
        extra = {}
    # one untimed pass creates the rocFFT plans of the Toeplitz operator and warms the MLP
    pc.cg_comparison(reg_lhs, rhs, L, model, data.dataset.noise_var, maxiter=5, with_probe=False,
                     extra_models=extra)
    res = pc.cg_comparison(reg_lhs, rhs, L, model, data.dataset.noise_var, rtol=1e-6, maxiter=cap,
                           extra_models=extra, with_shell_probe=True)
    res["workload"] = (f"padded mean system (2L)^3 Toeplitz CG, L={L}, {n_images} images, r=0, 5% amplitude "
                       f"jitter, nearest back-projection us=2, rtol 1e-6, x0 = direct estimate")
    res["backprojection_s"] = round(t2 - t1, 3)
    if model is not None:
        res["time_speedup_learned_vs_plain"] = round(res["plain"]["seconds"] / res["learned"]["seconds"], 3)
        res["iteration_ratio_learned_vs_plain"] = round(res["learned"]["iterations"] /
                                                        max(res["plain"]["iterations"], 1), 3)
    if extra:
        res["time_speedup_learned_voxel_vs_plain"] = round(res["plain"]["seconds"] / res["learned_voxel"]["seconds"], 3)
    del data, lhs, rhs, reg_lhs
    torch.cuda.empty_cache()
    return res


def nufft_rates(kern, L, r, B, d, w=7):
    """Algorithmic HBM GB/s and FP32 TFLOP/s of the interpolation and spreading launches."""
    nf = 2 * L
    bytes_per_launch = B * (12 * d + 8 * r * d) + 8 * r * nf**3
    flops_per_launch = B * d * (w**3 * r * 4)
    out = {}
    for k in ("interp", "spread"):
        ms = kern[k]["avg_ms"]
        if ms:
            out[k] = {"avg_ms": round(ms, 3), "achieved_GBs": round(bytes_per_launch / (ms * 1e-3) / 1e9, 1),
                      "achieved_TFLOPs": round(flops_per_launch / (ms * 1e-3) / 1e12, 2),
                      "frac_fp32_peak": round(flops_per_launch / (ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4)}
    return out


def pipeline_run(pkg, dev, n_images, steps, cg_iters, L=256, r=20, B=1024):
    """The L=256 pipeline on one GPU with the images streamed from pinned host memory: covariance
    SGD steps of the chosen form (batches of B copied on a second stream behind the step), the mean
    system (back-projection, plain CG and learned-diagonal PCG for a fixed number of iterations)
    and the Wiener coordinates of every image (FP64 r x r Cholesky per image)."""
    import numpy as np
    import torch

    dsm = importlib.import_module(pkg + ".dataset")
    mean = importlib.import_module(pkg + ".mean")
    pc = importlib.import_module(pkg + ".precond")
    wc = importlib.import_module(pkg + ".wiener_coords")
    ml_ops = importlib.import_module(pkg + ".ml_ops")
    cfg = dsm.SyntheticConfig(L, n_images, r, list(np.geomspace(10, 0.1, r)), 1.0, 0.01, (0.5, 3.5), 0.0, 1005)
    t0 = time.perf_counter()
    data = dsm.make_synthetic(cfg, dev, n_images=n_images, subtract_mean=False, host_images=True)
    torch.cuda.synchronize()
    log(f"pipeline L={L}: {n_images} images generated into pinned host memory in {time.perf_counter() - t0:.1f} s")
    out = {"workload": (f"L={L}, r={r}, batch {B}, {n_images} images of {L}x{L} in pinned host memory "
                        f"(pixel 1.0 A, SNR 0.01, defocus U[0.5,3.5] um), streamed to the device per batch"),
           "images": n_images}
    steps = min(steps, n_images // B - 1)
    dt, kern = sgd_run(pkg, data, L, r, B, steps, 1, "fused", timing=True)
    d = int(ml_ops.half_plane_pixels(L, "cpu").numel())
    out["sgd"] = {"value": round(steps * B / dt, 1), "unit": "images/s", "ms_per_step": round(1e3 * dt / steps, 2),
                  "steps": steps, "nufft_kernels": nufft_rates(kern, L, r, B, d)}
    log(f"pipeline L={L}: SGD {out['sgd']['value']} images/s ({out['sgd']['ms_per_step']} ms/step)")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lhs, rhs = mean.backproject_lhs_rhs(data.dataset, batch_size=B)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    reg_lhs = mean.regularize_lhs(lhs)
    try:
        model = pc.load_default(dev)
    except FileNotFoundError:
        model = None
    pc.cg_comparison(reg_lhs, rhs, L, model, data.dataset.noise_var, maxiter=3, with_probe=False)
    res = pc.cg_comparison(reg_lhs, rhs, L, model, data.dataset.noise_var, rtol=1e-6, maxiter=cg_iters,
                           with_probe=False)
    res["backprojection_images_per_s"] = round(n_images / (t2 - t1), 1)
    res["note"] = f"each solve stops after {cg_iters} iterations; the residual it reached is final_residual"
    out["mean_cg"] = res
    log(f"pipeline L={L}: mean CG {res['plain'].get('ms_per_iteration')} ms/it plain, "
        f"{(res.get('learned') or {}).get('ms_per_iteration')} ms/it learned")
    del lhs, rhs, reg_lhs
    ev = (data.eigenvecs).to(dev)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    coords = wc.latentMAP(data.dataset, ev, data.eigenvals.to(dev), batch_size=B)
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    out["wiener_coords"] = {"value": round(n_images / (t4 - t3), 1), "unit": "images/s",
                            "finite": bool(torch.isfinite(coords).all())}
    log(f"pipeline L={L}: Wiener coordinates {out['wiener_coords']['value']} images/s")
    del data, coords, ev
    torch.cuda.empty_cache()
    return out


def cpu_baseline(L, r, B_cpu, steps=3, noise_var=1.0):
    """The FP32 CPU port (oracle/cpu_nufft.py + cpu_nufft.cpp, OpenMP) timed on this host: the
    median of ``steps`` SGD steps (setpts, type-2 NUFFT, ML cost and gradient on the half plane,
    type-1 NUFFT, Adam) of B_cpu images at L, r, with the step's own point set each time."""
    import numpy as np

    cpu = importlib.import_module("oracle.cpu_nufft")
    o = importlib.import_module("oracle.oracle")
    o.build(force=False)
    cpu.build(force=False)
    rng = np.random.default_rng(0)
    vols = (rng.standard_normal((r, L, L, L)) / L).astype(np.float32)
    m, v = np.zeros_like(vols), np.zeros_like(vols)
    nu = cpu.CpuNufft(L, r, 1e-6)
    times = []
    for k in range(steps):
        R = np.linalg.qr(rng.standard_normal((B_cpu, 3, 3)))[0]
        pts = o.pts_rot(R, L).astype(np.float32)
        ctf = o.ctf(np.tile(np.array([[15000, 15000, 0, 300, 2.7, 0.1, 1.5]]), (B_cpu, 1)), L).astype(np.float32)
        images = rng.standard_normal((B_cpu, L, L)).astype(np.float32)
        t0 = time.perf_counter()
        _, g = cpu.sgd_step_cpu(vols, images, pts, ctf, noise_var, nu)
        cpu.adam_step(vols, g, m, v, k + 1, 1e-6 * B_cpu)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {
        "value": B_cpu / med,
        "unit": "images/s",
        "cores": nu.workers,
        "kind": "port",
        "dtype": "f32",
        "sample": (f"median of {steps} SGD steps of {B_cpu} images at L={L}, r={r} (FP32 C++/OpenMP ES-NUFFT "
                   f"spreading/interpolation, pruned scipy.fft fine-grid FFTs, numpy ML terms on the half plane, "
                   f"Adam); step seconds {[round(t, 2) for t in times]}"),
        "pinned_by": "tests/test_cpu_port.py (NUFFT vs FP64 NUDFT <= 1e-5; ML step vs FP64 oracle <= 1e-4)",
    }


def nudft_check(pkg, data, L, r, B, sub=512, per_image=4, n_modes=64):
    """Error of the benchmark's NUFFT outputs against the FP64 direct NUDFT on a 512-image subset of
    the first batch (the chosen form's half-plane points): type 2 of the ground-truth eigenvolumes at
    ``per_image`` random pixels of each subset image; type 1 of the resulting values on the subset's
    points at ``n_modes`` random output modes."""
    import numpy as np
    import torch

    o = importlib.import_module("oracle.oracle")
    o.build(force=False)
    npl = importlib.import_module(pkg + ".nufft_plan")
    ml = importlib.import_module(pkg + ".ml_ops")
    dev = data.dataset.images.device
    _, pts, _, _ = data.dataset[0:B]
    pix = ml.half_plane_pixels(L, dev).long()
    npx = pix.numel()
    P = pts[:, :, pix].transpose(0, 1).reshape(3, -1)
    P = (torch.remainder(P + torch.pi, 2 * torch.pi) - torch.pi).contiguous()
    vols = (data.eigenvecs * data.eigenvals.sqrt().reshape(-1, 1, 1, 1)).float()[:r].contiguous()
    t0 = time.perf_counter()
    plan = npl.NufftPlan((L,) * 3, batch_size=r, device=dev)
    plan.set_points_flat(P, B)
    c = plan.execute_forward_flat(vols)
    rng = np.random.default_rng(5)
    cols = (np.arange(sub)[:, None] * npx + rng.integers(0, npx, (sub, per_image))).ravel()
    ci = torch.tensor(cols, device=dev)
    ref2 = o.nudft_type2(vols.cpu().numpy().astype(np.complex128), P[:, ci].cpu().numpy().astype(np.float64))
    e2 = float(np.linalg.norm(c[:, ci].cpu().numpy() - ref2) / np.linalg.norm(ref2))
    del plan
    Ps = P[:, : sub * npx].contiguous()
    cs = c[:, : sub * npx].contiguous()
    plan1 = npl.NufftPlan((L,) * 3, batch_size=r, device=dev)
    plan1.set_points_flat(Ps, sub)
    f1 = plan1.execute_adjoint(cs).reshape(r, -1)
    modes = rng.choice(L**3, n_modes, replace=False)
    ref1 = o.nudft_type1_modes(cs.cpu().numpy().astype(np.complex128), Ps.cpu().numpy().astype(np.float64),
                               (L, L, L), modes)
    e1 = float(np.linalg.norm(f1[:, torch.tensor(modes, device=dev)].cpu().numpy() - ref1) / np.linalg.norm(ref1))
    del plan1
    torch.cuda.empty_cache()
    return {"type2": e2, "type1": e1, "subset_images": sub,
            "type2_points": int(cols.size), "type1_points": int(Ps.shape[1]), "type1_modes": int(n_modes),
            "transforms": r, "reference": "FP64 direct NUDFT (oracle/nudft.c)",
            "seconds": round(time.perf_counter() - t0, 1)}


def ml_vs_fp64(pkg, data, L, r, n_images=48):
    """Cost and gradient of the three forms of the ML objective against the FP64 autograd objective
    (FP64 NUFFT plan, eps 1e-12) on the first ``n_images`` images of the benchmark data."""
    import torch

    npl = importlib.import_module(pkg + ".nufft_plan")
    sgd = importlib.import_module(pkg + ".covar_sgd")
    dev = data.dataset.images.device
    images, pts, ctf, _ = data.dataset[0:n_images]
    nv = data.dataset.noise_var
    vols = (data.eigenvecs * data.eigenvals.sqrt().reshape(-1, 1, 1, 1)).float()[:r]
    vols = vols + 0.3 * vols.std() * torch.randn(vols.shape, generator=torch.Generator().manual_seed(4)).to(dev)
    p64 = npl.NufftPlan((L,) * 3, batch_size=r, eps=1e-12, dtype=torch.float64, device=dev)
    p64.setpts(pts.double())
    v = vols.double().clone().requires_grad_(True)
    ref = sgd.cost_maximum_liklihood(v, images.double(), p64, ctf.double(), nv, form="autograd")
    ref.backward()
    ref_c, ref_g = ref.item(), v.grad
    del p64
    plan = npl.NufftPlan((L,) * 3, batch_size=r, device=dev)
    plan.setpts(pts)
    out = {}
    for tag, form, prec in (("chosen", "fused", 1), ("fused_fp32", "fused", 0), ("plain_fp32", "autograd", 0)):
        w = vols.clone().requires_grad_(True)
        cst = sgd.cost_maximum_liklihood(w, images, plan, ctf, nv, form=form, precision=prec)
        cst.backward()
        out[tag] = {"cost": abs(cst.item() - ref_c) / abs(ref_c),
                    "gradient": float((w.grad.double() - ref_g).norm() / ref_g.norm())}
    del plan
    torch.cuda.empty_cache()
    out["images"] = n_images
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import torch

    pkg = "3dcovarestimation_amd"
    importlib.import_module(pkg)._native.load()
    dsm = importlib.import_module(pkg + ".dataset")
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    dev = torch.device("cuda:0")
    L, r, B = args.L, args.rank, args.batch
    plain_steps = args.plain_steps or max(2, args.steps // 2)
    n_img = max((args.warmup + max(args.steps, plain_steps, args.bilinear_steps)) * B, args.subspace_batches * B)
    cfg = dsm.config_by_name("L128") if L == 128 else dsm.SyntheticConfig(L, n_img, r, [1.0] * r, 1.5, 0.02,
                                                                             (0.5, 3.0))
    n_img = min(n_img, cfg.n_images)
    cfg.rank = r
    cfg.eigenvalues = list(cfg.eigenvalues)[:r] + [0.1] * max(0, r - len(cfg.eigenvalues))
    log(f"generating {n_img} synthetic images L={L} r={r}")
    t0 = time.perf_counter()
    data = dsm.make_synthetic(cfg, dev, n_images=n_img)
    torch.cuda.synchronize()
    log(f"generated in {time.perf_counter() - t0:.1f} s")

    log("chosen form (bf16x3 MFMA + hand adjoint)")
    dt, kern = sgd_run(pkg, data, L, r, B, args.steps, args.warmup, "fused", timing=True)
    ips = args.steps * B / dt
    log(f"chosen: {ips:.1f} images/s ({1e3 * dt / args.steps:.2f} ms/step)")
    plain = fused32 = None
    if not args.no_plain:
        log("fused FP32 form (FP32 Gram + hand adjoint)")
        dtf, _ = sgd_run(pkg, data, L, r, B, plain_steps, 1, "fused", precision=0)
        fused32 = {"value": plain_steps * B / dtf, "ms_per_step": 1e3 * dtf / plain_steps, "steps": plain_steps}
        log(f"fused FP32: {fused32['value']:.1f} images/s")
        log("plain FP32 autograd form")
        dtp, _ = sgd_run(pkg, data, L, r, B, plain_steps, 1, "autograd")
        plain = {"value": plain_steps * B / dtp, "ms_per_step": 1e3 * dtp / plain_steps, "steps": plain_steps}
        log(f"plain: {plain['value']:.1f} images/s")

    # roofline of the dominant NUFFT kernel.  Both NUFFT kernels are arithmetic-bound (about 80 flop per
    # algorithmic byte against an FP32 ridge of about 20), so the binding figure is FP32 vector
    # throughput; the algorithmic HBM rate is reported beside it.
    nf = 2 * L
    # points per image on the chosen path: one pixel of every +-k pair (Hermitian half plane)
    ml_ops = importlib.import_module(pkg + ".ml_ops")
    d = int(ml_ops.half_plane_pixels(L, "cpu").numel())
    dom = max(("interp", "spread"), key=lambda k: kern[k]["total_ms"] or 0)
    avg_ms = kern[dom]["avg_ms"]
    # per point: 3 coordinates (12 B) + r complex values (8r B); plus the r fine grids once
    bytes_per_launch = B * (12 * d + 8 * r * d) + 8 * r * nf**3
    w = 7
    # per point and transform: w^3 taps, each a real weight times a complex value (2 FMAs = 4 flops)
    flops_per_launch = B * d * (w**3 * r * 4)
    achieved_gbs = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    achieved_tf = flops_per_launch / (avg_ms * 1e-3) / 1e12
    # measured HBM bytes per launch of the same kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a
    # 1 GiB copy) from the committed profile of this workload: not measured in this run
    traffic, traffic_src = None, None
    tj = ROOT / "profiles" / f"hbm_traffic_sgd_L{L}_r{r}_B{B}.json"
    if tj.exists():
        try:
            traffic = round(json.loads(tj.read_text())[f"k_{dom}"]["traffic_GB"] * 1e9)
            traffic_src = f"profiles/{tj.name} (committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE run, not this run)"
        except (KeyError, ValueError):
            traffic = None
    roof = {
        "kernel": f"k_{dom} (ES w=7, nf={nf}, r={r}, {B} images x {d} half-plane points)",
        "bound": "valu",
        "achieved": round(achieved_tf, 2),
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tf / FP32_PEAK_TFLOPS, 4),
        "frac_of_nonpacked_peak": round(achieved_tf / (FP32_PEAK_TFLOPS / 2), 4),
        "traffic": traffic,
        "traffic_source": traffic_src,
        "algorithmic_flops": flops_per_launch,
        "algorithmic_bytes": bytes_per_launch,
        "hbm_achieved_GBs": round(achieved_gbs, 1),
        "hbm_frac": round(achieved_gbs / HBM_PEAK_GBS, 4),
        "avg_launch_ms": round(avg_ms, 3),
        "timing": "HIP events on the launch stream around every launch in the timed steps (cov3d_timing_*)",
    }
    if kern["interp"]["avg_ms"]:
        # the interpolation is LDS-bound (SQ counters: LDS active ~65 %, VALU ~68 %): every point reads its
        # w^3 taps of every transform from the LDS tile once, 8 bytes per tap and transform
        lds_bytes = B * d * w**3 * 8 * r
        lds_tbs = lds_bytes / (kern["interp"]["avg_ms"] * 1e-3) / 1e12
        roof["interp_lds"] = {"kernel": "k_interp_pipe", "bytes_per_launch": lds_bytes,
                              "achieved_TBs": round(lds_tbs, 1), "peak_TBs": LDS_STREAM_TBS,
                              "frac": round(lds_tbs / LDS_STREAM_TBS, 4),
                              "peak_source": "measured conflict-free ds_read_b128 stream (no published figure)"}
    if args.copy_probe:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 3 x 1 GiB device-to-device copies")
    contr = None
    try:
        contr = contraction_run(pkg, data, L, r, B)
        log(f"contractions: {contr}")
    except Exception as e:  # reported, never required
        contr = {"error": repr(e)}
    # achieved algorithmic rates of both NUFFT kernels
    per_kernel = {k: {"avg_ms": round(kern[k]["avg_ms"], 3),
                      "achieved_GBs": round(bytes_per_launch / (kern[k]["avg_ms"] * 1e-3) / 1e9, 1),
                      "achieved_TFLOPs": round(flops_per_launch / (kern[k]["avg_ms"] * 1e-3) / 1e12, 2)}
                  for k in ("interp", "spread") if kern[k]["avg_ms"]}
    bil = None
    if args.bilinear_steps > 0:
        log("discretised bilinear path (the reference's default trainer): fused vs autograd Fourier-domain ML")
        bil = {}
        for tag, form in (("fused", "fused"), ("autograd", "autograd")):
            dtb, _ = sgd_run(pkg, data, L, r, B, args.bilinear_steps, 1, form, nufft_disc="bilinear")
            bil[tag] = {"value": round(args.bilinear_steps * B / dtb, 1),
                        "ms_per_step": round(1e3 * dtb / args.bilinear_steps, 2)}
        bil["speedup_fused_vs_autograd"] = round(bil["fused"]["value"] / bil["autograd"]["value"], 3)
        bil["workload"] = (f"covariance SGD step, bilinear slice path (us=2), Fourier-domain ML, L={L}, r={r}, "
                           f"batch={B}; fused = native slice gather/scatter + MFMA Gram + hand adjoint")
        log(f"bilinear path: {bil}")
    nud = mlf = None
    if not args.no_nudft:
        log("NUFFT outputs vs the FP64 NUDFT on a 512-image subset")
        nud = nudft_check(pkg, data, L, r, B)
        log(f"NUDFT check: {nud}")
        log("ML cost and gradient of the three forms vs FP64")
        mlf = ml_vs_fp64(pkg, data, L, r)
        log(f"ML vs FP64: {mlf}")
    sub = None
    if args.subspace_batches > 0:
        log(f"subspace agreement vs FP64 over {args.subspace_batches} steps")
        sub = subspace_run(pkg, data, L, r, B, args.subspace_batches, args.fp64_eps)
        log(f"subspace: {sub}")
    cg = None
    if not args.no_cg:
        del data
        torch.cuda.empty_cache()
        log("mean CG leg (plain CG vs learned-diagonal PCG)")
        cg = mean_cg_run(pkg, L, args.cg_images, dev, args.cg_cap)
        log(f"mean CG: plain {cg['plain']['iterations']} it / {cg['plain']['seconds']} s; learned "
            f"{cg.get('learned', {}).get('iterations')} it / {cg.get('learned', {}).get('seconds')} s")
    pipe = None
    if args.pipeline_images > 0:
        log(f"L=256 pipeline leg: {args.pipeline_images} images streamed from pinned host memory")
        pipe = pipeline_run(pkg, dev, args.pipeline_images, max(2, args.steps // 2), args.pipeline_cg_iters)
    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline: {args.cpu_steps} steps of {args.cpu_images} images")
        try:
            cpu = cpu_baseline(L, r, args.cpu_images, args.cpu_steps)
            log(f"CPU baseline: {cpu['value']:.2f} images/s on {cpu['cores']} threads")
        except Exception as e:  # the baseline is reported, never required
            cpu = {"error": repr(e)}
    out = {
        "metric": "SGD images/s at L=128,r=20 (BF16+adjoint vs plain FP32); CG time to 1e-6",
        "value": round(ips, 2),
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 NUFFT, bf16x3 MFMA contractions, f64 Cholesky",
        "data": "synthetic (Gaussian-blob eigenvolumes, SO(3)-uniform poses, CTF, SNR 0.02), resident in HBM",
        "config": {"workload": f"covariance SGD step, NUFFT path, ML objective, L={L}, r={r}, batch={B}",
                   "L": L, "rank": r, "batch": B, "form": "bf16x3 Gram + hand-written adjoint"},
        "plain_fp32_autograd": plain,
        "fused_fp32": fused32,
        "speedup_vs_plain": round(ips / plain["value"], 3) if plain else None,
        "speedup_vs_fused_fp32": round(ips / fused32["value"], 3) if fused32 else None,
        "kernels_ms": {k: round(v["avg_ms"], 3) if v["avg_ms"] else None for k, v in kern.items()},
        "roofline": roof,
        "nufft_kernels": per_kernel,
        "contractions": contr,
        "bilinear_path": bil,
        "nudft_rel_err": nud,
        "ml_vs_fp64": mlf,
        "subspace_error": sub,
        "mean_cg": cg,
        "pipeline_L256": pipe,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
