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
plain CG and for PCG with the learned diagonal (and the probed exact
diagonal as a reference point), reported under ``mean_cg``.

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
    ap.add_argument("--cpu-images", type=int, default=256)
    ap.add_argument("--no-cg", action="store_true", help="skip the mean-CG leg")
    ap.add_argument("--subspace-batches", type=int, default=4,
                    help="SGD steps of the FP64-trajectory subspace comparison (0 skips it)")
    ap.add_argument("--cg-images", type=int, default=100000)
    ap.add_argument("--cg-cap", type=int, default=20000, help="iteration cap of each CG solve")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 1 GiB device copy (calibrates HBM byte counters)")
    return ap.parse_args()


def sgd_run(pkg, data, L, r, B, steps, warmup, form, timing=False):
    import torch

    covar_mod = importlib.import_module(pkg + ".covar")
    sgd = importlib.import_module(pkg + ".covar_sgd")
    dsm = importlib.import_module(pkg + ".dataset")
    native = importlib.import_module(pkg + "._native")
    dev = data.dataset.images.device
    torch.manual_seed(0)
    covar = covar_mod.Covar(L, r, pixel_var_estimate=1.0 / L).to(dev)
    loader = dsm.create_dataloader(data.dataset, batch_size=B)
    trainer = sgd.CovarTrainer(covar, loader, dev)
    trainer.setup_training(nufft_disc="nufft", objective_func="ml", form=form)
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


def subspace_run(pkg, data, L, r, B, n_batches):
    """Subspace of the chosen and plain forms vs an FP64 trajectory (same start, same batches)."""
    import torch

    val = importlib.import_module(pkg + ".validation")
    ds = data.dataset
    signal_var = max(float(ds.images[: n_batches * B].var()) - ds.noise_var, 1e-6)
    v0 = torch.randn((r,) + (L,) * 3, generator=torch.Generator().manual_seed(0)) * (signal_var / L) ** 0.5
    # lr 1e-6 per image: the reference workflow's default for the eigenvector fit (workflow.py:377)
    bases, secs = val.sgd_subspaces(ds, r, v0, batch_size=B, epochs=1, n_images=n_batches * B, lr=1e-6,
                                    forms=("fp64", "chosen", "plain_fp32"))
    err = val.subspace_errors(bases)
    q0, _ = torch.linalg.qr(v0.reshape(r, -1).double().T.to(bases["fp64"].device))
    return {
        "chosen_vs_fp64": err["chosen"],
        "plain_fp32_vs_fp64": err["plain_fp32"],
        "fp64_moved_from_start": val.subspace_distance(q0, bases["fp64"]),
        "steps": n_batches,
        "seconds": {k: round(v, 2) for k, v in secs.items()},
        "definition": (f"RMS sine of principal angles between r={r} eigen-subspaces after {n_batches} Adam steps "
                       f"of {B} images (lr 1e-6 per image) from the same start; reference trajectory in FP64 (parameters, images, "
                       f"NUFFT eps 1e-12, plain autograd objective)"),
    }


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
    # one untimed pass creates the rocFFT plans of the Toeplitz operator and warms the MLP
    pc.cg_comparison(reg_lhs, rhs, L, model, data.dataset.noise_var, maxiter=5, with_probe=False)
    res = pc.cg_comparison(reg_lhs, rhs, L, model, data.dataset.noise_var, rtol=1e-6, maxiter=cap)
    res["workload"] = (f"padded mean system (2L)^3 Toeplitz CG, L={L}, {n_images} images, r=0, 5% amplitude "
                       f"jitter, nearest back-projection us=2, rtol 1e-6, x0 = direct estimate")
    res["backprojection_s"] = round(t2 - t1, 3)
    if model is not None:
        res["time_speedup_learned_vs_plain"] = round(res["plain"]["seconds"] / res["learned"]["seconds"], 3)
        res["iteration_ratio_learned_vs_plain"] = round(res["learned"]["iterations"] /
                                                        max(res["plain"]["iterations"], 1), 3)
    del data, lhs, rhs, reg_lhs
    torch.cuda.empty_cache()
    return res


def cpu_baseline(L, r, B_cpu, noise_var=1.0):
    import numpy as np

    cpu = importlib.import_module("oracle.cpu_nufft")
    o = importlib.import_module("oracle.oracle")
    o.build(force=False)
    rng = np.random.default_rng(0)
    vols = rng.standard_normal((r, L, L, L)) / L
    R = np.linalg.qr(rng.standard_normal((B_cpu, 3, 3)))[0]
    pts = o.pts_rot(R, L)
    ctf = o.ctf(np.tile(np.array([[15000, 15000, 0, 300, 2.7, 0.1, 1.5]]), (B_cpu, 1)), L)
    images = rng.standard_normal((B_cpu, L, L))
    cores = os.cpu_count() or 1
    nu = cpu.CpuNufft(L, r, 1e-6, workers=cores)
    t0 = time.perf_counter()
    cpu.sgd_step_cpu(vols, images, pts, ctf, noise_var, nu)
    dt = time.perf_counter() - t0
    return {
        "value": B_cpu / dt,
        "unit": "images/s",
        "cores": cores,
        "kind": "port",
        "sample": f"one ML cost+gradient step (type-2 + type-1 ES NUFFT in FP64, numpy ML terms) on "
        f"{B_cpu} images at L={L}, r={r}; {dt:.1f} s",
    }


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
    n_img = (args.warmup + max(args.steps, plain_steps)) * B
    cfg = dsm.config_by_name("L128") if L == 128 else dsm.SyntheticConfig(L, n_img, r, [1.0] * r, 1.5, 0.02,
                                                                             (0.5, 3.0))
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
    plain = None
    if not args.no_plain:
        log("plain FP32 autograd form")
        dtp, _ = sgd_run(pkg, data, L, r, B, plain_steps, 1, "autograd")
        plain = {"value": plain_steps * B / dtp, "ms_per_step": 1e3 * dtp / plain_steps, "steps": plain_steps}
        log(f"plain: {plain['value']:.1f} images/s")

    # roofline of the dominant NUFFT kernel (algorithmic bytes per launch)
    nf = 2 * L
    # points per image on the chosen path: one pixel of every +-k pair (Hermitian half plane)
    ml_ops = importlib.import_module(pkg + ".ml_ops")
    d = int(ml_ops.half_plane_pixels(L, "cpu").numel())
    dom = max(("interp", "spread"), key=lambda k: kern[k]["total_ms"] or 0)
    avg_ms = kern[dom]["avg_ms"]
    # per point: 3 coordinates (12 B) + r complex values (8r B); plus the r fine grids once
    bytes_per_launch = B * (12 * d + 8 * r * d) + 8 * r * nf**3
    w = 7
    flops_per_launch = B * d * (w**3 * r * 4)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    # measured HBM bytes per launch of the same kernel (rocprofv3 FETCH_SIZE / WRITE_SIZE,
    # scaled by a 1 GiB copy) from the committed profile, when it covers this workload
    traffic = None
    tj = ROOT / "profiles" / f"hbm_traffic_sgd_L{L}_r{r}_B{B}.json"
    if tj.exists():
        try:
            traffic = round(json.loads(tj.read_text())[f"k_{dom}"]["traffic_GB"] * 1e9)
        except (KeyError, ValueError):
            traffic = None
    roof = {
        "kernel": f"k_{dom} (ES w=7, nf={nf}, r={r}, {B} images x {d} half-plane points)",
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "algorithmic_bytes": bytes_per_launch,
        "valu_tflops": round(flops_per_launch / (avg_ms * 1e-3) / 1e12, 2),
        "valu_frac_fp32_peak": round(flops_per_launch / (avg_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4),
        "avg_launch_ms": round(avg_ms, 3),
    }
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
    # achieved algorithmic GB/s of both NUFFT kernels
    per_kernel = {k: {"avg_ms": round(kern[k]["avg_ms"], 3),
                      "achieved_GBs": round(bytes_per_launch / (kern[k]["avg_ms"] * 1e-3) / 1e9, 1)}
                  for k in ("interp", "spread") if kern[k]["avg_ms"]}
    sub = None
    if args.subspace_batches > 0:
        log(f"subspace agreement vs FP64 over {args.subspace_batches} steps")
        sub = subspace_run(pkg, data, L, r, B, args.subspace_batches)
        log(f"subspace: chosen {sub['chosen_vs_fp64']:.2e}, plain {sub['plain_fp32_vs_fp64']:.2e} ({sub['seconds']})")
    cg = None
    if not args.no_cg:
        del data
        torch.cuda.empty_cache()
        log("mean CG leg (plain CG vs learned-diagonal PCG)")
        cg = mean_cg_run(pkg, L, args.cg_images, dev, args.cg_cap)
        log(f"mean CG: plain {cg['plain']['iterations']} it / {cg['plain']['seconds']} s; learned "
            f"{cg.get('learned', {}).get('iterations')} it / {cg.get('learned', {}).get('seconds')} s")
    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline on {args.cpu_images} images")
        try:
            cpu = cpu_baseline(L, r, args.cpu_images)
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
        "speedup_vs_plain": round(ips / plain["value"], 3) if plain else None,
        "kernels_ms": {k: round(v["avg_ms"], 3) if v["avg_ms"] else None for k, v in kern.items()},
        "roofline": roof,
        "nufft_kernels": per_kernel,
        "contractions": contr,
        "subspace_error": sub,
        "mean_cg": cg,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
