#!/usr/bin/env python
"""Benchmark of the ReconNet reconstruction path on one MI355X.

Metric: slices/s to reach the plain SQS-TV RMSE (HU) on 512^2 fan-beam slices
(32 slices of 512^2 at 0.7 mm, 1152 views over 2 pi, 736 bins of 1.0 mm,
SID 595 mm, SDD 1085.6 mm, 12 ordered subsets, TV prior beta = 1e-4 with
delta = 1e-3, Poisson noise at I0 = 1e5).  The workload is synthetic: seeded
random-ellipse phantoms projected by the library's own Siddon projector.

Plain SQS-TV is first run to convergence (untimed calibration): k_conv is the
first iteration whose RMSE is within 1 % of the RMSE after k_max iterations and
whose relative update ||dx||/||x|| is below 1e-4.  A timed step is one
complete reconstruction of the 32-slice batch: k_conv OS-SQS-TV passes from a
zero image, inputs resident in HBM.  When trained weights of the learned
unrolled network are present (reconnet_amd/weights/unrolled_fan512.npz) the
step is the 10-iteration learned reconstruction instead, and it counts only if
its RMSE does not exceed the plain converged RMSE.

Prints one JSON line.  Options: --gpus (1 only), --steps, --warmup.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--slices", type=int, default=32)
    ap.add_argument("--kmax", type=int, default=100)
    ap.add_argument("--subsets", type=int, default=12)
    ap.add_argument("--beta", type=float, default=1e-4)
    ap.add_argument("--delta", type=float, default=1e-3)
    ap.add_argument("--i0", type=float, default=1e5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def cpu_baseline(geo, truth_np, noisy_np, w_np, k_conv, nsub, threads):
    """One OS-SQS data pass of the FP32 CPU oracle on slice 0 (OpenMP), scaled
    to k_conv passes per slice."""
    import oracle as orc
    from reconnet_amd.geometry import interleaved_order

    if threads > 0:
        orc.set_threads(threads)
    order = interleaved_order(geo.nview, nsub)
    deg = geo.angles()[order]
    b = np.ascontiguousarray(noisy_np[0][:, order])[None, :, :, None, None]
    w = np.ascontiguousarray(w_np[0][:, order])[None, :, :, None, None]
    g = geo.oracle_geo()
    ones = np.ones((1, geo.nx, geo.ny, 1, 1), np.float32)
    fp1 = orc.siddon_fan_fp(ones, deg, g, geo.nu, precision="f32")
    D = orc.siddon_fan_bp((fp1 * w).astype(np.float32), deg, g, (geo.nx, geo.ny, 1), precision="f32")
    x0 = np.zeros((1, geo.nx, geo.ny, 1, 1), np.float32)
    t0 = time.perf_counter()
    orc.os_sqs(x0, b, D, w, deg, g, nsub, lb=0.0, precision="f32")
    t = time.perf_counter() - t0
    return {"value": 1.0 / (k_conv * t), "unit": "slices/s", "cores": orc.num_threads(), "kind": "port",
            "sample": "one 12-subset OS-SQS pass (Siddon FP+BP, FP32 OpenMP oracle) on slice 0, %.2f s, "
                      "scaled to k_conv=%d passes per slice" % (t, k_conv)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box of this build has a single MI355X; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from reconnet_amd.geometry import HU_WATER, WORKLOADS
    from reconnet_amd.phantom import add_poisson_noise, phantom_batch
    from reconnet_amd.sqs import OsSqsTV
    from reconnet_amd.torch_ops import FanPlan

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    geo = WORKLOADS["fan512"]
    ns = args.slices
    log("workload: %d slices %dx%d, %d views, %d bins" % (ns, geo.nx, geo.ny, geo.nview, geo.nu))

    truth = phantom_batch(range(ns), geo.nx, geo.ny, geo.dx, geo.dy, device=dev)
    plan = FanPlan(geo)
    clean = plan.forward(truth.contiguous(), "siddon")
    noisy, w = add_poisson_noise(clean, args.i0, seed=2024)
    mask = torch.from_numpy(geo.fov_mask()).to(dev)
    npix = float(mask.sum())

    def rmse_hu(x):
        d = (x - truth) * mask
        return (1000.0 / HU_WATER) * torch.sqrt((d * d).sum(dim=(1, 2)) / npix)

    eng = OsSqsTV(geo, args.subsets, ns)
    eng.set_data(noisy, w)
    eng.compute_norm()
    zeros = torch.zeros((ns, geo.nx, geo.ny), dtype=torch.float32, device=dev)
    sqs_kw = dict(beta=args.beta, eps=args.delta ** 2, lb=0.0)

    # ---- calibration: plain SQS-TV to k_max, convergence iteration ----
    eng.set_image(zeros)
    rm, upd = [], []
    t0 = time.perf_counter()
    for k in range(args.kmax):
        dd, xx = eng.run_pass(track=True, **sqs_kw)
        upd.append(float(np.sqrt(dd.sum() / max(xx.sum(), 1e-30))))
        rm.append(rmse_hu(eng.image()).cpu().numpy())
        if (k + 1) % 20 == 0:
            log("calibration pass %d: mean RMSE %.2f HU, update %.2e" % (k + 1, rm[-1].mean(), upd[-1]))
    torch.cuda.synchronize()
    log("calibration %.1f s" % (time.perf_counter() - t0))
    rm = np.array(rm)
    r_final = rm[-1].mean()
    k_conv = args.kmax
    for k in range(args.kmax):
        if rm[k].mean() <= 1.01 * r_final and upd[k] < 1e-4:
            k_conv = k + 1
            break
    r_star = float(rm[k_conv - 1].mean())
    log("k_conv = %d, R* = %.3f HU (RMSE at k_max %.3f HU)" % (k_conv, r_star, r_final))

    # ---- timed steps: k_conv passes per reconstruction ----
    def step():
        eng.set_image(zeros)
        for _ in range(k_conv):
            eng.run_pass(**sqs_kw)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    eng.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        log("step %d done" % (i + 1))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kt = eng.kernel_times()
    eng.set_profiling(False)
    rec_rmse = float(rmse_hu(eng.image()).mean())
    value = ns * args.steps / elapsed
    ms_per_step = 1000.0 * elapsed / args.steps
    log("plain SQS-TV: %.2f slices/s, %.1f ms per step, RMSE %.3f HU" % (value, ms_per_step, rec_rmse))

    # ---- roofline of the dominant kernel ----
    npers = eng.npers
    bytes_proj = 4.0 * (geo.nx * geo.ny + geo.nu * npers) * ns
    bytes_upd = 4.0 * 5 * geo.nx * geo.ny * ns
    kernels = {
        "siddon_fan_bp": (kt["bp"], bytes_proj),
        "siddon_fan_fp_residual": (kt["fp"], bytes_proj),
        "sqs_update": (kt["update"], bytes_upd),
    }
    name = max(kernels, key=lambda k: kernels[k][0])
    tot_ms, nbytes = kernels[name]
    avg_ms = tot_ms / max(kt["steps"], 1)
    achieved = nbytes / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": None, "kernel": name,
                "avg_launch_ms": avg_ms, "bytes_per_launch": nbytes}
    per_kernel = {k: {"avg_ms": v[0] / max(kt["steps"], 1), "bytes": v[1],
                      "gbs": v[1] / (v[0] / max(kt["steps"], 1) * 1e-3) / 1e9} for k, v in kernels.items()}

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (FP32 oracle, slice 0) ...")
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(geo, truth.cpu().numpy(), noisy.cpu().numpy(), w.cpu().numpy(), k_conv, args.subsets,
                           threads)
        log("CPU baseline: %.5f slices/s on %d threads" % (cpu["value"], cpu["cores"]))

    line = {
        "metric": "slices/s to plain SQS-TV RMSE (HU), learned-unrolled vs plain, 512² fan",
        "value": value,
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipse phantoms, Poisson noise I0=%g)" % args.i0,
        "config": {"workload": "plain OS-SQS-TV to convergence, 32 fan slices 512x512 @0.7mm, 1152 views, "
                               "736 bins, 12 subsets", "slices": ns, "k_conv": k_conv, "rmse_hu": rec_rmse,
                   "r_star_hu": r_star, "beta": args.beta, "delta": args.delta, "method": "plain_sqs_tv"},
        "roofline": roofline,
        "kernels": per_kernel,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
