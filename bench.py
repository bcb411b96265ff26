#!/usr/bin/env python
"""Benchmark of the ReconNet reconstruction path on one MI355X.

Metric: slices/s to reach the plain SQS-TV RMSE (HU) on 512^2 fan-beam slices,
learned unrolled OS-SQS against plain SQS-TV on the same GPU.

Workload: 32 fan-beam slices of 512^2 at 0.7 mm, 1152 views over 2 pi, 736
bins of 1.0 mm, SID 595 mm, SDD 1085.6 mm, 12 ordered subsets, Poisson noise.
Synthetic data: seeded random-ellipse phantoms (held-out seeds 100000+) projected
by the library's Siddon projector.

1. Plain SQS-TV (beta = 1e-4, delta = 1e-3) is run to convergence, untimed:
   k_conv is the first iteration whose RMSE is within 1 % of the RMSE after
   k_max iterations and whose relative update ||dx||/||x|| is below 1e-4;
   R* = RMSE(k_conv).  Its throughput, 32 / t(k_conv passes), is then timed.
2. With trained weights (reconnet_amd/weights/unrolled_fan512.npz) a timed step
   is one 10-iteration learned reconstruction of the batch (10 OS-SQS data
   passes + 10 CNN priors on BF16 MFMA), at the noise level it was trained for;
   value = its slices/s, valid when its RMSE <= R*.  Without weights the step is
   the plain reconstruction and value = plain slices/s.

Inputs are resident in HBM before timing.  Prints one JSON line.
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
PEAK_BF16_TFLOPS = 2500.0


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
    ap.add_argument("--i0", type=float, default=None, help="photons per ray (default: the weights' training dose, "
                                                           "else 1e5)")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--plain", action="store_true", help="time plain SQS-TV even if weights exist")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def cpu_baseline(geo, noisy_np, w_np, k_conv, nsub, threads, passes_per_slice, what):
    """One OS-SQS data pass of the FP32 CPU oracle on slice 0 (OpenMP), scaled
    to the passes per slice of the timed method."""
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
    return {"value": 1.0 / (passes_per_slice * t), "unit": "slices/s", "cores": orc.num_threads(), "kind": "port",
            "sample": "one 12-subset OS-SQS data pass (Siddon FP+BP, FP32 OpenMP oracle) on slice 0 took %.2f s; "
                      "scaled to %d passes per slice (%s); prior steps not included" % (t, passes_per_slice, what)}


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
    from reconnet_amd.unrolled import WEIGHTS_PATH, UnrolledOSSQS, load_nets

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    geo = WORKLOADS["fan512"]
    ns = args.slices

    wpath = args.weights or WEIGHTS_PATH
    nets = None
    i0 = args.i0
    if os.path.exists(wpath) and not args.plain:
        nets = load_nets(wpath)
        z = np.load(wpath, allow_pickle=False)
        if i0 is None and "meta_i0" in z.files:
            i0 = float(z["meta_i0"])
    if i0 is None:
        i0 = 1e5
    log("workload: %d slices %dx%d, %d views, %d bins, I0 %g, learned=%s"
        % (ns, geo.nx, geo.ny, geo.nview, geo.nu, i0, nets is not None))

    seeds = list(range(100000, 100000 + ns))
    truth = phantom_batch(seeds, geo.nx, geo.ny, geo.dx, geo.dy, device=dev).contiguous()
    plan = FanPlan(geo)
    clean = plan.forward(truth, "siddon")
    noisy, w = add_poisson_noise(clean, i0, seed=2024)
    del clean
    mask = torch.from_numpy(geo.fov_mask()).to(dev).float()
    npix = float(mask.sum())

    def rmse_hu(x):
        d = (x - truth) * mask
        return (1000.0 / HU_WATER) * torch.sqrt((d * d).sum(dim=(1, 2)) / npix)

    eng = OsSqsTV(geo, args.subsets, ns)
    eng.set_data(noisy, w)
    eng.compute_norm()
    zeros = torch.zeros((ns, geo.nx, geo.ny), dtype=torch.float32, device=dev)
    sqs_kw = dict(beta=args.beta, eps=args.delta ** 2, lb=0.0)

    # ---- plain SQS-TV to convergence (untimed) ----
    eng.set_image(zeros)
    rm, upd = [], []
    t0 = time.perf_counter()
    for k in range(args.kmax):
        dd, xx = eng.run_pass(track=True, **sqs_kw)
        upd.append(float(np.sqrt(dd.sum() / max(xx.sum(), 1e-30))))
        rm.append(rmse_hu(eng.image()).cpu().numpy())
        if (k + 1) % 20 == 0:
            log("plain pass %d: mean RMSE %.2f HU, update %.2e" % (k + 1, rm[-1].mean(), upd[-1]))
    torch.cuda.synchronize()
    log("convergence run %.1f s" % (time.perf_counter() - t0))
    rm = np.array(rm)
    r_final = float(rm[-1].mean())
    k_conv = args.kmax
    for k in range(args.kmax):
        if rm[k].mean() <= 1.01 * r_final and upd[k] < 1e-4:
            k_conv = k + 1
            break
    r_star = float(rm[k_conv - 1].mean())
    r_best = float(rm.mean(axis=1).min())
    log("k_conv = %d, R* = %.3f HU (k_max %.3f HU, best %.3f HU)" % (k_conv, r_star, r_final, r_best))

    def plain_step():
        eng.set_image(zeros)
        for _ in range(k_conv):
            eng.run_pass(**sqs_kw)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for i in range(steps):
            fn()
            log("step %d done" % (i + 1))
        torch.cuda.synchronize()
        return time.perf_counter() - t

    # plain throughput (1 step; the headline step when no weights are present)
    eng.set_profiling(True)
    plain_steps = args.steps if nets is None else 1
    t_plain = timed(plain_step, plain_steps, args.warmup if nets is None else 0)
    kt = eng.kernel_times()
    eng.set_profiling(False)
    plain_sps = ns * plain_steps / t_plain
    plain_rmse = float(rmse_hu(eng.image()).mean())
    log("plain SQS-TV: %.2f slices/s (%d passes), RMSE %.3f HU" % (plain_sps, k_conv, plain_rmse))

    conv_ms = None
    learned = None
    if nets is not None:
        model = UnrolledOSSQS(nets, lb=0.0)

        def learned_step():
            model.reconstruct(eng, zeros)

        eng.set_profiling(True)
        elapsed = timed(learned_step, args.steps, args.warmup)
        kt = eng.kernel_times()
        eng.set_profiling(False)
        x = eng.image()
        l_rmse = float(rmse_hu(x).mean())
        value = ns * args.steps / elapsed
        ms_per_step = 1000.0 * elapsed / args.steps
        # CNN cost (events on the current stream around one prior application)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        xx = x.clone()
        e0.record()
        for net in nets:
            xx = net.forward(xx)
        e1.record()
        torch.cuda.synchronize()
        conv_ms = e0.elapsed_time(e1) / len(nets)
        learned = {"slices_per_s": value, "rmse_hu": l_rmse, "iterations": len(nets),
                   "reaches_r_star": bool(l_rmse <= r_star), "speedup_vs_plain": value / plain_sps,
                   "cnn_ms_per_application": conv_ms}
        log("learned: %.2f slices/s, RMSE %.3f HU (R* %.3f), %.1fx plain" % (value, l_rmse, r_star,
                                                                          value / plain_sps))
        method = "learned_unrolled_%d" % len(nets)
        passes = len(nets)
    else:
        value = plain_sps
        ms_per_step = 1000.0 * t_plain / plain_steps
        method = "plain_sqs_tv"
        passes = k_conv

    # ---- roofline of the dominant kernel (HIP events on the SQS stream) ----
    npers = eng.npers
    bytes_proj = 4.0 * (geo.nx * geo.ny + geo.nu * npers) * ns
    bytes_upd = 4.0 * 5 * geo.nx * geo.ny * ns
    steps_k = max(kt["steps"], 1)
    kernels = {
        "siddon_fan_bp": (kt["bp"], bytes_proj),
        "siddon_fan_fp_residual": (kt["fp"], bytes_proj),
        "sqs_update": (kt["update"], bytes_upd),
    }
    name = max(kernels, key=lambda k: kernels[k][0])
    tot_ms, nbytes = kernels[name]
    avg_ms = tot_ms / steps_k
    achieved = nbytes / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": None, "kernel": name, "avg_launch_ms": avg_ms,
                "bytes_per_launch": nbytes, "launches": steps_k}
    per_kernel = {k: {"avg_ms": v[0] / steps_k, "bytes_per_launch": v[1],
                      "gbs": v[1] / (v[0] / steps_k * 1e-3) / 1e9} for k, v in kernels.items()}
    if conv_ms is not None:
        flops = 2.0 * geo.nx * geo.ny * ns * 9 * (32 + 3 * 32 * 32 + 32)
        per_kernel["cnn_prior"] = {"avg_ms": conv_ms, "flops_per_application": flops,
                                   "tflops": flops / (conv_ms * 1e-3) / 1e12}

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (FP32 oracle, slice 0) ...")
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(geo, noisy.cpu().numpy(), w.cpu().numpy(), k_conv, args.subsets, threads, passes, method)
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
        "dtype": "f32" if nets is None else "f32 (projectors), bf16 MFMA / f32 accumulate (prior)",
        "data": "synthetic (seeded ellipse phantoms, Poisson noise I0=%g)" % i0,
        "config": {"workload": "32 fan slices 512x512 @0.7mm, 1152 views, 736 bins, 12 subsets; " + method,
                   "slices": ns, "method": method, "i0": i0, "k_conv": k_conv, "r_star_hu": r_star,
                   "plain_slices_per_s": plain_sps, "plain_rmse_hu": plain_rmse, "learned": learned,
                   "beta": args.beta, "delta": args.delta},
        "roofline": roofline,
        "kernels": per_kernel,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
