#!/usr/bin/env python
"""Benchmark of the ReconNet reconstruction path on one MI355X.

Metric: slices/s to reach the plain SQS-TV RMSE (HU) on 512^2 fan-beam slices,
learned unrolled OS-SQS against plain SQS-TV on the same GPU.

Workload: 32 fan-beam slices of 512^2 at 0.7 mm, 1152 views over 2 pi, 736
bins of 1.0 mm, SID 595 mm, SDD 1085.6 mm, 12 ordered subsets, Poisson noise.
Synthetic data: seeded random-ellipse phantoms projected by the library's
Siddon projector; evaluation seeds 100000 .. 100199 (held out from training).

Both methods start from the FBP image (Hann-windowed ramp filter, FBP-weighted
DD backprojection); the FBP is part of every timed step.

1. Plain SQS-TV (beta = 1e-4, delta = 1e-3) runs k_max passes on the 200
   evaluation slices (untimed).  k_conv is the first iteration whose mean RMSE
   is within 1 % of the mean RMSE after k_max passes and whose relative update
   ||dx||/||x|| is below 1e-4; R* = mean RMSE(k_conv).
2. A timed plain step is FBP + k_conv passes over the 32-slice batch (the
   passes replayed from one HIP graph).
3. With trained weights (reconnet_amd/weights/unrolled_fan512.npz, used at the
   dose they were trained for) a timed step is FBP + the 10-iteration learned
   reconstruction of the batch (10 OS-SQS data passes + 10 CNN priors on BF16
   MFMA, replayed from one HIP graph); value = its slices/s, valid when its
   mean RMSE over the 200 evaluation slices is <= R*.  Without weights the
   headline step is the plain one.

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


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (lines "traffic <kernel> <bytes> ..."), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_fan512.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and parts[0] == "traffic" and parts[1] == kernel:
                return float(parts[2])
    return None


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--slices", type=int, default=32)
    ap.add_argument("--kmax", type=int, default=200)
    ap.add_argument("--eval", type=int, default=200, help="held-out slices for k_conv, R* and the learned RMSE")
    ap.add_argument("--no-graph", action="store_true", help="launch the timed passes eagerly")
    ap.add_argument("--probe", action="store_true",
                    help="profiling probe: FBP + 2 eager OS-SQS passes over the batch, a 256 MiB elementwise "
                         "copy (the reference kernel for counter scaling) and one CNN prior; no timing line")
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
    from reconnet_amd.unrolled import WEIGHTS_PATH, UnrolledGraph, UnrolledOSSQS, load_nets

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

    from reconnet_amd.metrics import convergence_pass
    from reconnet_amd.sqs import fbp_initial_image

    plan = FanPlan(geo)
    mask = torch.from_numpy(geo.fov_mask()).to(dev).float()
    npix = float(mask.sum())
    sqs_kw = dict(beta=args.beta, eps=args.delta ** 2, lb=0.0)

    def rmse_hu(x, truth):
        d = (x - truth) * mask
        return (1000.0 / HU_WATER) * torch.sqrt((d * d).sum(dim=(1, 2)) / npix)

    def simulate(seeds):
        truth = phantom_batch(seeds, geo.nx, geo.ny, geo.dx, geo.dy, device=dev).contiguous()
        noisy, w = add_poisson_noise(plan.forward(truth, "siddon"), i0, seed=2024 + seeds[0])
        return truth, noisy, w

    if args.probe:
        truth, noisy, w = simulate(list(range(100000, 100000 + ns)))
        eng = OsSqsTV(geo, args.subsets, ns)
        eng.set_data(noisy, w)
        eng.compute_norm()
        eng.set_image(fbp_initial_image(plan, noisy))
        for _ in range(2):
            eng.run_pass(**sqs_kw)
        # reference kernel for counter scaling: an elementwise kernel that reads
        # 256 MiB and writes 256 MiB (a plain copy_ would go to the DMA blit path)
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        if nets is not None:
            xx = eng.image()
            for net in nets:
                xx = net.forward(xx)
        else:
            from reconnet_amd.cnn import PriorCNN, init_params
            PriorCNN(init_params(seed=0, device=dev), scale=50.0).forward(eng.image())
        torch.cuda.synchronize()
        log("probe done: 2 passes over %d slices, CNN prior, one elementwise copy of %d bytes" % (ns, src.numel() * 4))
        return

    # ---- plain SQS-TV to convergence on the evaluation slices (untimed) ----
    ne = max(args.eval, ns)
    etruth, enoisy, ew = simulate(list(range(100000, 100000 + ne)))
    eeng = OsSqsTV(geo, args.subsets, ne)
    eeng.set_data(enoisy, ew)
    eeng.compute_norm()
    ex0 = fbp_initial_image(plan, enoisy)
    fbp_rmse = float(rmse_hu(ex0, etruth).mean())
    eeng.set_image(ex0)
    rm, upd = [], []
    t0 = time.perf_counter()
    for k in range(args.kmax):
        dd, xx = eeng.run_pass(track=True, **sqs_kw)
        upd.append(float(np.sqrt(dd.sum() / max(xx.sum(), 1e-30))))
        rm.append(rmse_hu(eeng.image(), etruth).cpu().numpy())
        if (k + 1) % 20 == 0:
            log("plain pass %d: mean RMSE %.2f HU, update %.2e" % (k + 1, rm[-1].mean(), upd[-1]))
    torch.cuda.synchronize()
    log("convergence run on %d slices: %.1f s (FBP start %.2f HU)" % (ne, time.perf_counter() - t0, fbp_rmse))
    rm = np.array(rm)
    r_final = float(rm[-1].mean())
    k_conv = convergence_pass(rm.mean(axis=1), upd)
    k_conv += k_conv % 2 * (args.subsets % 2)     # pass graphs need an even subset-step count
    r_star = float(rm[min(k_conv, args.kmax) - 1].mean())
    r_best = float(rm.mean(axis=1).min())
    log("k_conv = %d, R* = %.3f HU (k_max %.3f HU, best %.3f HU at pass %d)"
        % (k_conv, r_star, r_final, r_best, int(rm.mean(axis=1).argmin()) + 1))

    learned_eval = None
    if nets is not None:
        model = UnrolledOSSQS(nets, lb=0.0)
        xl = model.reconstruct(eeng, ex0, chunk=32)
        l_all = rmse_hu(xl, etruth).cpu().numpy()
        learned_eval = float(l_all.mean())
        log("learned on %d evaluation slices: mean RMSE %.3f HU (R* %.3f)" % (ne, learned_eval, r_star))
        del xl
    del eeng, enoisy, ew, ex0, etruth

    # ---- timed workload: the first 32 evaluation slices ----
    truth, noisy, w = simulate(list(range(100000, 100000 + ns)))
    eng = OsSqsTV(geo, args.subsets, ns)
    eng.set_data(noisy, w)
    eng.compute_norm()
    x0 = torch.empty((ns, geo.nx, geo.ny), dtype=torch.float32, device=dev)

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

    fbp_initial_image(plan, noisy, out=x0)
    eng.set_image(x0)
    pgraph = None if args.no_graph else eng.capture_passes(k_conv, **sqs_kw)

    def plain_step():
        fbp_initial_image(plan, noisy, out=x0)
        eng.set_image(x0)
        if pgraph is not None:
            pgraph.replay()
        else:
            for _ in range(k_conv):
                eng.run_pass(**sqs_kw)

    plain_steps = args.steps if nets is None else 1
    t_plain = timed(plain_step, plain_steps, args.warmup if nets is None else 0)
    plain_sps = ns * plain_steps / t_plain
    plain_rmse = float(rmse_hu(eng.image(), truth).mean())
    log("plain SQS-TV: %.2f slices/s (FBP + %d passes), batch RMSE %.3f HU" % (plain_sps, k_conv, plain_rmse))
    del pgraph

    conv_ms = None
    learned = None
    if nets is not None:
        ugraph = None if args.no_graph else UnrolledGraph(model, eng, x0)

        def learned_step():
            fbp_initial_image(plan, noisy, out=x0)
            if ugraph is not None:
                return ugraph.replay(x0)
            return model.reconstruct(eng, x0)

        elapsed = timed(learned_step, args.steps, args.warmup)
        x = learned_step()
        l_rmse = float(rmse_hu(x, truth).mean())
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
        learned = {"slices_per_s": value, "batch_rmse_hu": l_rmse, "eval_rmse_hu": learned_eval,
                   "iterations": len(nets), "reaches_r_star": bool(learned_eval <= r_star),
                   "speedup_vs_plain": value / plain_sps, "cnn_ms_per_application": conv_ms}
        log("learned: %.2f slices/s, eval RMSE %.3f HU (R* %.3f), %.1fx plain" % (value, learned_eval, r_star,
                                                                               value / plain_sps))
        method = "learned_unrolled_%d" % len(nets)
        passes = len(nets)
        del ugraph
    else:
        value = plain_sps
        ms_per_step = 1000.0 * t_plain / plain_steps
        method = "plain_sqs_tv"
        passes = k_conv

    # per-kernel HIP-event times of eager passes on the library stream
    eng.set_image(x0)
    eng.set_profiling(True)
    for _ in range(3):
        eng.run_pass(**sqs_kw)
    kt = eng.kernel_times()
    eng.set_profiling(False)

    # ---- roofline of the dominant kernel (HIP events on the SQS stream) ----
    npers = eng.npers
    bytes_proj = 4.0 * (geo.nx * geo.ny + geo.nu * npers) * ns
    # update (x, g, D read, x written) + interleaved copies (x read, two 4-slice layouts written)
    bytes_upd = 4.0 * 7 * geo.nx * geo.ny * ns
    steps_k = max(kt["steps"], 1)
    kernels = {"siddon_fan_fp_residual": (kt["fp"], bytes_proj)}
    if kt["update"] < 0.1 * kt["bp"]:
        # the update ran in the backprojection's epilogue: residual, x and D read,
        # the new image and its two interleaved copies written
        bytes_bpu = 4.0 * (geo.nu * npers + 5 * geo.nx * geo.ny) * ns
        kernels["siddon_fan_bp_sqs_update"] = (kt["bp"] + kt["update"], bytes_bpu)
    else:
        kernels["siddon_fan_bp"] = (kt["bp"], bytes_proj)
        kernels["sqs_update"] = (kt["update"], bytes_upd)
    name = max(kernels, key=lambda k: kernels[k][0])
    tot_ms, nbytes = kernels[name]
    avg_ms = tot_ms / steps_k
    achieved = nbytes / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": profiled_traffic("k_" + name), "kernel": name,
                "avg_launch_ms": avg_ms, "bytes_per_launch": nbytes, "launches": steps_k}
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
                   "slices": ns, "method": method, "i0": i0, "init": "fbp_hann", "eval_slices": ne,
                   "k_max": args.kmax, "k_conv": k_conv, "r_star_hu": r_star, "fbp_rmse_hu": fbp_rmse,
                   "plain_slices_per_s": plain_sps, "plain_batch_rmse_hu": plain_rmse, "learned": learned,
                   "beta": args.beta, "delta": args.delta, "hip_graph": not args.no_graph},
        "roofline": roofline,
        "kernels": per_kernel,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
