#!/usr/bin/env python
"""Benchmark of the ReconNet reconstruction path on one MI355X.

Metric: slices/s to reach the plain SQS-TV RMSE (HU) on 512^2 fan-beam slices,
learned unrolled OS-SQS against plain SQS-TV on the same GPU.

Workload: 32 fan-beam slices of 512^2 at 0.7 mm, 1152 views over 2 pi, 736
bins of 1.0 mm, SID 595 mm, SDD 1085.6 mm, 12 ordered subsets, Poisson noise
at the dose the learned weights were trained for (I0 = 2.5e4; else 1e5),
statistical weights counts / I0.  Synthetic data: seeded random-ellipse
phantoms projected by the library's Siddon projector; evaluation seeds
100000 .. 100199 (held out from training).

Both methods start from the FBP image (Hann-windowed ramp filter, FBP-weighted
DD backprojection); the FBP is part of every timed step.

The plain judge.  With these weights a TV weight of 1e-4 is ~1e-6 of the data
curvature: such a run is unregularised OS-SQS, whose RMSE falls for ~20
passes and then rises as it fits the noise, so it has no converged RMSE to be
judged by.  The judge is therefore plain SQS-TV at the TV weight whose
converged RMSE on these slices is lowest (beta = 0.01, delta = 1e-3, from a
sweep of beta = 1e-3 .. 0.3 run with --sweep-beta): its RMSE decreases
monotonically to a plateau.  It runs k_max = 2000 passes on the 200
evaluation slices (untimed); k_conv is the first pass whose mean RMSE is
within 1 % of the mean RMSE after k_max passes and whose relative update
||dx||/||x|| is below 1e-4; R* = mean RMSE(k_conv).  The run at the nominal
beta = 1e-4 (k_max = 200) is reported beside it.

Timed steps (inputs resident in HBM, warm-up steps untimed):
  plain   = FBP + k_conv SQS-TV passes over the 32-slice batch (HIP graphs of
            passes replayed);
  learned = FBP + the 10-iteration learned reconstruction (10 OS-SQS data
            passes + 10 CNN priors on BF16 MFMA, one HIP graph).
value = the learned method's slices/s when its mean RMSE over the 200
evaluation slices is <= R*, else the plain judge's slices/s.

Secondary measurements: per-kernel HIP-event times inside the OS-SQS passes
(roofline of the dominant kernel), per-call GPU times of the 256^2 fan
operators, the 256^2 x 128 cone pair and the 16-slice uneven batch, and the
FP32 OpenMP CPU oracle timed on the host cores for the same work.
Prints one JSON line.
"""

import argparse
import json
import math
import os
import subprocess
import sys
import time

# CPU baseline threads stay on the cores they start on (OpenMP reads this when
# its runtime starts, so it is set before torch loads one)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0
# helical workload (3 turns, pitch 1): raw flat-detector views and the rebinned parallel-cone rows
_ZROT = 64 * 1.0 * 595.0 / 1085.6
HELICAL = {"nu": 512, "nview": 2160, "nv": 64,
           "rebin": dict(du_f=1.2, dv_f=1.0, beta0=0.0, dbeta=math.pi / 360, nu_p=512, nview_p=2160, nv_p=64,
                         du_p=1.2 * 595.0 / 1085.6, dv_p=1.0, theta0=0.0, dtheta=math.pi / 360, dsd=1085.6,
                         dso=595.0),
           "bp": dict(nview_per_pi=360, theta0=0.0, vol_z0=-11.0, zrot=_ZROT, nx=256, ny=256, nz=128, dx=1.0,
                      dy=1.0, dz=1.0, du=1.2 * 595.0 / 1085.6, dv=1.0, dsd=1085.6, dso=595.0)}

JUDGE_BETA = 0.01
SPEC_BETA = 1e-4


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
    ap.add_argument("--eval", type=int, default=200, help="held-out slices for k_conv, R* and the learned RMSE")
    ap.add_argument("--subsets", type=int, default=12)
    ap.add_argument("--beta", type=float, default=JUDGE_BETA, help="TV weight of the plain judge")
    ap.add_argument("--kmax", type=int, default=2000, help="passes of the judge's convergence run")
    ap.add_argument("--spec-beta", type=float, default=SPEC_BETA, help="nominal TV weight, reported beside")
    ap.add_argument("--spec-kmax", type=int, default=200)
    ap.add_argument("--delta", type=float, default=1e-3)
    ap.add_argument("--i0", type=float, default=None, help="photons per ray (default: the weights' training dose, "
                                                           "else 1e5)")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--plain", action="store_true", help="time plain SQS-TV only, even if weights exist")
    ap.add_argument("--no-graph", action="store_true", help="launch the timed passes eagerly")
    ap.add_argument("--probe", action="store_true",
                    help="profiling probe: FBP + 2 eager OS-SQS passes over the batch, a 256 MiB elementwise "
                         "copy (the reference kernel for counter scaling) and one CNN prior; no timing line")
    ap.add_argument("--probe-configs", action="store_true",
                    help="profiling probe of the 256^2 fan operators, the cone pair and the uneven batch")
    ap.add_argument("--sweep-beta", default=None,
                    help="comma-separated TV weights: print the judge's convergence table for each on the "
                         "evaluation slices and exit")
    ap.add_argument("--no-configs", action="store_true", help="skip the secondary GPU/CPU measurements")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


# ----------------------------------------------------------------------------------------------
# host description and CPU timing helpers

def host_info():
    info = {"nproc": os.cpu_count(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
        pairs = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
        cores = {ln for ln in pairs.splitlines() if ln and not ln.startswith("#")}
        info["physical_cores"] = len(cores)
    except Exception as e:      # lscpu missing: leave the fields out
        info["lscpu_error"] = str(e)
    return info


def cpu_threads(args, info):
    """All physical cores, capped by the host share this process is given
    (OMP_NUM_THREADS, set per GPU on a shared host)."""
    if args.cpu_threads:
        return args.cpu_threads
    n = info.get("physical_cores") or os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def median_time(fn, reps=3, warm=1):
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


# ----------------------------------------------------------------------------------------------

def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box of this build has a single MI355X; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from reconnet_amd.geometry import HU_WATER, WORKLOADS
    from reconnet_amd.metrics import convergence_pass
    from reconnet_amd.phantom import add_poisson_noise, phantom_batch
    from reconnet_amd.sqs import OsSqsTV, fbp_initial_image
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

    plan = FanPlan(geo)
    mask = torch.from_numpy(geo.fov_mask()).to(dev).float()
    npix = float(mask.sum())

    def sqs_kw(beta):
        return dict(beta=beta, eps=args.delta ** 2, lb=0.0)

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
            eng.run_pass(**sqs_kw(args.beta))
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
    if args.probe_configs:
        config_gpu_timings(torch, dev, reps=3)
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        torch.mul(src, 1.0, out=torch.empty_like(src))
        torch.cuda.synchronize()
        log("config probe done")
        return

    # ---- plain SQS-TV convergence runs on the evaluation slices (untimed) ----
    ne = max(args.eval, ns)
    etruth, enoisy, ew = simulate(list(range(100000, 100000 + ne)))
    eeng = OsSqsTV(geo, args.subsets, ne)
    eeng.set_data(enoisy, ew)
    eeng.compute_norm()
    ex0 = fbp_initial_image(plan, enoisy)
    fbp_rmse = float(rmse_hu(ex0, etruth).mean())

    def convergence_run(beta, kmax):
        eeng.set_image(ex0)
        rm, upd = [], []
        t0 = time.perf_counter()
        for k in range(kmax):
            dd, xx = eeng.run_pass(track=True, **sqs_kw(beta))
            upd.append(float(np.sqrt(dd.sum() / max(xx.sum(), 1e-30))))
            rm.append(float(rmse_hu(eeng.image(), etruth).mean()))
            if (k + 1) % 250 == 0 or k + 1 == kmax:
                log("beta %g pass %d: mean RMSE %.2f HU, update %.2e" % (beta, k + 1, rm[-1], upd[-1]))
        torch.cuda.synchronize()
        rm, upd = np.array(rm), np.array(upd)
        k_conv = convergence_pass(rm, upd)
        converged = bool(((rm <= 1.01 * rm[-1]) & (upd < 1e-4)).any())
        res = {"beta": beta, "delta": args.delta, "k_max": kmax, "k_conv": k_conv, "k_conv_converged": converged,
               "r_star_hu": float(rm[k_conv - 1]), "r_final_hu": float(rm[-1]), "r_best_hu": float(rm.min()),
               "k_best": int(rm.argmin()) + 1, "update_at_k_conv": float(upd[k_conv - 1]),
               "update_final": float(upd[-1]), "seconds": time.perf_counter() - t0}
        # how k_conv would move with a shorter convergence run
        res["k_conv_if_k_max"] = {str(m): convergence_pass(rm[:m], upd[:m]) for m in (kmax // 2, 3 * kmax // 4)}
        res["rmse_at"] = {str(k): float(rm[k - 1]) for k in (10, 20, 50, 100, 200, 500, 1000, 1500, 2000)
                          if k <= kmax}
        return res

    if args.sweep_beta:
        table = [convergence_run(float(b), args.kmax) for b in args.sweep_beta.split(",")]
        print(json.dumps({"sweep": table, "fbp_rmse_hu": fbp_rmse, "eval_slices": ne, "i0": i0}), flush=True)
        return

    judge = convergence_run(args.beta, args.kmax)
    spec = convergence_run(args.spec_beta, args.spec_kmax)
    k_conv = judge["k_conv"]
    r_star = judge["r_star_hu"]
    log("judge beta %g: k_conv %d (converged %s), R* %.3f HU, final %.3f HU; nominal beta %g: best %.3f HU at "
        "pass %d, %.3f HU at pass %d (FBP %.2f HU)"
        % (args.beta, k_conv, judge["k_conv_converged"], r_star, judge["r_final_hu"], args.spec_beta,
           spec["r_best_hu"], spec["k_best"], spec["r_final_hu"], args.spec_kmax, fbp_rmse))

    learned_eval = None
    model = None
    if nets is not None:
        model = UnrolledOSSQS(nets, lb=0.0)
        xl = model.reconstruct(eeng, ex0, chunk=32)
        learned_eval = float(rmse_hu(xl, etruth).mean())
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
    chunk = 10     # passes per HIP graph; a step replays it k // chunk times and runs the rest eagerly

    def plain_runner(beta, npass):
        g = None if args.no_graph or npass < chunk else eng.capture_passes(chunk, **sqs_kw(beta))

        def step():
            fbp_initial_image(plan, noisy, out=x0)
            eng.set_image(x0)
            done = 0
            if g is not None:
                for _ in range(npass // chunk):
                    g.replay()
                done = npass // chunk * chunk
            for _ in range(npass - done):
                eng.run_pass(**sqs_kw(beta))
        return step

    step = plain_runner(args.beta, k_conv)
    t_plain = timed(step, args.steps, args.warmup)
    plain_sps = ns * args.steps / t_plain
    plain_rmse = float(rmse_hu(eng.image(), truth).mean())
    log("plain SQS-TV (beta %g): %.2f slices/s (FBP + %d passes), batch RMSE %.3f HU"
        % (args.beta, plain_sps, k_conv, plain_rmse))
    judge.update({"plain_slices_per_s": plain_sps, "plain_batch_rmse_hu": plain_rmse,
                  "plain_ms_per_step": 1000.0 * t_plain / args.steps})
    # the nominal-beta run stopped at its best iterate
    step = plain_runner(args.spec_beta, spec["k_best"])
    t_spec = timed(step, args.steps, args.warmup)
    spec.update({"plain_slices_per_s_at_k_best": ns * args.steps / t_spec,
                 "plain_batch_rmse_hu_at_k_best": float(rmse_hu(eng.image(), truth).mean())})

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
        learned_sps = ns * args.steps / elapsed
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        xx = x.clone()
        e0.record()
        for net in nets:
            xx = net.forward(xx)
        e1.record()
        torch.cuda.synchronize()
        conv_ms = e0.elapsed_time(e1) / len(nets)
        reaches = bool(learned_eval <= r_star)
        learned = {"slices_per_s": learned_sps, "ms_per_step": 1000.0 * elapsed / args.steps,
                   "batch_rmse_hu": l_rmse, "eval_rmse_hu": learned_eval, "iterations": len(nets),
                   "reaches_r_star": reaches, "speedup_vs_plain": learned_sps / plain_sps,
                   "reaches_nominal_best": bool(learned_eval <= spec["r_best_hu"]),
                   "speedup_vs_nominal_best_plain": learned_sps / spec["plain_slices_per_s_at_k_best"],
                   "cnn_ms_per_application": conv_ms}
        log("learned: %.2f slices/s, eval RMSE %.3f HU (R* %.3f), %.1fx the plain judge"
            % (learned_sps, learned_eval, r_star, learned_sps / plain_sps))
        del ugraph
        if reaches:
            method, value, ms_per_step, passes = "learned_unrolled_%d" % len(nets), learned_sps, \
                learned["ms_per_step"], len(nets)
        else:
            method, value, ms_per_step, passes = "plain_sqs_tv", plain_sps, judge["plain_ms_per_step"], k_conv
    else:
        method, value, ms_per_step, passes = "plain_sqs_tv", plain_sps, judge["plain_ms_per_step"], k_conv

    # ---- per-kernel HIP-event times of eager passes on the library stream ----
    eng.set_image(x0)
    beta_prof = 0.0 if method.startswith("learned") else args.beta
    eng.set_profiling(True)
    for _ in range(3):
        eng.run_pass(**sqs_kw(beta_prof))
    kt = eng.kernel_times()
    eng.set_profiling(False)

    npers = eng.npers
    bytes_proj = 4.0 * (geo.nx * geo.ny + geo.nu * npers) * ns
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
                                   "tflops": flops / (conv_ms * 1e-3) / 1e12,
                                   "mfma_frac": flops / (conv_ms * 1e-3) / 1e12 / PEAK_BF16_TFLOPS}

    configs = None
    if not args.no_configs:
        configs = config_gpu_timings(torch, dev, reps=5)

    info = host_info()
    cpu = None
    if not args.no_cpu_baseline:
        threads = cpu_threads(args, info)
        log("CPU baseline (FP32 OpenMP oracle, %d threads) ..." % threads)
        cpu = cpu_baseline(geo, noisy, w, args, threads, passes, method, k_conv, nets, info)
        if configs is not None:
            cpu_config_timings(configs, threads)
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
        "dtype": "f32" if not method.startswith("learned") else
                 "f32 (projectors), bf16 MFMA / f32 accumulate (prior)",
        "data": "synthetic (seeded ellipse phantoms, Poisson noise I0=%g, weights counts/I0)" % i0,
        "config": {"workload": "32 fan slices 512x512 @0.7mm, 1152 views, 736 bins, 12 subsets; " + method,
                   "slices": ns, "method": method, "i0": i0, "init": "fbp_hann", "eval_slices": ne,
                   "fbp_rmse_hu": fbp_rmse, "k_conv": k_conv, "r_star_hu": r_star,
                   "judge": judge, "nominal_beta_run": spec, "learned": learned,
                   "hip_graph": not args.no_graph},
        "roofline": roofline,
        "kernels": per_kernel,
        "configs": configs,
        "host": info,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# ----------------------------------------------------------------------------------------------
# secondary workloads: 256^2 fan operators, the cone pair, the uneven batch

def config_gpu_timings(torch, dev, reps=5):
    """Per-call GPU times (events on the current stream, which the library's
    device-pointer entry points launch on) with algorithmic bytes 4 (unique
    input + output elements) per call."""
    from reconnet_amd.geometry import CONE_WORKLOAD, WORKLOADS, circular_cone_vectors, uneven_fan_batch
    from reconnet_amd.phantom import phantom, phantom3d
    from reconnet_amd.ragged import RaggedFanBatch
    from reconnet_amd.torch_ops import ConePlan, FanPlan

    def gpu_ms(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {}

    def rec(key, ms, nbytes, **extra):
        out[key] = dict(gpu_ms=ms, bytes_per_call=nbytes, gbs=nbytes / (ms * 1e-3) / 1e9, **extra)
        log("%-22s %8.3f ms  %7.1f GB/s" % (key, ms, out[key]["gbs"]))

    # fan 256^2, 360 views, 512 bins
    g = WORKLOADS["fan256"]
    p = FanPlan(g)
    x = torch.from_numpy(phantom(0, g.nx, g.ny, g.dx, g.dy)[None]).to(dev)
    y = p.forward(x, "dd")
    nb = 4.0 * (g.nx * g.ny + g.nu * g.nview)
    rec("fan256_dd_fp", gpu_ms(lambda: p.forward(x, "dd")), nb)
    rec("fan256_dd_bp", gpu_ms(lambda: p.backward(y, "dd")), nb)
    rec("fan256_dd_adjoint", gpu_ms(lambda: p.backward(y, "dd_adjoint")), nb)
    rec("fan256_pixel_bp", gpu_ms(lambda: p.backward(y, "pixel")), nb)
    rec("fan256_siddon_fp", gpu_ms(lambda: p.forward(x, "siddon")), nb)
    rec("fan256_siddon_bp", gpu_ms(lambda: p.backward(y, "siddon")), nb)
    rec("fan256_ramp_filter", gpu_ms(lambda: p.filter(y, 0)), 8.0 * g.nu * g.nview)
    rec("fan256_fbp", gpu_ms(lambda: p.fbp(y, 0)), 4.0 * (g.nx * g.ny + g.nu * g.nview))
    # the same operators on a batch of 32 slices
    xb = x.expand(32, -1, -1).contiguous()
    yb = p.forward(xb, "dd")
    rec("fan256x32_dd_fp", gpu_ms(lambda: p.forward(xb, "dd")), 32 * nb)
    rec("fan256x32_dd_bp", gpu_ms(lambda: p.backward(yb, "dd")), 32 * nb)
    rec("fan256x32_pixel_bp", gpu_ms(lambda: p.backward(yb, "pixel")), 32 * nb)
    rec("fan256x32_dd_adjoint", gpu_ms(lambda: p.backward(yb, "dd_adjoint")), 32 * nb)
    del p

    # cone 256^2 x 128, 720 views, 64 x 512
    c = CONE_WORKLOAD
    vecs = circular_cone_vectors(c["nview"], c["dso"], c["dsd"])
    cp = ConePlan(c["nx"], c["ny"], c["nz"], c["dx"], c["dy"], c["dz"], c["nu"], c["nv"], c["du"], c["dv"], *vecs)
    v = phantom3d(0, c["nx"], c["ny"], c["nz"], c["dx"], c["dy"], c["dz"], device=dev)[None, ..., None].contiguous()
    pr = cp.forward(v, "dd")
    nbc = 4.0 * (c["nx"] * c["ny"] * c["nz"] + c["nu"] * c["nview"] * c["nv"])
    rec("cone_dd_fp", gpu_ms(lambda: cp.forward(v, "dd")), nbc)
    rec("cone_dd_adjoint", gpu_ms(lambda: cp.backward(pr, "dd")), nbc)
    prs = cp.forward(v, "siddon")
    rec("cone_siddon_fp", gpu_ms(lambda: cp.forward(v, "siddon")), nbc)
    rec("cone_siddon_bp", gpu_ms(lambda: cp.backward(prs, "siddon")), nbc)
    del cp

    # uneven batch: 6 x 256^2/360, 6 x 384^2/720, 4 x 512^2/1152
    geos = uneven_fan_batch()
    rb = RaggedFanBatch(geos)
    imgs = [torch.from_numpy(phantom(s, q.nx, q.ny, q.dx, q.dy)).to(dev) for s, q in enumerate(geos)]
    flat = rb.flatten_images(imgs)
    sino = rb.forward(flat, "dd")
    nbr = 4.0 * (rb.n_image + rb.n_sino)
    rec("ragged16_dd_fp", gpu_ms(lambda: rb.forward(flat, "dd")), nbr)
    rec("ragged16_dd_adjoint", gpu_ms(lambda: rb.backward(sino, "dd")), nbr)
    rec("ragged16_siddon_fp", gpu_ms(lambda: rb.forward(flat, "siddon")), nbr)
    rec("ragged16_siddon_bp", gpu_ms(lambda: rb.backward(sino, "siddon")), nbr)
    del rb

    # helical 256^2 x 128 at 1 mm, 3 turns of 720 views at pitch 1 (SID 595, SDD 1085.6 mm):
    # flat 64 x 512 views (1.0 x 1.2 mm) -> rebinned parallel-cone rows -> weighted BP
    from reconnet_amd.helical import helical_backproject, rebin_fan_to_parallel

    hw = HELICAL
    fan = torch.rand((1, hw["nu"], hw["nview"], hw["nv"], 1), generator=torch.Generator(device=dev).manual_seed(5),
                     device=dev)
    par = rebin_fan_to_parallel(fan, **hw["rebin"])
    nbh = 4.0 * (2 * hw["nu"] * hw["nview"] * hw["nv"])
    rec("helical_rebin", gpu_ms(lambda: rebin_fan_to_parallel(fan, out=par, **hw["rebin"])), nbh)
    vol = helical_backproject(par, **hw["bp"])
    nbb = 4.0 * (hw["nu"] * hw["nview"] * hw["nv"] + 256 * 256 * 128)
    rec("helical_bp", gpu_ms(lambda: helical_backproject(par, out=vol, **hw["bp"])), nbb)
    out["helical_bp"]["voxel_views_per_s"] = 256 * 256 * 128 * hw["bp"]["nview_per_pi"] * 3 / (
        out["helical_bp"]["gpu_ms"] * 1e-3)
    return out


def cpu_config_timings(configs, threads):
    """The FP32 OpenMP oracle on the same work as config_gpu_timings (median
    of 3 after a warm-up; the cone pair on every 20th view, scaled by 20)."""
    import oracle as orc
    from reconnet_amd.geometry import CONE_WORKLOAD, WORKLOADS, circular_cone_vectors, uneven_fan_batch
    from reconnet_amd.phantom import phantom, phantom3d

    orc.set_threads(threads)
    g = WORKLOADS["fan256"]
    deg = g.angles()
    go = g.oracle_geo()
    x = phantom(0, g.nx, g.ny, g.dx, g.dy)[None, :, :, None, None]
    y = orc.dd_fan_fp(x, deg, go, g.nu, precision="f32").astype(np.float32)
    shp = (g.nx, g.ny, 1)
    t = {
        "fan256_dd_fp": median_time(lambda: orc.dd_fan_fp(x, deg, go, g.nu, precision="f32")),
        "fan256_dd_bp": median_time(lambda: orc.dd_fan_bp(y, deg, go, shp, precision="f32")),
        "fan256_pixel_bp": median_time(lambda: orc.pixel_fan_bp(y, deg, go, shp, precision="f32")),
        "fan256_siddon_fp": median_time(lambda: orc.siddon_fan_fp(x, deg, go, g.nu, precision="f32")),
        "fan256_siddon_bp": median_time(lambda: orc.siddon_fan_bp(y, deg, go, shp, precision="f32")),
        "fan256_ramp_filter": median_time(lambda: orc.ramp_filter(y, go, 0, precision="f32")),
        "fan256_fbp": median_time(lambda: orc.dd_fan_bp(orc.ramp_filter(y, go, 0, precision="f32"), deg, go, shp,
                                                        type_projector=1, precision="f32")),
    }
    t["fan256x32_dd_fp"] = 32 * t["fan256_dd_fp"]
    t["fan256x32_dd_bp"] = 32 * t["fan256_dd_bp"]
    t["fan256x32_pixel_bp"] = 32 * t["fan256_pixel_bp"]
    c = CONE_WORKLOAD
    sub = 20
    vecs = [a[::sub].copy() for a in circular_cone_vectors(c["nview"], c["dso"], c["dsd"])]
    cg = dict(dx=c["dx"], dy=c["dy"], dz=c["dz"], du=c["du"], dv=c["dv"])
    vol = phantom3d(0, c["nx"], c["ny"], c["nz"], c["dx"], c["dy"], c["dz"])[None, ..., None]
    pr = orc.dd_cone_fp(vol, *vecs, cg, c["nu"], c["nv"], precision="f32").astype(np.float32)
    t["cone_dd_fp"] = sub * median_time(lambda: orc.dd_cone_fp(vol, *vecs, cg, c["nu"], c["nv"], precision="f32"),
                                        reps=1, warm=0)
    t["cone_dd_adjoint"] = sub * median_time(
        lambda: orc.dd_cone_adj(pr, *vecs, cg, (c["nx"], c["ny"], c["nz"]), precision="f32"), reps=1, warm=0)
    # one slice of each size class, scaled by the class's slice count
    geos = uneven_fan_batch()
    tot = {"dd_fp": 0.0, "siddon_fp": 0.0, "siddon_bp": 0.0}
    for s, q in enumerate(geos):
        if s > 0 and geos[s - 1] == q:
            continue
        count = sum(1 for r in geos if r == q)
        xi = phantom(s, q.nx, q.ny, q.dx, q.dy)[None, :, :, None, None]
        d, qo = q.angles(), q.oracle_geo()
        yi = orc.siddon_fan_fp(xi, d, qo, q.nu, precision="f32").astype(np.float32)
        tot["dd_fp"] += count * median_time(lambda: orc.dd_fan_fp(xi, d, qo, q.nu, precision="f32"), reps=1, warm=0)
        tot["siddon_fp"] += count * median_time(lambda: orc.siddon_fan_fp(xi, d, qo, q.nu, precision="f32"),
                                                reps=1, warm=0)
        tot["siddon_bp"] += count * median_time(
            lambda: orc.siddon_fan_bp(yi, d, qo, (q.nx, q.ny, 1), precision="f32"), reps=1, warm=0)
    for k in ("dd_fp", "siddon_fp", "siddon_bp"):
        t["ragged16_" + k] = tot[k]
    # helical BP: 2 of the 128 slices (z = 29, 30 mm: whole pi-windows), FP64 oracle, scaled by 64
    hb = HELICAL["bp"]
    par = np.random.default_rng(5).uniform(0, 1, (1, HELICAL["nu"], HELICAL["nview"], HELICAL["nv"], 1))
    par = par.astype(np.float32)
    hg = dict(dx=hb["dx"], dy=hb["dy"], dz=hb["dz"], du=hb["du"], dv=hb["dv"], dsd=hb["dsd"], dso=hb["dso"])
    t["helical_bp"] = 64 * median_time(lambda: orc.helical_bp(par, hg, (hb["nx"], hb["ny"], 2), hb["nview_per_pi"],
                                                              hb["theta0"], hb["vol_z0"] + 40, hb["zrot"]),
                                       reps=1, warm=0)
    del par
    for key, sec in t.items():
        if key in configs:
            configs[key]["cpu_ms"] = 1000.0 * sec
            configs[key]["gpu_speedup"] = 1000.0 * sec / configs[key]["gpu_ms"]
    configs["cpu_notes"] = ("FP32 OpenMP oracle, %d threads, median of 3 after a warm-up; cone pair timed on "
                            "36 of 720 views and scaled by 20; uneven batch: one slice of each size class, scaled by "
                            "the class's slice count (the oracle's DD adjoint is not instantiated in FP32, so "
                            "ragged16_dd_adjoint has no CPU figure); helical BP: FP64 oracle on 2 of 128 slices, "
                            "scaled by 64"
                            % threads)


def cpu_baseline(geo, noisy, w, args, threads, passes, method, k_conv, nets, info):
    """The timed method's work on slice 0, FP32 OpenMP oracle on the host
    cores: the FBP start (ramp filter + DD FBP backprojection) and one
    12-subset OS-SQS data pass (Siddon FP + BP), plus, for the learned method,
    one CNN prior (the CPU network restatement, FP64 numpy) or, for plain
    SQS-TV, the TV terms of every subset; scaled to the method's passes."""
    import oracle as orc
    import oracle.cnn as ocnn
    from reconnet_amd.geometry import interleaved_order

    orc.set_threads(threads)
    order = interleaved_order(geo.nview, args.subsets)
    deg = geo.angles()[order]
    nz = noisy[0].cpu().numpy()
    b = np.ascontiguousarray(nz[:, order])[None, :, :, None, None]
    wt = np.ascontiguousarray(w[0].cpu().numpy()[:, order])[None, :, :, None, None]
    g = geo.oracle_geo()
    ones = np.ones((1, geo.nx, geo.ny, 1, 1), np.float32)
    fp1 = orc.siddon_fan_fp(ones, deg, g, geo.nu, precision="f32")
    D = orc.siddon_fan_bp((fp1 * wt).astype(np.float32), deg, g, (geo.nx, geo.ny, 1), precision="f32")
    sino = nz[None, :, :, None, None]

    def fbp():
        f = orc.ramp_filter(sino, g, 2, precision="f32")
        return orc.dd_fan_bp(f.astype(np.float32), geo.angles(), g, (geo.nx, geo.ny, 1), type_projector=1,
                             precision="f32")

    t0 = time.perf_counter()
    x0 = np.clip(fbp(), 0, None).astype(np.float32)
    t_fbp = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.os_sqs(x0, b, D, wt, deg, g, args.subsets, lb=0.0, precision="f32")
    t_pass = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.os_sqs(x0, b, D, wt, deg, g, args.subsets, lb=0.0, precision="f64")
    t_pass64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(args.subsets):
        orc.tv_sqs(x0[0, :, :, 0, 0], precision="f32")
    t_tv = time.perf_counter() - t0
    t_cnn = 0.0
    if nets is not None:
        params = _cpu_net_params(nets[0])
        t0 = time.perf_counter()
        ocnn.forward(x0[:, :, :, 0, 0], params)
        t_cnn = time.perf_counter() - t0
    if method.startswith("learned"):
        per_slice = t_fbp + passes * (t_pass + t_cnn)
        what = "FBP + %d x (OS-SQS data pass + CNN prior)" % passes
    else:
        per_slice = t_fbp + passes * (t_pass + t_tv)
        what = "FBP + %d x (OS-SQS pass + 12 TV surrogates)" % passes
    plain_per_slice = t_fbp + k_conv * (t_pass + t_tv)
    return {"value": 1.0 / per_slice, "unit": "slices/s", "cores": orc.num_threads(), "kind": "port",
            "precision": "f32",
            "sample": "slice 0 of the batch on the FP32 OpenMP oracle (%d threads, OMP_PROC_BIND=%s): FBP %.3f s, "
                      "one 12-subset OS-SQS data pass %.3f s, 12 TV surrogates %.3f s, one CNN prior %.3f s "
                      "(FP64 numpy restatement); scaled to %s"
                      % (orc.num_threads(), os.environ.get("OMP_PROC_BIND"), t_fbp, t_pass, t_tv, t_cnn, what),
            "f64_pass_s": t_pass64, "fbp_s": t_fbp, "pass_s": t_pass, "tv_s": t_tv, "cnn_s": t_cnn,
            "plain_judge_slices_per_s": 1.0 / plain_per_slice,
            "cpu_model": info.get("cpu_model"), "physical_cores": info.get("physical_cores")}


def _cpu_net_params(net):
    """The network's weights in the CPU restatement's layout (w1 [9][32],
    w2..w4 [9][32 out][32 in], w5 [9][32]), FP64, with the input / output
    scale s folded as the GPU network folds it (w1 s, w5 / s, b5 / s)."""
    a = {k: np.asarray(v, np.float64) for k, v in net.to_numpy().items()}
    s = float(a.pop("scale"))
    a["w1"] = a["w1"] * s
    a["w5"] = a["w5"] / s
    a["b5"] = a["b5"] / s
    return a


if __name__ == "__main__":
    main()
