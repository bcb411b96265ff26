"""Train the learned unrolled OS-SQS network on one GPU.

    python -m reconnet_amd.train_unrolled --train 2000 --out reconnet_amd/weights/unrolled_fan512.npz

Data: seeded random-ellipse phantoms (seeds 0 .. N-1) on the 512^2 fan
geometry (0.7 mm pixels, 1152 views, 736 bins), projected by the library's
Siddon projector and corrupted with Poisson noise at I0 (quarter dose 2.5e4 by
default).  All N slices stay resident in HBM in one OS-SQS engine.

Every slice starts from its FBP image (Hann-windowed ramp, FBP-weighted DD
backprojection), as in the reconstruction itself.  Training is greedy: stage k
takes the current images of all slices, applies one
OS-SQS data pass (12 subsets), and fits the stage-k CNN prior (initialised from
stage k-1) to map the result to the phantom (Adam on the batch mean of the
per-slice RMSE in HU over the field of view, the evaluation metric).  The
trained stage is applied to every slice before the next stage starts.  Every
forward and backward convolution runs on the library's BF16 MFMA kernels.
"""

import argparse
import json
import sys
import time

import numpy as np
import torch

from .cnn import PriorCNN, init_params
from .geometry import HU_WATER, WORKLOADS
from .phantom import add_poisson_noise, phantom_batch
from .sqs import OsSqsTV, fbp_initial_image
from .torch_ops import FanPlan
from .unrolled import save_nets

HU = 1000.0 / HU_WATER


def log(msg):
    print("[train] " + msg, flush=True)


def simulate(geo, seeds, i0, noise_seed, chunk=64):
    """Phantoms, noisy sinograms, weights and FBP starting images for the given
    seeds (CUDA tensors)."""
    plan = FanPlan(geo)
    truths, sinos, wts, inits = [], [], [], []
    for i in range(0, len(seeds), chunk):
        s = list(seeds[i:i + chunk])
        t = phantom_batch(s, geo.nx, geo.ny, geo.dx, geo.dy, device="cuda").contiguous()
        p = plan.forward(t, "siddon")
        n, w = add_poisson_noise(p, i0, seed=noise_seed + int(s[0]))
        truths.append(t)
        sinos.append(n)
        wts.append(w)
        inits.append(fbp_initial_image(plan, n))
    return torch.cat(truths), torch.cat(sinos), torch.cat(wts), torch.cat(inits)


def masked_rmse_hu(x, truth, mask):
    d = (x - truth) * mask
    return HU * torch.sqrt((d * d).sum(dim=(1, 2)) / mask.sum())


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--val", type=int, default=32)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--subsets", type=int, default=12)
    ap.add_argument("--i0", type=float, default=2.5e4)
    ap.add_argument("--epochs", type=int, default=16)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--scale", type=float, default=50.0)
    ap.add_argument("--out", default="unrolled_fan512.npz")
    ap.add_argument("--log", default="train_log_fan512.json")
    args = ap.parse_args(argv)

    torch.cuda.set_device(0)
    geo = WORKLOADS["fan512"]
    mask = torch.from_numpy(geo.fov_mask()).cuda().float()
    t0 = time.time()
    truth, sino, w, x0 = simulate(geo, list(range(args.train)), args.i0, noise_seed=7)
    vtruth, vsino, vw, vx0 = simulate(geo, list(range(50000, 50000 + args.val)), args.i0, noise_seed=7)
    log("simulated %d + %d slices in %.1f s" % (args.train, args.val, time.time() - t0))

    eng = OsSqsTV(geo, args.subsets, args.train)
    eng.set_data(sino, w)
    del sino, w
    eng.compute_norm()
    veng = OsSqsTV(geo, args.subsets, args.val)
    veng.set_data(vsino, vw)
    veng.compute_norm()
    del vsino, vw
    log("normalisers ready (%.1f s)" % (time.time() - t0))

    eng.set_image(x0)
    veng.set_image(vx0)
    log("FBP start: val RMSE %.2f HU" % float(masked_rmse_hu(vx0, vtruth, mask).mean()))
    del x0, vx0
    nets, history = [], []
    params = init_params(seed=0)
    gen = torch.Generator(device="cpu").manual_seed(1)
    for k in range(args.iters):
        ts = time.time()
        eng.run_pass(beta=0.0, lb=0.0)
        veng.run_pass(beta=0.0, lb=0.0)
        xin = eng.image()
        vin = veng.image()
        net = PriorCNN(params, scale=args.scale)
        opt = torch.optim.Adam(net.parameters(), lr=args.lr)
        nsteps = args.epochs * (args.train // args.batch)
        sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, nsteps)
        losses, epoch_means = [], []
        npix = float(mask.sum())
        for ep in range(args.epochs):
            perm = torch.randperm(args.train, generator=gen).cuda()
            first = len(losses)
            for i in range(0, args.train - args.batch + 1, args.batch):
                idx = perm[i:i + args.batch]
                xb = xin.index_select(0, idx)
                tb = truth.index_select(0, idx)
                y = net.forward(xb, keep=True)
                r = (y - tb) * mask
                # loss: mean over the batch of the per-slice RMSE in HU (the evaluation metric)
                rms = torch.sqrt((r * r).sum(dim=(1, 2)) / npix).clamp_min(1e-12)
                losses.append(float(HU * rms.mean()))
                grads = net.backward((HU / (args.batch * npix)) * r / rms[:, None, None])
                opt.zero_grad(set_to_none=False)
                for name, p in net.p.items():
                    p.grad = grads[name].reshape(p.shape).contiguous()
                opt.step()
                sched.step()
                net.refresh()
            epoch_means.append(float(np.mean(losses[first:])))
        with torch.no_grad():
            xout = torch.cat([net.forward(xin[i:i + 64]) for i in range(0, args.train, 64)])
            vout = net.forward(vin)
        eng.set_image(xout)
        veng.set_image(vout)
        v_before = float(masked_rmse_hu(vin, vtruth, mask).mean())
        v_after = float(masked_rmse_hu(vout, vtruth, mask).mean())
        t_after = float(masked_rmse_hu(xout, truth, mask).mean())
        history.append({"stage": k, "loss_first": losses[0], "loss_last": float(np.mean(losses[-50:])),
                        "epoch_mean_rmse_hu": epoch_means,
                        "val_rmse_in": v_before, "val_rmse_out": v_after, "train_rmse_out": t_after,
                        "seconds": time.time() - ts})
        log("stage %d: val RMSE %.2f -> %.2f HU (train %.2f), epoch mean batch RMSE %s, %.1f s"
            % (k, v_before, v_after, t_after, " ".join("%.1f" % v for v in epoch_means), time.time() - ts))
        nets.append(net)
        params = {n: p.detach().clone() for n, p in net.p.items()}
        del xin, xout
    save_nets(args.out, nets, meta={"i0": args.i0, "train": args.train, "subsets": args.subsets,
                                    "epochs": args.epochs})
    with open(args.log, "w") as f:
        json.dump({"args": vars(args), "history": history, "seconds": time.time() - t0}, f, indent=1)
    log("saved %s (%.1f s total)" % (args.out, time.time() - t0))


if __name__ == "__main__":
    sys.exit(main())
