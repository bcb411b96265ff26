"""What the trained prior gains: on held-out slices (seeds 100000+, the dose
the weights were trained for), the 10-iteration learned reconstruction against
the same 10 OS-SQS data passes without the CNN, and against plain SQS-TV
(beta = 1e-4, delta = 1e-3) after the same passes, all from the FBP image."""

import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_trained_prior_beats_the_same_passes_without_it():
    from reconnet_amd.geometry import WORKLOADS
    from reconnet_amd.metrics import rmse_hu
    from reconnet_amd.phantom import add_poisson_noise, phantom_batch
    from reconnet_amd.sqs import OsSqsTV, fbp_initial_image
    from reconnet_amd.torch_ops import FanPlan
    from reconnet_amd.unrolled import WEIGHTS_PATH, UnrolledOSSQS, load_nets

    if not os.path.exists(WEIGHTS_PATH):
        pytest.skip("no trained weights in the package")
    z = np.load(WEIGHTS_PATH, allow_pickle=False)
    i0 = float(z["meta_i0"])
    nets = load_nets(WEIGHTS_PATH)
    geo = WORKLOADS["fan512"]
    ns = 8
    plan = FanPlan(geo)
    truth = phantom_batch(list(range(100000, 100000 + ns)), geo.nx, geo.ny, geo.dx, geo.dy, device="cuda")
    noisy, w = add_poisson_noise(plan.forward(truth.contiguous(), "siddon"), i0, seed=77)
    mask = torch.from_numpy(geo.fov_mask()).cuda().float()
    eng = OsSqsTV(geo, 12, ns)
    eng.set_data(noisy, w)
    eng.compute_norm()
    x0 = fbp_initial_image(plan, noisy)

    learned = rmse_hu(UnrolledOSSQS(nets, lb=0.0).reconstruct(eng, x0), truth, mask).mean().item()
    eng.set_image(x0)
    for _ in nets:
        eng.run_pass(beta=0.0, lb=0.0)
    no_prior = rmse_hu(eng.image(), truth, mask).mean().item()
    eng.set_image(x0)
    for _ in nets:
        eng.run_pass(beta=1e-4, eps=1e-6, lb=0.0)
    tv = rmse_hu(eng.image(), truth, mask).mean().item()
    fbp = rmse_hu(x0, truth, mask).mean().item()
    print("RMSE HU: fbp %.2f, %d passes %.2f, %d passes SQS-TV %.2f, learned %.2f"
          % (fbp, len(nets), no_prior, len(nets), tv, learned))
    assert learned < 0.9 * no_prior
    assert learned < 0.9 * tv
