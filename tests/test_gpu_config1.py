"""OS-SQS parity at the full 32-slice workload: 512^2 slices at 0.7 mm,
1152 views, 736 bins, 12 interleaved subsets, Poisson noise at I0 = 1e5,
started from the FBP image, with the launch shapes the autotuner picks for
this batch (the interleaved 4-slice layout and the backprojection with the
update fused into its epilogue).

Slice 0 and slice 29 (in the last 4-slice group) of the engine's pass are
compared with the FP64 oracle: one plain pass (beta = 0, ossqs.cu:91-149)
and one SQS-TV pass against its definition composed from the oracle's
Siddon pair and TVSQSKernel terms (TotalVariation.cu:9-68).
"""

import numpy as np
import pytest

import oracle as orc
from reconnet_amd.geometry import WORKLOADS
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5
SLICES = (0, 29)


@pytest.fixture(scope="module")
def c1():
    import torch
    from reconnet_amd.phantom import add_poisson_noise, phantom_batch
    from reconnet_amd.sqs import OsSqsTV, fbp_initial_image
    from reconnet_amd.torch_ops import FanPlan

    geo = WORKLOADS["fan512"]
    ns = 32
    plan = FanPlan(geo)
    truth = phantom_batch(range(1000, 1000 + ns), geo.nx, geo.ny, geo.dx, geo.dy, device="cuda")
    clean = plan.forward(truth, "siddon")
    b, w = add_poisson_noise(clean, 1e5, seed=7)
    x0 = fbp_initial_image(plan, b)
    eng = OsSqsTV(geo, 12, ns)
    eng.set_data(b, w)
    eng.compute_norm()
    D = eng.norm()
    torch.cuda.synchronize()
    host = dict(b=b.cpu().numpy(), w=w.cpu().numpy(), x0=x0.cpu().numpy(), D=D.cpu().numpy())
    yield geo, eng, x0, host
    eng.close()


def _oracle_args(geo, eng, host, s):
    order = eng.order
    deg = geo.angles()[order]
    b = host["b"][s][:, order][None, :, :, None, None]
    w = host["w"][s][:, order][None, :, :, None, None]
    x0 = host["x0"][s][None, :, :, None, None]
    D = host["D"][s][None, :, :, None, None]
    return deg, b, w, x0, D


def test_plain_pass_matches_oracle(c1):
    geo, eng, x0, host = c1
    eng.set_image(x0)
    eng.run_pass(beta=0.0, lb=0.0)
    out = eng.image().cpu().numpy()
    for s in SLICES:
        deg, b, w, xs, D = _oracle_args(geo, eng, host, s)
        ref = orc.os_sqs(xs, b, D, w, deg, geo.oracle_geo(), 12, lb=0.0)[0, :, :, 0, 0]
        assert np.abs(ref - xs[0, :, :, 0, 0]).max() > 1e-3
        assert rel_l2(out[s], ref) <= TOL, s


def test_sqs_tv_pass_matches_oracle_composition(c1):
    # beta large enough that the TV terms change the update visibly
    geo, eng, x0, host = c1
    beta, nsub = 5.0, 12
    eng.set_image(x0)
    eng.run_pass(beta=beta, eps=1e-6, lb=0.0)
    out = eng.image().cpu().numpy()
    g5 = geo.oracle_geo()
    npers = geo.nview // nsub
    for s in SLICES:
        deg, b, w, xs, D = _oracle_args(geo, eng, host, s)
        x = xs[0, :, :, 0, 0].astype(np.float64)
        moved_tv = 0.0
        for k in range(nsub):
            v = slice(k * npers, (k + 1) * npers)
            xi = x.astype(np.float32)[None, :, :, None, None]
            fp = orc.siddon_fan_fp(xi, deg[v], g5, geo.nu)
            r = (w[:, :, v] * (fp - b[:, :, v])).astype(np.float32)
            gd = orc.siddon_fan_bp(r, deg[v], g5, (geo.nx, geo.ny, 1))[0, :, :, 0, 0]
            gt, nt, _ = orc.tv_sqs(x.astype(np.float32))
            num = nsub * gd + 2 * beta * gt
            den = D[0, :, :, 0, 0] + 2 * beta * nt
            moved_tv = max(moved_tv, float(np.abs(2 * beta * gt / (den + 1e-6)).max()))
            x = np.maximum(x - num / (den + 1e-6), 0.0)
        assert moved_tv > 1e-4   # the prior is not negligible in this test
        assert rel_l2(out[s], x) <= TOL, s
