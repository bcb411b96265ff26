"""GPU parity of the BF16 MFMA convolution kernels against the FP64 CNN oracle.

Forward: checked against the oracle evaluated with the same BF16 roundings
(weights of the 32->32 layers, every stored activation), which isolates the
kernels' indexing and accumulation (bar 1e-4 relative L2), and against the pure
FP64 network (bar 2e-2, the BF16 storage error).  Backward: every parameter
gradient against the FP64 oracle (bar 3e-2) plus an exact-arithmetic check
of the MFMA weight-gradient kernel on BF16-exact inputs.
"""

import numpy as np
import pytest
import torch

from oracle import cnn as ocnn
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def bf16_round(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def oracle_forward_bf16(x, p):
    a0 = np.asarray(x, np.float64)[..., None]
    a = bf16_round(np.maximum(ocnn.conv_taps(a0, np.asarray(p["w1"])[:, :, None]) + p["b1"], 0))
    for k in (2, 3, 4):
        a = bf16_round(np.maximum(ocnn.conv_taps(a, bf16_round(p["w%d" % k])) + p["b%d" % k], 0))
    return x + ocnn.conv_taps(a, np.asarray(p["w5"])[:, None, :])[..., 0] + float(p["b5"])


@pytest.fixture(scope="module")
def setup():
    from reconnet_amd.cnn import PriorCNN, init_params

    params = init_params(seed=5, device="cuda", last_scale=1.0)
    for k in ("b1", "b2", "b3", "b4"):
        params[k] = torch.randn(32, generator=torch.Generator().manual_seed(7)).cuda() * 0.05
    params["b5"] = torch.tensor(0.01).cuda()
    net = PriorCNN(params)
    pn = {k: v.cpu().numpy().astype(np.float64) for k, v in net.p.items()}
    return net, pn


@pytest.mark.parametrize("shape", [(2, 24, 40), (1, 64, 64), (1, 17, 50)])
def test_forward_matches_oracle(setup, shape):
    net, pn = setup
    x = np.random.default_rng(1).uniform(0, 1, shape).astype(np.float32)
    y = net.forward(torch.from_numpy(x).cuda()).cpu().numpy().astype(np.float64)
    ref_bf = oracle_forward_bf16(x, pn)
    ref, _ = ocnn.forward(x, pn)
    assert rel_l2(y - x, ref_bf - x) <= 1e-4
    assert rel_l2(y - x, ref - x) <= 2e-2


def test_backward_matches_oracle(setup):
    net, pn = setup
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 1, (2, 24, 40)).astype(np.float32)
    c = rng.normal(size=x.shape).astype(np.float32)
    net.forward(torch.from_numpy(x).cuda(), keep=True)
    g = {k: v.cpu().numpy() for k, v in net.backward(torch.from_numpy(c).cuda()).items()}
    _, acts = ocnn.forward(x, pn)
    ref = ocnn.backward(c, pn, acts)
    for k in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4", "w5", "b5"):
        assert rel_l2(g[k], ref[k]) <= 3e-2, k


def test_wgrad_mid_exact_on_bf16_inputs():
    # inputs exactly representable in BF16 and small integers: every product and
    # partial sum is exact in FP32, so the MFMA result must equal the oracle exactly
    from reconnet_amd.cnn import _p, _sigs, _st

    h = _sigs()
    rng = np.random.default_rng(3)
    B, H, W = 1, 9, 70
    a = rng.integers(-3, 4, (B, H, W, 32)).astype(np.float32)
    d = rng.integers(-3, 4, (B, H, W, 32)).astype(np.float32)
    ta = torch.from_numpy(a).cuda().to(torch.bfloat16)
    td = torch.from_numpy(d).cuda().to(torch.bfloat16)
    gw = torch.empty(9, 32, 32, device="cuda")
    assert h.rn_cnn_wgrad_mid(_p(ta), _p(td), _p(gw), B, H, W, _st()) == 0
    ref = np.stack([np.einsum("bhwo,bhwi->oi", d, ocnn._shift(a, dy, dx)) for dy, dx in ocnn.OFFS])
    assert np.array_equal(gw.cpu().numpy(), ref.astype(np.float32))


def test_training_step_reduces_loss():
    from reconnet_amd.cnn import PriorCNN, init_params

    torch.manual_seed(0)
    net = PriorCNN(init_params(seed=1))
    x = torch.rand(2, 32, 32, device="cuda")
    target = torch.nn.functional.avg_pool2d(x[:, None], 3, 1, 1)[:, 0].contiguous()
    opt = torch.optim.Adam(net.parameters(), lr=1e-3)
    losses = []
    for _ in range(30):
        y = net.forward(x, keep=True)
        r = y - target
        losses.append(float((r * r).mean()))
        grads = net.backward(2 * r / r.numel())
        opt.zero_grad()
        for k, p in net.p.items():
            p.grad = grads[k].reshape(p.shape)
        opt.step()
        net.refresh()
    assert losses[-1] < 0.5 * losses[0]
