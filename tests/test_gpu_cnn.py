"""GPU parity of the learned prior's convolution kernels against the CNN oracle.

Exact checks: every entry point on small-integer inputs, where every product
and partial sum is exact in FP32 and the single BF16 rounding of a stored
result is deterministic, so the kernel must equal the oracle bit for bit
(this pins the MFMA lane layout, the tap offsets, borders, masks and bias).
Network checks: forward and all parameter gradients against the oracle
evaluated with the kernels' BF16 roundings (bars 2e-3 / 1e-2 relative L2: an
accumulation-order difference can flip the BF16 rounding of a stored value
by one unit) and against the exact FP64 network (bar 2e-2 forward; for
gradients the BF16 storage error is itself large, see the oracle's docstring,
so the FP64 comparison only bounds it).
"""

import numpy as np
import pytest
import torch

from oracle import cnn as ocnn
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _h():
    from reconnet_amd.cnn import _sigs

    return _sigs()


def _p(t):
    from reconnet_amd.cnn import _p as p

    return p(t)


def _st():
    from reconnet_amd.cnn import _st as s

    return s()


def _ints(rng, lo, hi, shape):
    return rng.integers(lo, hi + 1, shape).astype(np.float32)


def _bf(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda().to(torch.bfloat16)


def _f(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


SHAPES = [(1, 9, 70), (2, 17, 33), (1, 64, 64)]


@pytest.mark.parametrize("shape", SHAPES)
def test_conv_in_exact(shape):
    B, H, W = shape
    rng = np.random.default_rng(11)
    x, w, b = _ints(rng, -3, 3, shape), _ints(rng, -2, 2, (9, 32)), _ints(rng, -2, 2, 32)
    out = torch.empty(B, H, W, 32, dtype=torch.bfloat16, device="cuda")
    assert _h().rn_cnn_conv_in(_p(_f(x)), _p(_f(w)), _p(_f(b)), _p(out), B, H, W, _st()) == 0
    ref = ocnn.bf16_round(np.maximum(ocnn.conv_taps(x[..., None], w[:, :, None]) + b, 0))
    assert np.array_equal(out.float().cpu().numpy(), ref.astype(np.float32))


N_CONV_MID = 5      # launch candidates of the 32 -> 32 layer (LDS-staged rows x 4, per-wave walk)


def _conv_mid(a, w, b, mask, mode, shape, cand=None):
    from reconnet_amd._lib import autotune_pin

    B, H, W = shape
    out = torch.empty(B, H, W, 32, dtype=torch.bfloat16, device="cuda")
    tb, tm = _f(b), _bf(mask)
    if cand is not None:
        autotune_pin("conv_mid", cand)
    try:
        st = _h().rn_cnn_conv_mid(_p(_bf(a)), _p(_bf(w)), _p(tb) if mode else None, None if mode else _p(tm),
                                  _p(out), B, H, W, mode, _st())
    finally:
        if cand is not None:
            autotune_pin("conv_mid", None)
    assert st == 0
    return out.float().cpu().numpy()


@pytest.mark.parametrize("cand", [None] + list(range(N_CONV_MID)))
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("mode", [1, 0])
def test_conv_mid_exact(shape, mode, cand):
    B, H, W = shape
    rng = np.random.default_rng(12 + mode)
    a, w = _ints(rng, -3, 3, (B, H, W, 32)), _ints(rng, -2, 2, (9, 32, 32))
    b = _ints(rng, -4, 4, 32)
    mask = _ints(rng, -1, 1, (B, H, W, 32))
    out = _conv_mid(a, w, b, mask, mode, shape, cand)
    z = ocnn.conv_taps(a, w)
    ref = np.maximum(z + b, 0) if mode else z * (mask > 0)
    assert np.array_equal(out, ocnn.bf16_round(ref).astype(np.float32))


@pytest.mark.parametrize("mode", [1, 0])
def test_conv_mid_launch_shapes_agree_bitwise(mode):
    # real-valued inputs: every candidate runs the same MFMA sequence per output
    shape = (2, 45, 150)
    rng = np.random.default_rng(40 + mode)
    a = rng.normal(0, 1, shape + (32,)).astype(np.float32)
    w = rng.normal(0, 0.1, (9, 32, 32)).astype(np.float32)
    b = rng.normal(0, 0.1, 32).astype(np.float32)
    mask = rng.normal(0, 1, shape + (32,)).astype(np.float32)
    outs = [_conv_mid(a, w, b, mask, mode, shape, c) for c in range(N_CONV_MID)]
    assert np.abs(outs[0]).max() > 0.1
    for c in range(1, N_CONV_MID):
        assert np.array_equal(outs[c], outs[0]), c


@pytest.mark.parametrize("shape", SHAPES)
def test_conv_out_and_out_dgrad_exact(shape):
    B, H, W = shape
    rng = np.random.default_rng(14)
    a4, w5 = _ints(rng, 0, 3, (B, H, W, 32)), _ints(rng, -2, 2, (9, 32))
    x, dy = _ints(rng, -5, 5, shape), _ints(rng, -3, 3, shape)
    y = torch.empty(B, H, W, device="cuda")
    assert _h().rn_cnn_conv_out(_p(_bf(a4)), _p(_f(w5)), 0.5, _p(_f(x)), _p(y), B, H, W, _st()) == 0
    ref = x + ocnn.conv_taps(a4, w5[:, None, :])[..., 0] + 0.5
    assert np.array_equal(y.cpu().numpy(), ref.astype(np.float32))
    mask = _ints(rng, -1, 1, (B, H, W, 32))
    dz = torch.empty(B, H, W, 32, dtype=torch.bfloat16, device="cuda")
    assert _h().rn_cnn_out_dgrad(_p(_f(dy)), _p(_f(w5)), _p(_bf(mask)), _p(dz), B, H, W, _st()) == 0
    refd = 0.0
    for t, (ry, rx) in enumerate(ocnn.OFFS):
        refd = refd + ocnn._shift(dy[..., None].astype(np.float64), -ry, -rx) * w5[t][None, None, None, :]
    refd = ocnn.bf16_round(refd * (mask > 0))
    assert np.array_equal(dz.float().cpu().numpy(), refd.astype(np.float32))


@pytest.mark.parametrize("shape", SHAPES)
def test_weight_and_bias_gradients_exact(shape):
    B, H, W = shape
    rng = np.random.default_rng(15)
    x, d1 = _ints(rng, -3, 3, shape), _ints(rng, -3, 3, (B, H, W, 32))
    buf = torch.empty(9 * 32 + 32, device="cuda")
    assert _h().rn_cnn_wgrad_narrow(_p(_f(x)), _p(_bf(d1)), _p(buf), B, H, W, 0, _st()) == 0
    gw = np.stack([np.einsum("bhwo,bhw->o", d1, ocnn._shift(x[..., None], ry, rx)[..., 0]) for ry, rx in ocnn.OFFS])
    got = buf.cpu().numpy()
    assert np.array_equal(got[:288].reshape(9, 32), gw.astype(np.float32))
    assert np.array_equal(got[288:], d1.sum(axis=(0, 1, 2)).astype(np.float32))
    dy, a4 = _ints(rng, -3, 3, shape), _ints(rng, 0, 3, (B, H, W, 32))
    buf5 = torch.empty(9 * 32 + 1, device="cuda")
    assert _h().rn_cnn_wgrad_narrow(_p(_f(dy)), _p(_bf(a4)), _p(buf5), B, H, W, 1, _st()) == 0
    g5 = np.stack([np.einsum("bhw,bhwc->c", dy, ocnn._shift(a4, ry, rx)) for ry, rx in ocnn.OFFS])
    got5 = buf5.cpu().numpy()
    assert np.array_equal(got5[:288].reshape(9, 32), g5.astype(np.float32))
    assert got5[288] == np.float32(dy.sum())
    gb = torch.empty(32, device="cuda")
    assert _h().rn_cnn_bias_grad(_p(_bf(d1)), _p(gb), B * H * W, _st()) == 0
    assert np.array_equal(gb.cpu().numpy(), d1.sum(axis=(0, 1, 2)).astype(np.float32))


def test_wgrad_mid_exact():
    rng = np.random.default_rng(3)
    B, H, W = 1, 9, 70
    a = _ints(rng, -3, 3, (B, H, W, 32))
    d = _ints(rng, -3, 3, (B, H, W, 32))
    gw = torch.empty(9, 32, 32, device="cuda")
    assert _h().rn_cnn_wgrad_mid(_p(_bf(a)), _p(_bf(d)), _p(gw), B, H, W, _st()) == 0
    ref = np.stack([np.einsum("bhwo,bhwi->oi", d, ocnn._shift(a, dy, dx)) for dy, dx in ocnn.OFFS])
    assert np.array_equal(gw.cpu().numpy(), ref.astype(np.float32))


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
    ref_bf, _ = ocnn.forward_bf16(x, pn)
    ref, _ = ocnn.forward(x, pn)
    assert rel_l2(y - x, ref_bf - x) <= 2e-3
    assert rel_l2(y - x, ref - x) <= 2e-2


def test_backward_matches_oracle(setup):
    net, pn = setup
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 1, (2, 24, 40)).astype(np.float32)
    c = rng.normal(size=x.shape).astype(np.float32)
    net.forward(torch.from_numpy(x).cuda(), keep=True)
    g = {k: v.cpu().numpy() for k, v in net.backward(torch.from_numpy(c).cuda()).items()}
    _, acts_bf = ocnn.forward_bf16(x, pn)
    ref_bf = ocnn.backward_bf16(c, pn, acts_bf)
    _, acts = ocnn.forward(x, pn)
    ref = ocnn.backward(c, pn, acts)
    for k in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4", "w5", "b5"):
        assert rel_l2(g[k], ref_bf[k]) <= 1e-2, k
        assert rel_l2(g[k], ref[k]) <= 0.3, k


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


@pytest.mark.parametrize("shape", [(1, 9, 70), (2, 17, 33), (1, 64, 64), (3, 130, 200), (1, 64, 89), (1, 200, 97)])
def test_fused_forward_equals_layer_by_layer_bitwise(shape):
    # one launch of the whole network against conv_in, 3 x conv_mid, conv_out:
    # strips of 56 or 88 output columns overlapping by 8, each workgroup's
    # share of (image, strip, row) walked as segments (a share crosses strips
    # and images at these sizes), borders; every launch shape
    from reconnet_amd.cnn import PriorCNN, init_params

    B, H, W = shape
    p = init_params(seed=3)
    for k in ("b1", "b2", "b3", "b4"):
        p[k] = torch.randn(p[k].shape, generator=torch.Generator().manual_seed(len(k) + B)).cuda() * 0.1
    p["b5"] = torch.tensor(0.01).cuda()
    net = PriorCNN(p, scale=50.0)
    g = torch.Generator(device="cuda").manual_seed(H * W)
    x = torch.rand((B, H, W), generator=g, device="cuda") * 0.04
    ref = net.forward(x, fused=False)
    assert (ref - x).abs().max() > 1e-4
    from reconnet_amd._lib import autotune_pin

    for cand in [None, 0, 1, 2, 3]:
        if cand is not None:
            autotune_pin("cnn_forward", cand)
        try:
            out = net.forward(x)
        finally:
            autotune_pin("cnn_forward", None)
        assert torch.equal(out, ref), cand
