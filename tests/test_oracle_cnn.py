"""Pins of the FP64 CNN oracle (oracle/cnn.py): the convolution against
scipy's independent correlation routine, gradients against central finite
differences."""

import numpy as np
import pytest
import scipy.signal

from oracle import cnn


def test_conv_taps_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(2, 9, 11, 3))
    w = rng.normal(size=(9, 4, 3))
    out = cnn.conv_taps(a, w)
    for b in range(2):
        for co in range(4):
            ref = 0.0
            for ci in range(3):
                k = w[:, co, ci].reshape(3, 3)   # k[dy+1, dx+1]
                ref = ref + scipy.signal.correlate2d(a[b, :, :, ci], k, mode="same", boundary="fill")
            assert np.allclose(out[b, :, :, co], ref, rtol=1e-12, atol=1e-12)


def test_residual_identity_when_last_layer_zero():
    p = cnn.init_params(1)
    p["w5"] = np.zeros_like(p["w5"])
    p["b5"] = np.array(0.0)
    x = np.random.default_rng(2).uniform(size=(1, 8, 8))
    y, _ = cnn.forward(x, p)
    assert np.array_equal(y, x)


@pytest.mark.parametrize("name", ["w1", "b1", "w2", "b3", "w4", "w5", "b5"])
def test_gradients_match_finite_differences(name):
    rng = np.random.default_rng(3)
    p = cnn.init_params(4)
    x = rng.uniform(size=(2, 7, 6))
    c = rng.normal(size=(2, 7, 6))
    y, acts = cnn.forward(x, p)
    g = cnn.backward(c, p, acts)[name]
    base = np.asarray(p[name], np.float64)
    flat = base.reshape(-1)
    idx = rng.choice(flat.size, size=min(6, flat.size), replace=False)
    h = 1e-6
    for i in idx:
        q = {k: np.array(v, np.float64, copy=True) for k, v in p.items()}
        q[name].reshape(-1)[i] += h
        fp = np.sum(c * cnn.forward(x, q)[0])
        q[name].reshape(-1)[i] -= 2 * h
        fm = np.sum(c * cnn.forward(x, q)[0])
        assert (fp - fm) / (2 * h) == pytest.approx(np.asarray(g).reshape(-1)[i], rel=1e-5, abs=1e-8)


def test_bf16_round_matches_torch_conversion():
    import torch

    rng = np.random.default_rng(5)
    v = np.concatenate([rng.normal(size=4000) * 10.0 ** rng.integers(-6, 6, 4000),
                        # exact ties between two BF16 values (odd and even mantissa)
                        (np.arange(1, 200) * 2 + 1).astype(np.float64) * 2.0 ** -8 + 1.0]).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(cnn.bf16_round(v), ref)


def _sparse_int_params(seed):
    rng = np.random.default_rng(seed)
    p = {"w1": rng.integers(-1, 2, (9, 32)).astype(np.float64), "b1": rng.integers(0, 2, 32).astype(np.float64)}
    for k in (2, 3, 4):
        p["w%d" % k] = (rng.integers(-1, 2, (9, 32, 32)) * (rng.uniform(size=(9, 32, 32)) < 0.05)).astype(np.float64)
        p["b%d" % k] = rng.integers(0, 2, 32).astype(np.float64)
    p["w5"] = rng.integers(-1, 2, (9, 32)).astype(np.float64)
    p["b5"] = np.array(1.0)
    return p


def test_bf16_variant_equals_fp64_when_nothing_rounds():
    # small integers everywhere: every stored value is BF16-exact, so the two
    # variants must agree exactly, values and gradients
    p = _sparse_int_params(6)
    rng = np.random.default_rng(7)
    x = rng.integers(0, 3, (1, 10, 12)).astype(np.float64)
    c = rng.integers(-1, 2, (1, 10, 12)).astype(np.float64)
    y, acts = cnn.forward(x, p)
    yb, actsb = cnn.forward_bf16(x, p)
    assert max(np.abs(a).max() for a in acts) <= 256
    assert np.array_equal(y, yb)
    g = cnn.backward(c, p, acts)
    gb = cnn.backward_bf16(c, p, actsb)
    for k in g:
        assert np.array_equal(g[k], gb[k]), k


def test_bf16_variant_rounds_activations():
    p = cnn.init_params(8)
    x = np.random.default_rng(9).uniform(size=(1, 6, 6))
    _, acts = cnn.forward_bf16(x, p)
    for a in acts[1:]:
        assert np.array_equal(a, cnn.bf16_round(a))
    _, acts64 = cnn.forward(x, p)
    assert not np.array_equal(acts[2], acts64[2])
