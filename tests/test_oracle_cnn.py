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
