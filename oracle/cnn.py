"""FP64 CPU oracle of the learned prior's 3x3 network.

TEST INFRASTRUCTURE ONLY (no reference counterpart: the reference has no
learned prior).  Five 3x3 "same" cross-correlations with zero padding,
1 -> 32 -> 32 -> 32 -> 32 -> 1 channels, ReLU after the first four, residual
output y = x + conv5(a4) + b5.  Tap t = 3 (dy + 1) + (dx + 1) reads the input
at (row + dy, col + dx).  Weight layouts: w1 [9][32] (tap, out),
wm [9][32 out][32 in], w5 [9][32] (tap, in).

forward / backward are the network in exact FP64.  forward_bf16 /
backward_bf16 evaluate it with the roundings the GPU kernels perform: the
weights of the 32 -> 32 layers and every stored activation and data gradient
rounded to BF16 (round to nearest even), all sums in FP64.  A kernel that
matches the BF16 version closely but the FP64 version only loosely has the
right arithmetic; the FP64 gap is the BF16 storage error.
"""

import numpy as np

__all__ = ["conv_taps", "forward", "backward", "forward_bf16", "backward_bf16", "bf16_round", "init_params"]

OFFS = [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def _shift(a, dy, dx):
    """out[..., i, j, :] = a[..., i + dy, j + dx, :] (zero outside); a: [B, H, W, C]."""
    B, H, W, C = a.shape
    out = np.zeros_like(a)
    ys = slice(max(0, -dy), min(H, H - dy))
    xs = slice(max(0, -dx), min(W, W - dx))
    yd = slice(max(0, dy), min(H, H + dy))
    xd = slice(max(0, dx), min(W, W + dx))
    out[:, ys, xs, :] = a[:, yd, xd, :]
    return out


def conv_taps(a, w):
    """a: [B, H, W, Cin]; w: [9][Cout][Cin] -> [B, H, W, Cout] (FP64)."""
    a = np.asarray(a, np.float64)
    out = 0.0
    for t, (dy, dx) in enumerate(OFFS):
        out = out + _shift(a, dy, dx) @ np.asarray(w[t], np.float64).T
    return out


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {"w1": rng.normal(0, np.sqrt(2 / 9), (9, 32)), "b1": rng.normal(0, 0.01, 32)}
    for k in (2, 3, 4):
        p["w%d" % k] = rng.normal(0, np.sqrt(2 / 288), (9, 32, 32))
        p["b%d" % k] = rng.normal(0, 0.01, 32)
    p["w5"] = rng.normal(0, np.sqrt(1 / 288), (9, 32))
    p["b5"] = np.array(rng.normal(0, 0.001))
    return p


def forward(x, p):
    """x: [B, H, W] -> (y, cache)."""
    x = np.asarray(x, np.float64)
    a0 = x[..., None]
    z1 = conv_taps(a0, np.asarray(p["w1"])[:, :, None]) + p["b1"]
    a1 = np.maximum(z1, 0)
    acts = [a0, a1]
    a = a1
    for k in (2, 3, 4):
        z = conv_taps(a, p["w%d" % k]) + p["b%d" % k]
        a = np.maximum(z, 0)
        acts.append(a)
    y = x + conv_taps(a, np.asarray(p["w5"])[:, None, :])[..., 0] + float(p["b5"])
    return y, acts


def backward(dy, p, acts):
    """Gradients of sum(dy * y) with respect to all parameters (FP64)."""
    dy = np.asarray(dy, np.float64)
    g = {}
    a4 = acts[4]
    d = dy[..., None]
    g["w5"] = np.stack([np.einsum("bhw,bhwc->c", dy, _shift(a4, dyy, dxx)) for dyy, dxx in OFFS])
    g["b5"] = np.array(dy.sum())
    # data gradient of a cross-correlation: correlate with the flipped kernel
    w5 = np.asarray(p["w5"])
    dz = 0.0
    for t, (dyy, dxx) in enumerate(OFFS):
        dz = dz + _shift(d, -dyy, -dxx) * w5[t][None, None, None, :]
    dz = dz * (a4 > 0)
    for k in (4, 3, 2):
        a_in = acts[k - 1]
        g["w%d" % k] = np.stack([np.einsum("bhwo,bhwi->oi", dz, _shift(a_in, dyy, dxx)) for dyy, dxx in OFFS])
        g["b%d" % k] = dz.sum(axis=(0, 1, 2))
        w = np.asarray(p["w%d" % k])
        da = 0.0
        for t, (dyy, dxx) in enumerate(OFFS):
            da = da + _shift(dz, -dyy, -dxx) @ w[t]
        dz = da * (a_in > 0)
    a0 = acts[0]
    g["w1"] = np.stack([np.einsum("bhwo,bhw->o", dz, _shift(a0, dyy, dxx)[..., 0]) for dyy, dxx in OFFS])
    g["b1"] = dz.sum(axis=(0, 1, 2))
    return g


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), returned as FP64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(x, p):
    """forward() with BF16 storage of a1..a4 and BF16 weights w2..w4."""
    x = np.asarray(x, np.float64)
    a0 = x[..., None]
    a = bf16_round(np.maximum(conv_taps(a0, np.asarray(p["w1"])[:, :, None]) + p["b1"], 0))
    acts = [a0, a]
    for k in (2, 3, 4):
        a = bf16_round(np.maximum(conv_taps(a, bf16_round(p["w%d" % k])) + p["b%d" % k], 0))
        acts.append(a)
    y = x + conv_taps(a, np.asarray(p["w5"])[:, None, :])[..., 0] + float(p["b5"])
    return y, acts


def backward_bf16(dy, p, acts):
    """backward() on forward_bf16's activations, with every data gradient dz
    stored in BF16 and the 32 -> 32 data gradients taken with BF16 weights."""
    dy = np.asarray(dy, np.float64)
    g = {}
    a4 = acts[4]
    d = dy[..., None]
    g["w5"] = np.stack([np.einsum("bhw,bhwc->c", dy, _shift(a4, dyy, dxx)) for dyy, dxx in OFFS])
    g["b5"] = np.array(dy.sum())
    w5 = np.asarray(p["w5"])
    dz = 0.0
    for t, (dyy, dxx) in enumerate(OFFS):
        dz = dz + _shift(d, -dyy, -dxx) * w5[t][None, None, None, :]
    dz = bf16_round(dz * (a4 > 0))
    for k in (4, 3, 2):
        a_in = acts[k - 1]
        g["w%d" % k] = np.stack([np.einsum("bhwo,bhwi->oi", dz, _shift(a_in, dyy, dxx)) for dyy, dxx in OFFS])
        g["b%d" % k] = dz.sum(axis=(0, 1, 2))
        w = bf16_round(p["w%d" % k])
        da = 0.0
        for t, (dyy, dxx) in enumerate(OFFS):
            da = da + _shift(dz, -dyy, -dxx) @ w[t]
        dz = bf16_round(da * (a_in > 0))
    a0 = acts[0]
    g["w1"] = np.stack([np.einsum("bhwo,bhw->o", dz, _shift(a0, dyy, dxx)[..., 0]) for dyy, dxx in OFFS])
    g["b1"] = dz.sum(axis=(0, 1, 2))
    return g

