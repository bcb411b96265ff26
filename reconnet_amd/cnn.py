"""Learned 3x3 prior of the unrolled OS-SQS network (no reference counterpart).

Five 3x3 convolutions, 1 -> 32 -> 32 -> 32 -> 32 -> 1 channels, ReLU after
the first four, residual output y = x + conv5(a4) + b5 (28,353 parameters).
Forward and backward run on the library's HIP kernels (rn_cnn_*): BF16
activations in NHWC, BF16 weights on the gfx950 matrix cores for the 32 -> 32
layers, FP32 accumulation everywhere, FP32 master weights.  Tap t = 3 (dy+1) +
(dx+1) reads the input at (row + dy, col + dx).
"""

import ctypes

import numpy as np
import torch

from ._lib import check, lib

__all__ = ["PriorCNN", "init_params"]

_P = ctypes.c_void_p


def _sigs():
    h = lib()
    if getattr(h, "_cnn_sigs", False):
        return h
    ci, cf, cl = ctypes.c_int, ctypes.c_float, ctypes.c_long
    for name, args in [
        ("rn_cnn_conv_in", [_P, _P, _P, _P, ci, ci, ci, _P]),
        ("rn_cnn_conv_mid", [_P, _P, _P, _P, _P, ci, ci, ci, ci, _P]),
        ("rn_cnn_conv_out", [_P, _P, cf, _P, _P, ci, ci, ci, _P]),
        ("rn_cnn_forward", [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, cf, _P, ci, ci, ci, _P]),
        ("rn_cnn_out_dgrad", [_P, _P, _P, _P, ci, ci, ci, _P]),
        ("rn_cnn_wgrad_narrow", [_P, _P, _P, ci, ci, ci, ci, _P]),
        ("rn_cnn_bias_grad", [_P, _P, cl, _P]),
        ("rn_cnn_wgrad_mid", [_P, _P, _P, ci, ci, ci, _P]),
    ]:
        fn = getattr(h, name)
        fn.restype = ci
        fn.argtypes = args
    h._cnn_sigs = True
    return h


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _st():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def init_params(seed=0, device="cuda", last_scale=0.1):
    """He-initialised FP32 parameters (w5 scaled down so the prior starts near identity)."""
    g = torch.Generator().manual_seed(int(seed))
    p = {
        "w1": torch.randn(9, 32, generator=g) * np.sqrt(2 / 9),
        "b1": torch.zeros(32),
    }
    for k in (2, 3, 4):
        p["w%d" % k] = torch.randn(9, 32, 32, generator=g) * np.sqrt(2 / 288)
        p["b%d" % k] = torch.zeros(32)
    p["w5"] = torch.randn(9, 32, generator=g) * np.sqrt(1 / 288) * last_scale
    p["b5"] = torch.zeros(())
    return {k: v.float().to(device) for k, v in p.items()}


class PriorCNN:
    """Holds FP32 master parameters; forward/backward on device tensors [B, H, W]."""

    NAMES = ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4", "w5", "b5")

    def __init__(self, params, scale=1.0):
        """scale s: the network sees s * x and its residual is divided by s, so
        that images of attenuation values (~0.02 /mm) reach the layers at unit
        scale; folded into w1 and w5 / b5 at launch time."""
        _sigs()
        self.scale = float(scale)
        self.p = {k: params[k].detach().float().contiguous().cuda() for k in self.NAMES}
        self._prep()

    @classmethod
    def from_numpy(cls, arrays, prefix=""):
        scale = float(arrays[prefix + "scale"]) if (prefix + "scale") in arrays else 1.0
        return cls({k: torch.from_numpy(np.asarray(arrays[prefix + k], np.float32)) for k in cls.NAMES}, scale)

    def to_numpy(self, prefix=""):
        d = {prefix + k: v.detach().cpu().numpy() for k, v in self.p.items()}
        d[prefix + "scale"] = np.float32(self.scale)
        return d

    def _prep(self):
        """BF16 copies of the 32 -> 32 weights: forward [t][co][ci] and data-gradient
        [t][ci][co] with the taps flipped (correlation with the mirrored kernel)."""
        self.wt = {}
        self.wd = {}
        for k in (2, 3, 4):
            w = self.p["w%d" % k]
            self.wt[k] = w.to(torch.bfloat16).contiguous()
            self.wd[k] = w.flip(0).transpose(1, 2).to(torch.bfloat16).contiguous()
        self.w1e = (self.p["w1"] * self.scale).contiguous()
        self.w5e = (self.p["w5"] / self.scale).contiguous()
        self.b5 = float(self.p["b5"]) / self.scale

    def parameters(self):
        return [self.p[k] for k in self.NAMES]

    def refresh(self):
        self._prep()

    def forward(self, x, keep=False, fused=True):
        """y = x + prior(x) for x [B, H, W].  Without `keep` the whole network
        runs as one launch (rn_cnn_forward, no activations in memory); with
        `keep` (or fused=False) layer by layer, keeping the activations for
        backward.  Both give the same bits."""
        h = lib()
        x = x.contiguous()
        B, H, W = x.shape
        if fused and not keep:
            y = torch.empty_like(x)
            check(h.rn_cnn_forward(_p(x), _p(self.w1e), _p(self.p["b1"]), _p(self.wt[2]), _p(self.p["b2"]),
                                   _p(self.wt[3]), _p(self.p["b3"]), _p(self.wt[4]), _p(self.p["b4"]),
                                   _p(self.w5e), self.b5, _p(y), B, H, W, _st()), "rn_cnn_forward")
            self._cache = None
            return y
        acts = []
        a = torch.empty((B, H, W, 32), dtype=torch.bfloat16, device=x.device)
        check(h.rn_cnn_conv_in(_p(x), _p(self.w1e), _p(self.p["b1"]), _p(a), B, H, W, _st()), "rn_cnn_conv_in")
        acts.append(a)
        for k in (2, 3, 4):
            o = torch.empty_like(a)
            check(h.rn_cnn_conv_mid(_p(a), _p(self.wt[k]), _p(self.p["b%d" % k]), None, _p(o), B, H, W, 1, _st()),
                  "rn_cnn_conv_mid")
            a = o
            acts.append(a)
        y = torch.empty_like(x)
        check(h.rn_cnn_conv_out(_p(a), _p(self.w5e), self.b5, _p(x), _p(y), B, H, W, _st()), "rn_cnn_conv_out")
        if keep:
            self._cache = (x, acts)
        else:
            self._cache = None
        return y

    def backward(self, dy):
        """Parameter gradients of sum(dy * y) for the last forward(keep=True)."""
        h = lib()
        x, acts = self._cache
        dy = dy.contiguous().float()
        B, H, W = x.shape
        npx = B * H * W
        g = {}
        buf = torch.empty(9 * 32 + 32, dtype=torch.float32, device=x.device)
        check(h.rn_cnn_wgrad_narrow(_p(dy), _p(acts[3]), _p(buf), B, H, W, 1, _st()), "rn_cnn_wgrad_narrow")
        g["w5"] = buf[:288].view(9, 32) / self.scale
        g["b5"] = buf[288] / self.scale
        dz = torch.empty_like(acts[3])
        check(h.rn_cnn_out_dgrad(_p(dy), _p(self.w5e), _p(acts[3]), _p(dz), B, H, W, _st()), "rn_cnn_out_dgrad")
        for k in (4, 3, 2):
            a_in = acts[k - 2]
            gw = torch.empty(9, 32, 32, dtype=torch.float32, device=x.device)
            check(h.rn_cnn_wgrad_mid(_p(a_in), _p(dz), _p(gw), B, H, W, _st()), "rn_cnn_wgrad_mid")
            gb = torch.empty(32, dtype=torch.float32, device=x.device)
            check(h.rn_cnn_bias_grad(_p(dz), _p(gb), npx, _st()), "rn_cnn_bias_grad")
            g["w%d" % k], g["b%d" % k] = gw, gb
            nz = torch.empty_like(dz)
            check(h.rn_cnn_conv_mid(_p(dz), _p(self.wd[k]), None, _p(a_in), _p(nz), B, H, W, 0, _st()),
                  "rn_cnn_conv_mid")
            dz = nz
        buf = torch.empty(9 * 32 + 32, dtype=torch.float32, device=x.device)
        check(h.rn_cnn_wgrad_narrow(_p(x), _p(dz), _p(buf), B, H, W, 0, _st()), "rn_cnn_wgrad_narrow")
        g["w1"] = buf[:288].view(9, 32) * self.scale
        g["b1"] = buf[288:].clone()
        return g
