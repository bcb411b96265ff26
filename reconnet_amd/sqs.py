"""Device-resident OS-SQS(-TV) over a batch of 2D fan slices.

Wraps the C-ABI engine rn_sqs_* (ossqs.cu semantics with the Siddon fan
projector, plus the TotalVariation.cu surrogate as prior).  Subsets are
contiguous blocks of ceil(nview / M) views of the plan's angle order; with
``order="interleaved"`` the angles and sinograms are permuted so that those
blocks are the interleaved subsets {k, k+M, ...} (the reference's `order`
argument, ReconNet_no_tf.py:419-422).
"""

import ctypes
import math

import numpy as np
import torch

from ._lib import check, lib
from .geometry import FanGeometry, interleaved_order
from .torch_ops import FanPlan, _ptr, _stream

__all__ = ["OsSqsTV", "PassGraph", "fbp_initial_image"]


def fbp_initial_image(plan, sino, filter_type=2, chunk=32, out=None):
    """Non-negative FBP images [B, nx, ny] of sinograms [B, nu, nview] (natural
    view order): windowed ramp filter (2 = Hann) and the FBP-weighted DD
    backprojection, the usual starting image of the iterative methods."""
    B = sino.shape[0]
    g = plan.geo
    if out is None:
        out = torch.empty((B, g.nx, g.ny), dtype=torch.float32, device=sino.device)
    for i in range(0, B, chunk):
        out[i:i + chunk] = plan.fbp(sino[i:i + chunk].contiguous(), filter_type, "dd_fbp")
    return out.clamp_(min=0.0)


class OsSqsTV:
    def __init__(self, geo: FanGeometry, nsubsets, nslices, order="interleaved"):
        self.geo = geo
        self.nsubsets = int(nsubsets)
        self.nslices = int(nslices)
        if order == "interleaved":
            self.order = interleaved_order(geo.nview, self.nsubsets)
        elif order is None:
            self.order = np.arange(geo.nview)
        else:
            self.order = np.asarray(order)
        self.plan = FanPlan(geo, geo.angles()[self.order])
        self.npers = int(math.ceil(geo.nview / float(self.nsubsets)))
        self._h = lib().rn_sqs_create(self.plan.handle, self.nslices, self.npers, self.nsubsets)
        if not self._h:
            check(-1, "rn_sqs_create")
        self._order_t = None

    def close(self):
        if getattr(self, "_h", None):
            lib().rn_sqs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _reorder(self, sino):
        if self._order_t is None or self._order_t.device != sino.device:
            self._order_t = torch.as_tensor(self.order, dtype=torch.long, device=sino.device)
        return sino.index_select(2, self._order_t).contiguous()

    def set_data(self, sino, weights=None):
        """sino, weights: [nslices, nu, nview] float32 CUDA, natural view order."""
        s = self._reorder(sino)
        w = self._reorder(weights) if weights is not None else None
        check(lib().rn_sqs_set_data(self._h, _ptr(s), _ptr(w) if w is not None else None, _stream()),
              "rn_sqs_set_data")

    def compute_norm(self):
        check(lib().rn_sqs_compute_norm(self._h, _stream()), "rn_sqs_compute_norm")

    def set_norm(self, norm):
        check(lib().rn_sqs_set_norm(self._h, _ptr(norm.contiguous()), _stream()), "rn_sqs_set_norm")

    def norm(self):
        g = self.geo
        out = torch.empty((self.nslices, g.nx, g.ny), dtype=torch.float32, device="cuda")
        check(lib().rn_sqs_get_norm(self._h, _ptr(out), _stream()), "rn_sqs_get_norm")
        return out

    def set_image(self, x):
        x = x.contiguous()
        check(lib().rn_sqs_set_image(self._h, _ptr(x), _stream()), "rn_sqs_set_image")

    def image(self, out=None):
        g = self.geo
        if out is None:
            out = torch.empty((self.nslices, g.nx, g.ny), dtype=torch.float32, device="cuda")
        check(lib().rn_sqs_get_image(self._h, _ptr(out), _stream()), "rn_sqs_get_image")
        return out

    def set_profiling(self, on):
        check(lib().rn_sqs_set_profiling(self._h, 1 if on else 0), "rn_sqs_set_profiling")

    def kernel_times(self):
        """Total HIP-event ms of (FP+residual, BP, update) launches and the step count."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_long()
        check(lib().rn_sqs_kernel_times(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                        ctypes.byref(n)), "rn_sqs_kernel_times")
        return {"fp": a.value, "bp": b.value, "update": c.value, "steps": n.value}

    def run_pass(self, beta=0.0, eps=1e-6, lb=None, ub=None, lam=1.0, track=False):
        """One pass over all subsets; returns per-slice (|dx|^2, |x|^2) when track."""
        lb = float(np.finfo(np.float32).min) if lb is None else float(lb)
        ub = float(np.finfo(np.float32).max) if ub is None else float(ub)
        buf = None
        if track:
            buf = (ctypes.c_double * (2 * self.nslices))()
        check(lib().rn_sqs_pass(self._h, float(lam), float(beta), float(eps), lb, ub,
                                buf if track else None, _stream()), "rn_sqs_pass")
        if track:
            a = np.frombuffer(buf, dtype=np.float64).reshape(self.nslices, 2)
            return a[:, 0].copy(), a[:, 1].copy()
        return None

    def capture_passes(self, npasses, beta=0.0, eps=1e-6, lb=None, ub=None, lam=1.0):
        """Record ``npasses`` full passes into a HIP graph (see PassGraph)."""
        return PassGraph(self, npasses, beta, eps, lb, ub, lam)


class PassGraph:
    """npasses OS-SQS(-TV) passes replayed as one HIP graph on the engine's
    current image.  The engine ping-pongs between two image buffers once per
    subset step; the graph is only valid when the recorded step count is even,
    so that every replay starts from the buffer the host side points to."""

    def __init__(self, eng, npasses, beta, eps, lb, ub, lam):
        steps = -(-eng.geo.nview // eng.npers) * int(npasses)
        if steps % 2:
            raise ValueError("a pass graph needs an even number of subset steps (got %d)" % steps)
        self.eng = eng
        x_keep = eng.image()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            eng.run_pass(beta, eps, lb, ub, lam)   # warm-up outside the capture
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            for _ in range(int(npasses)):
                eng.run_pass(beta, eps, lb, ub, lam)
        torch.cuda.synchronize()
        eng.set_image(x_keep)          # undo the warm-up pass (capturing runs nothing)

    def replay(self):
        self.graph.replay()

