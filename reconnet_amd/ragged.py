"""Differentiable projector over a batch of fan slices of uneven size.

``RaggedFanBatch`` wraps rn_ragged_* : one persistent launch projects (or
backprojects) every slice of the batch, each with its own image size, detector
and angle set.  ``RaggedProjection`` is the torch.autograd.Function whose
backward is the exact transpose of the forward (DD -> RN_BP_DD_ADJOINT,
Siddon -> RN_BP_SIDDON), as the reference's TensorFlow gradient pairs FP and BP
(python/ReconNet.py:773-830).
"""

import ctypes

import numpy as np
import torch

from ._lib import check, lib
from .torch_ops import _ptr, _stream

__all__ = ["RaggedFanBatch", "RaggedProjection"]

_KIND_FP = {"dd": 0, "siddon": 1}
_KIND_BP = {"dd": 2, "siddon": 3}


def _sigs():
    h = lib()
    if not getattr(h, "_ragged_sigs", False):
        P, ci = ctypes.c_void_p, ctypes.c_int
        h.rn_ragged_create.restype = P
        h.rn_ragged_create.argtypes = [ci, P, P, P]
        h.rn_ragged_destroy.restype = None
        h.rn_ragged_destroy.argtypes = [P]
        for n in ("rn_ragged_forward", "rn_ragged_backward"):
            getattr(h, n).restype = ci
            getattr(h, n).argtypes = [P, ci, P, P, P]
        h._ragged_sigs = True
    return h


class RaggedFanBatch:
    def __init__(self, geos, angles=None):
        h = _sigs()
        self.geos = list(geos)
        self.angles = [np.ascontiguousarray(g.angles() if angles is None else a, np.float32)
                       for g, a in zip(self.geos, angles if angles is not None else [None] * len(self.geos))]
        dims = np.array([[g.nx, g.ny, g.nu, g.nview] for g in self.geos], np.int32)
        geo = np.array([[g.dx, g.dy, g.da, g.off_a, g.dsd, g.dso] for g in self.geos], np.float32)
        ang = np.ascontiguousarray(np.concatenate(self.angles), np.float32)
        self._h = h.rn_ragged_create(len(self.geos), dims.ctypes.data_as(ctypes.c_void_p),
                                     geo.ctypes.data_as(ctypes.c_void_p), ang.ctypes.data_as(ctypes.c_void_p))
        if not self._h:
            check(-1, "rn_ragged_create")
        self.img_sizes = [g.nx * g.ny for g in self.geos]
        self.sino_sizes = [g.nu * g.nview for g in self.geos]

    def close(self):
        if getattr(self, "_h", None):
            lib().rn_ragged_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_image(self):
        return int(sum(self.img_sizes))

    @property
    def n_sino(self):
        return int(sum(self.sino_sizes))

    def flatten_images(self, imgs):
        return torch.cat([t.reshape(-1) for t in imgs]).contiguous()

    def split_images(self, flat):
        out, o = [], 0
        for g, n in zip(self.geos, self.img_sizes):
            out.append(flat[o:o + n].view(g.nx, g.ny))
            o += n
        return out

    def split_sinos(self, flat):
        out, o = [], 0
        for g, n in zip(self.geos, self.sino_sizes):
            out.append(flat[o:o + n].view(g.nu, g.nview))
            o += n
        return out

    def forward(self, img_flat, kind="dd"):
        img_flat = img_flat.contiguous()
        out = torch.empty(self.n_sino, dtype=torch.float32, device=img_flat.device)
        check(lib().rn_ragged_forward(self._h, _KIND_FP[kind], _ptr(img_flat), _ptr(out), _stream()),
              "rn_ragged_forward")
        return out

    def backward(self, sino_flat, kind="dd"):
        sino_flat = sino_flat.contiguous()
        out = torch.empty(self.n_image, dtype=torch.float32, device=sino_flat.device)
        check(lib().rn_ragged_backward(self._h, _KIND_BP[kind], _ptr(sino_flat), _ptr(out), _stream()),
              "rn_ragged_backward")
        return out


class RaggedProjection(torch.autograd.Function):
    """Flat concatenated images -> flat concatenated sinograms, y = A x."""

    @staticmethod
    def forward(ctx, img_flat, batch, kind="dd"):
        ctx.batch, ctx.kind = batch, kind
        return batch.forward(img_flat, kind)

    @staticmethod
    def backward(ctx, g):
        return ctx.batch.backward(g, ctx.kind), None, None
