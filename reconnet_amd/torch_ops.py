"""torch front end of the HIP projectors (device memory and streams only).

``FanPlan`` wraps one C-ABI plan (rn_fan_plan_create): geometry, angles, the
float32 X/Y split and workspaces.  Every call enqueues the library's kernels on
torch's current stream; inputs must be float32 CUDA tensors.

``FanProjection`` is the differentiable projector: forward = distance-driven
(or Siddon) FP, backward = its exact transpose (RN_BP_DD_ADJOINT /
RN_BP_SIDDON).  The reference pairs FP and BP the same way in its TensorFlow
gradients (python/ReconNet.py:773-830), with SiddonFan as the projector
(ReconNet4D/zOpsReconNet.cc:127-131).
"""

import ctypes

import numpy as np
import torch

from ._lib import check, lib
from .geometry import FanGeometry

__all__ = ["FanPlan", "FanProjection", "ConePlan", "ConeProjection", "FP_KINDS", "BP_KINDS", "CONE_KINDS"]

FP_KINDS = {"dd": 0, "siddon": 1}
BP_KINDS = {"dd": 0, "dd_fbp": 1, "dd_adjoint": 2, "siddon": 3, "pixel": 4}
ADJOINT_OF = {"dd": "dd_adjoint", "siddon": "siddon"}


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _check_tensor(t, shape, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("%s must be a CUDA tensor" % name)
    if t.dtype != torch.float32:
        raise TypeError("%s must be float32" % name)
    if tuple(t.shape) != tuple(shape):
        raise ValueError("%s has shape %s, expected %s" % (name, tuple(t.shape), tuple(shape)))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)


class FanPlan:
    """One fan geometry + angle set on the current CUDA device."""

    def __init__(self, geo: FanGeometry, angles=None):
        self.geo = geo
        self.angles = np.ascontiguousarray(geo.angles() if angles is None else angles, dtype=np.float32)
        if self.angles.shape[0] != geo.nview:
            raise ValueError("angle count %d != nview %d" % (self.angles.shape[0], geo.nview))
        g = geo
        self._h = lib().rn_fan_plan_create(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.nu, g.nview, g.nv, g.da, g.dv,
                                           g.off_a, g.off_v, g.dsd, g.dso,
                                           self.angles.ctypes.data_as(ctypes.c_void_p))
        if not self._h:
            check(-1, "rn_fan_plan_create")

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().rn_fan_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # shapes: images [B, nx, ny] (2D) or [B, nx, ny, nz, C]; sinograms [B, nu, nview] or [B, nu, nview, nv, C]
    def _img_shape(self, B, C=1, full=False):
        g = self.geo
        return (B, g.nx, g.ny, g.nz, C) if full else (B, g.nx, g.ny)

    def _sino_shape(self, B, C=1, full=False):
        g = self.geo
        return (B, g.nu, g.nview, g.nv, C) if full else (B, g.nu, g.nview)

    def forward(self, x, kind="dd", out=None):
        full = x.dim() == 5
        B, C = x.shape[0], (x.shape[4] if full else 1)
        _check_tensor(x, self._img_shape(B, C, full), "image")
        if out is None:
            out = torch.empty(self._sino_shape(B, C, full), dtype=torch.float32, device=x.device)
        check(lib().rn_fan_forward(self._h, FP_KINDS[kind], _ptr(x), _ptr(out), B, C, _stream()), "rn_fan_forward")
        return out

    def backward(self, y, kind="dd_adjoint", out=None):
        full = y.dim() == 5
        B, C = y.shape[0], (y.shape[4] if full else 1)
        _check_tensor(y, self._sino_shape(B, C, full), "sinogram")
        if out is None:
            out = torch.empty(self._img_shape(B, C, full), dtype=torch.float32, device=y.device)
        check(lib().rn_fan_backward(self._h, BP_KINDS[kind], _ptr(y), _ptr(out), B, C, _stream()), "rn_fan_backward")
        return out

    def filter(self, y, filter_type=0, fan=True, out=None):
        full = y.dim() == 5
        B, C = y.shape[0], (y.shape[4] if full else 1)
        _check_tensor(y, self._sino_shape(B, C, full), "sinogram")
        if out is None:
            out = torch.empty_like(y)
        check(lib().rn_fan_filter(self._h, _ptr(y), _ptr(out), B, C, int(filter_type), 1 if fan else 0, _stream()),
              "rn_fan_filter")
        return out

    def fbp(self, y, filter_type=0, bp="dd_fbp"):
        return self.backward(self.filter(y, filter_type), bp)


class FanProjection(torch.autograd.Function):
    """y = A x with A the DD (or Siddon) fan projector; dL/dx = A^T dL/dy."""

    @staticmethod
    def forward(ctx, x, plan, kind="dd"):
        ctx.plan = plan
        ctx.kind = kind
        return plan.forward(x.contiguous(), kind)

    @staticmethod
    def backward(ctx, gy):
        gx = ctx.plan.backward(gy.contiguous(), ADJOINT_OF[ctx.kind])
        return gx, None, None


CONE_KINDS = {"dd": 0, "siddon": 1}


def _cone_sigs():
    h = lib()
    if not getattr(h, "_cone_sigs", False):
        ci, cf, P = ctypes.c_int, ctypes.c_float, ctypes.c_void_p
        h.rn_cone_plan_create.restype = P
        h.rn_cone_plan_create.argtypes = [ci, ci, ci, cf, cf, cf, cf, cf, cf, ci, ci, ci, cf, cf, cf, cf, P, P, P, P]
        h.rn_cone_plan_destroy.restype = None
        h.rn_cone_plan_destroy.argtypes = [P]
        for n in ("rn_cone_forward", "rn_cone_backward"):
            getattr(h, n).restype = ci
            getattr(h, n).argtypes = [P, ci, P, P, ci, ci, P]
        h._cone_sigs = True
    return h


class ConePlan:
    """Circular cone-beam geometry on the current CUDA device.

    Volumes [B, nx, ny, nz, C] and projections [B, nu, nview, nv, C] (the
    reference layouts); det_center / det_u / det_v / src are [nview, 3] arrays as
    for cSiddonConeProjectionAbitrary.  kind "dd": distance-driven pair whose
    backward is the exact adjoint; kind "siddon": the Siddon pair."""

    def __init__(self, nx, ny, nz, dx, dy, dz, nu, nv, du, dv, det_center, det_u, det_v, src,
                 cx=0.0, cy=0.0, cz=0.0, off_u=0.0, off_v=0.0):
        h = _cone_sigs()
        self.shape_img = (nx, ny, nz)
        self.nu, self.nv = nu, nv
        vecs = [np.ascontiguousarray(a, np.float32).reshape(-1, 3) for a in (det_center, det_u, det_v, src)]
        self.nview = vecs[0].shape[0]
        self._vecs = vecs
        self._h = h.rn_cone_plan_create(nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, self.nview, nv, du, dv, off_u, off_v,
                                        *[v.ctypes.data_as(ctypes.c_void_p) for v in vecs])
        if not self._h:
            check(-1, "rn_cone_plan_create")

    def close(self):
        if getattr(self, "_h", None):
            lib().rn_cone_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, x, kind="dd"):
        B, C = x.shape[0], x.shape[4]
        _check_tensor(x, (B,) + self.shape_img + (C,), "volume")
        out = torch.empty((B, self.nu, self.nview, self.nv, C), dtype=torch.float32, device=x.device)
        check(lib().rn_cone_forward(self._h, CONE_KINDS[kind], _ptr(x), _ptr(out), B, C, _stream()), "rn_cone_forward")
        return out

    def backward(self, y, kind="dd"):
        B, C = y.shape[0], y.shape[4]
        _check_tensor(y, (B, self.nu, self.nview, self.nv, C), "projections")
        out = torch.empty((B,) + self.shape_img + (C,), dtype=torch.float32, device=y.device)
        check(lib().rn_cone_backward(self._h, CONE_KINDS[kind], _ptr(y), _ptr(out), B, C, _stream()),
              "rn_cone_backward")
        return out


class ConeProjection(torch.autograd.Function):
    """y = A x with A the cone DD (or Siddon) projector; dL/dx = A^T dL/dy."""

    @staticmethod
    def forward(ctx, x, plan, kind="dd"):
        ctx.plan, ctx.kind = plan, kind
        return plan.forward(x.contiguous(), kind)

    @staticmethod
    def backward(ctx, gy):
        return ctx.plan.backward(gy.contiguous(), ctx.kind), None, None
