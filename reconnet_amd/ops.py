"""The projector pair as torch.library custom operators.

The reference registers its projectors as TensorFlow ops whose gradients call
the opposite projector (ReconNet4D/zOpsReconNet.cc:26-82,270-321,
python/ReconNet.py:773-830).  Here the same contract is expressed with
torch.library: ``reconnet::fan_forward`` / ``reconnet::fan_backward`` (2D fan,
images [B, nx, ny] <-> sinograms [B, nu, nview]) and ``reconnet::cone_forward``
/ ``reconnet::cone_backward`` (circular cone, reference 5-D layouts), each with
a fake (meta) implementation for shape propagation and an autograd rule whose
backward is the exact adjoint.  Plans are referred to by an integer handle from
``register_plan`` so that the operators take only tensors and scalars and can
sit inside traced / compiled graphs.
"""

import itertools

import torch

from .torch_ops import ADJOINT_OF, ConePlan, FanPlan

__all__ = ["register_plan", "release_plan", "fan_project", "fan_backproject", "cone_project", "cone_backproject"]

_PLANS = {}
_IDS = itertools.count(1)
_CONE_ADJOINT = {"dd": "dd", "siddon": "siddon"}
_FAN_BP_KINDS = ("dd", "dd_fbp", "dd_adjoint", "siddon", "pixel")


def register_plan(plan):
    """Make a FanPlan / ConePlan addressable from the operators; returns its handle."""
    if not isinstance(plan, (FanPlan, ConePlan)):
        raise TypeError("register_plan takes a FanPlan or a ConePlan")
    pid = next(_IDS)
    _PLANS[pid] = plan
    return pid


def release_plan(pid):
    _PLANS.pop(int(pid), None)


def _plan(pid, cls):
    p = _PLANS.get(int(pid))
    if not isinstance(p, cls):
        raise ValueError("no %s registered under handle %r" % (cls.__name__, pid))
    return p


# ------------------------------------------------------------------ fan --

@torch.library.custom_op("reconnet::fan_forward", mutates_args=())
def _fan_forward(img: torch.Tensor, plan_id: int, kind: str) -> torch.Tensor:
    return _plan(plan_id, FanPlan).forward(img.contiguous(), kind)


@_fan_forward.register_fake
def _(img, plan_id, kind):
    g = _plan(plan_id, FanPlan).geo
    return img.new_empty((img.shape[0], g.nu, g.nview))


@torch.library.custom_op("reconnet::fan_backward", mutates_args=())
def _fan_backward(sino: torch.Tensor, plan_id: int, kind: str) -> torch.Tensor:
    return _plan(plan_id, FanPlan).backward(sino.contiguous(), kind)


@_fan_backward.register_fake
def _(sino, plan_id, kind):
    g = _plan(plan_id, FanPlan).geo
    return sino.new_empty((sino.shape[0], g.nx, g.ny))


def _fan_fwd_setup(ctx, inputs, output):
    ctx.plan_id, ctx.kind = inputs[1], inputs[2]


def _fan_fwd_grad(ctx, g):
    if ctx.kind not in ADJOINT_OF:
        raise RuntimeError("projector kind %r has no adjoint" % ctx.kind)
    return _fan_backward(g, ctx.plan_id, ADJOINT_OF[ctx.kind]), None, None


def _fan_bwd_setup(ctx, inputs, output):
    ctx.plan_id, ctx.kind = inputs[1], inputs[2]


def _fan_bwd_grad(ctx, g):
    # the transpose of an exact adjoint is its forward projector
    fwd = {v: k for k, v in ADJOINT_OF.items()}
    if ctx.kind not in fwd:
        raise RuntimeError("backprojector kind %r is not the adjoint of a forward projector" % ctx.kind)
    return _fan_forward(g, ctx.plan_id, fwd[ctx.kind]), None, None


torch.library.register_autograd("reconnet::fan_forward", _fan_fwd_grad, setup_context=_fan_fwd_setup)
torch.library.register_autograd("reconnet::fan_backward", _fan_bwd_grad, setup_context=_fan_bwd_setup)


def fan_project(img, plan_id, kind="dd"):
    """y = A x for images [B, nx, ny] (float32, CUDA); differentiable."""
    return torch.ops.reconnet.fan_forward(img, int(plan_id), kind)


def fan_backproject(sino, plan_id, kind="dd_adjoint"):
    """x = A^T y (or another backprojector kind) for sinograms [B, nu, nview]."""
    if kind not in _FAN_BP_KINDS:
        raise ValueError("unknown backprojector kind %r" % kind)
    return torch.ops.reconnet.fan_backward(sino, int(plan_id), kind)


# ----------------------------------------------------------------- cone --

@torch.library.custom_op("reconnet::cone_forward", mutates_args=())
def _cone_forward(vol: torch.Tensor, plan_id: int, kind: str) -> torch.Tensor:
    return _plan(plan_id, ConePlan).forward(vol.contiguous(), kind)


@_cone_forward.register_fake
def _(vol, plan_id, kind):
    p = _plan(plan_id, ConePlan)
    return vol.new_empty((vol.shape[0], p.nu, p.nview, p.nv, vol.shape[4]))


@torch.library.custom_op("reconnet::cone_backward", mutates_args=())
def _cone_backward(prj: torch.Tensor, plan_id: int, kind: str) -> torch.Tensor:
    return _plan(plan_id, ConePlan).backward(prj.contiguous(), kind)


@_cone_backward.register_fake
def _(prj, plan_id, kind):
    p = _plan(plan_id, ConePlan)
    return prj.new_empty((prj.shape[0],) + tuple(p.shape_img) + (prj.shape[4],))


def _cone_setup(ctx, inputs, output):
    ctx.plan_id, ctx.kind = inputs[1], inputs[2]


torch.library.register_autograd("reconnet::cone_forward",
                                lambda ctx, g: (_cone_backward(g, ctx.plan_id, _CONE_ADJOINT[ctx.kind]), None, None),
                                setup_context=_cone_setup)
torch.library.register_autograd("reconnet::cone_backward",
                                lambda ctx, g: (_cone_forward(g, ctx.plan_id, ctx.kind), None, None),
                                setup_context=_cone_setup)


def cone_project(vol, plan_id, kind="dd"):
    return torch.ops.reconnet.cone_forward(vol, int(plan_id), kind)


def cone_backproject(prj, plan_id, kind="dd"):
    return torch.ops.reconnet.cone_backward(prj, int(plan_id), kind)
