"""Differentiable projectors: torch.autograd over the hand-written exact
adjoints, single geometry and uneven-size persistent batches.

Batch: 16 slices, six 256^2 / 360 views, six 384^2 / 720 views, four 512^2 /
1152 views (736 bins).  Loss 0.5 ||Ax - y||^2 in FP32.  Checks: per-slice
results of the persistent launch equal independent single-geometry launches
bit for bit; the gradient equals A^T(Ax - y) from the FP64 oracle (Siddon,
where the oracle holds the exact transpose; DD through <g, z> = <r, A z> with
the oracle's forward operator); central finite differences (h = 1e-3) on 64
random pixels, for both.
"""

import numpy as np
import pytest
import torch

import oracle as orc
from reconnet_amd.geometry import FanGeometry
from reconnet_amd.phantom import phantom
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _uneven_geos():
    geos = []
    for n, nx, nview, dx in [(6, 256, 360, 1.4), (6, 384, 720, 0.93), (4, 512, 1152, 0.7)]:
        geos += [FanGeometry(nx=nx, ny=nx, dx=dx, dy=dx, nu=736, nview=nview, da=1.0 / 1085.6, dsd=1085.6,
                             dso=595.0)] * n
    return geos


@pytest.fixture(scope="module")
def batch():
    from reconnet_amd.ragged import RaggedFanBatch

    geos = _uneven_geos()
    rb = RaggedFanBatch(geos)
    imgs = [torch.from_numpy(phantom(s, g.nx, g.ny, g.dx, g.dy)).cuda() for s, g in enumerate(geos)]
    return geos, rb, imgs


@pytest.mark.parametrize("kind", ["dd", "siddon"])
def test_ragged_equals_per_slice_launches(batch, kind):
    from reconnet_amd.torch_ops import FanPlan

    geos, rb, imgs = batch
    y = rb.split_sinos(rb.forward(rb.flatten_images(imgs), kind))
    bp = rb.split_images(rb.backward(torch.cat([t.reshape(-1) for t in y]), kind))
    adj = {"dd": "dd_adjoint", "siddon": "siddon"}[kind]
    for s in (0, 6, 12, 15):
        plan = FanPlan(geos[s])
        ref = plan.forward(imgs[s][None].contiguous(), kind)[0]
        assert torch.equal(y[s], ref)
        refb = plan.backward(y[s][None].contiguous(), adj)[0]
        assert torch.equal(bp[s], refb)


def test_gradient_matches_oracle_and_finite_differences(batch):
    from reconnet_amd.ragged import RaggedProjection

    geos, rb, imgs = batch
    rng = np.random.default_rng(42)
    x0 = rb.flatten_images(imgs)
    meas = rb.forward(x0, "siddon") * 1.01 + 0.01
    x = (x0 + 0.002 * torch.from_numpy(rng.normal(size=x0.numel()).astype(np.float32)).cuda()).requires_grad_(True)
    r = RaggedProjection.apply(x, rb, "siddon") - meas
    loss = 0.5 * (r * r).sum()
    loss.backward()
    g = x.grad.detach()
    # FP64 oracle gradient A^T (A x - y) on two slices of different size
    xs = rb.split_images(x.detach())
    ms = rb.split_sinos(meas)
    gs = rb.split_images(g)
    for s in (0, 12):
        geo = geos[s]
        xa = xs[s].cpu().numpy()[None, :, :, None, None]
        ax = orc.siddon_fan_fp(xa, geo.angles(), geo.oracle_geo(), geo.nu)
        res = (ax - ms[s].cpu().numpy().astype(np.float64)[None, :, :, None, None]).astype(np.float32)
        ref = orc.siddon_fan_bp(res, geo.angles(), geo.oracle_geo(), (geo.nx, geo.ny, 1))[0, :, :, 0, 0]
        assert rel_l2(gs[s].cpu().numpy(), ref) <= 1e-4
    # central finite differences, h = 1e-3, on 64 random pixels
    idx = rng.choice(x0.numel(), size=64, replace=False)
    h = 1e-3
    with torch.no_grad():
        base = x.detach().clone()
        fd = []
        for i in idx:
            xp = base.clone(); xp[i] += h
            xm = base.clone(); xm[i] -= h
            lp = 0.5 * ((rb.forward(xp, "siddon") - meas).double() ** 2).sum()
            lm = 0.5 * ((rb.forward(xm, "siddon") - meas).double() ** 2).sum()
            fd.append(float((lp - lm) / (2 * h)))
    fd = np.array(fd)
    ga = g.cpu().numpy()[idx].astype(np.float64)
    assert rel_l2(ga, fd) <= 1e-2


def test_dd_gradient_through_uneven_batch_matches_oracle_and_finite_differences(batch):
    # DD: the gradient A^T (A x - y) of the persistent launch, against the FP64
    # oracle's DD projector through <g, z> = <r, A z> for random z (the adjoint
    # must be the transpose of the oracle's forward operator), and central
    # finite differences of the loss on 64 random pixels
    from reconnet_amd.ragged import RaggedProjection

    geos, rb, imgs = batch
    rng = np.random.default_rng(43)
    x0 = rb.flatten_images(imgs)
    meas = rb.forward(x0, "dd") * 1.01 + 0.01
    x = (x0 + 0.002 * torch.from_numpy(rng.normal(size=x0.numel()).astype(np.float32)).cuda()).requires_grad_(True)
    r = RaggedProjection.apply(x, rb, "dd") - meas
    loss = 0.5 * (r * r).sum()
    loss.backward()
    g = x.grad.detach()
    rs = rb.split_sinos(r.detach())
    gs = rb.split_images(g)
    for s in (0, 7, 12):
        geo = geos[s]
        z = rng.uniform(-1, 1, (geo.nx, geo.ny)).astype(np.float32)
        az = orc.dd_fan_fp(z[None, :, :, None, None], geo.angles(), geo.oracle_geo(), geo.nu)[0, :, :, 0, 0]
        lhs = float((gs[s].cpu().numpy().astype(np.float64) * z).sum())
        rhs = float((rs[s].cpu().numpy().astype(np.float64) * az).sum())
        assert abs(lhs - rhs) <= 1e-4 * np.sqrt((gs[s].double() ** 2).sum().item() * (z.astype(np.float64) ** 2).sum())
    idx = rng.choice(x0.numel(), size=64, replace=False)
    h = 1e-3
    with torch.no_grad():
        base = x.detach().clone()
        fd = []
        for i in idx:
            xp = base.clone(); xp[i] += h
            xm = base.clone(); xm[i] -= h
            lp = 0.5 * ((rb.forward(xp, "dd") - meas).double() ** 2).sum()
            lm = 0.5 * ((rb.forward(xm, "dd") - meas).double() ** 2).sum()
            fd.append(float((lp - lm) / (2 * h)))
    fd = np.array(fd)
    ga = g.cpu().numpy()[idx].astype(np.float64)
    assert np.abs(ga).max() > 1e-3
    assert rel_l2(ga, fd) <= 1e-2


def test_dd_autograd_single_geometry_dot_identity():
    from reconnet_amd.torch_ops import FanPlan, FanProjection

    geo = FanGeometry(nx=256, ny=256, dx=1.0, dy=1.0, nu=512, nview=360, da=1.2 / 1085.6, dsd=1085.6, dso=595.0)
    plan = FanPlan(geo)
    g = torch.Generator(device="cpu").manual_seed(0)
    x = torch.rand(2, 256, 256, generator=g).cuda().requires_grad_(True)
    v = torch.rand(2, 512, 360, generator=g).cuda()
    y = FanProjection.apply(x, plan, "dd")
    (y * v).sum().backward()
    lhs = float((y.detach().double() * v.double()).sum())
    rhs = float((x.detach().double() * x.grad.double()).sum())
    assert abs(lhs - rhs) / abs(lhs) <= 1e-5
