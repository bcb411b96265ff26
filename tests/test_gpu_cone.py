"""Circular cone-beam projectors on the GPU: the DD pair against the FP64
oracle, its exact adjoint (dot test at the 256^2 x 128 volume, 64 x 512
detector, 720-view size), the Siddon pair's dot test at that size, and the DD
forward against the Siddon forward on a refined detector (the cell average
of the line integrals) on a smooth volume."""

import numpy as np
import pytest
import torch

import oracle as orc
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def circ_cone(nview, dso, dsd, th0=0.0):
    th = (np.arange(nview, dtype=np.float32) * np.float32(2 * np.pi) / np.float32(nview)).astype(np.float64) + th0
    src = np.stack([dso * np.sin(th), -dso * np.cos(th), np.zeros_like(th)], 1)
    ctr = np.stack([-(dsd - dso) * np.sin(th), (dsd - dso) * np.cos(th), np.zeros_like(th)], 1)
    du = np.stack([np.cos(th), np.sin(th), np.zeros_like(th)], 1)
    dv = np.tile([0.0, 0.0, 1.0], (nview, 1))
    return [a.astype(np.float32) for a in (ctr, du, dv, src)]


def _plan(n, nz, nu, nv, du, dv, vecs, **kw):
    from reconnet_amd.torch_ops import ConePlan

    return ConePlan(n, n, nz, 1.0, 1.0, 1.0, nu, nv, du, dv, *vecs, **kw)


@pytest.mark.parametrize("nC,off,du,dv", [(1, 0.0, 1.3, 1.1), (2, 0.3, 1.3, 1.1), (1, 0.2, 2.6, 2.3),
                                          (2, -0.1, 1.3, 2.3)])
def test_dd_cone_matches_oracle(nC, off, du, dv):
    # cells of 1.3 x 1.1 project to footprints under one voxel on every plane
    # (the forward's 2 x 2-cell path); 2.6 x 2.3 and 1.3 x 2.3 to larger ones
    # (its general loop)
    n, nz, nu, nv, nview = 24, 12, 40, 20, 30
    vecs = circ_cone(nview, 120.0, 200.0, th0=0.2)
    geo = dict(dx=1.0, dy=1.0, dz=1.0, du=du, dv=dv, off_u=off, off_v=-off, cx=0.4, cy=-0.3, cz=0.2)
    plan = _plan(n, nz, nu, nv, du, dv, vecs, off_u=off, off_v=-off, cx=0.4, cy=-0.3, cz=0.2)
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 1, (1, n, n, nz, nC)).astype(np.float32)
    y = plan.forward(torch.from_numpy(x).cuda(), "dd").cpu().numpy()
    assert rel_l2(y, orc.dd_cone_fp(x, *vecs, geo, nu, nv)) <= TOL
    p = rng.uniform(0, 1, (1, nu, nview, nv, nC)).astype(np.float32)
    b = plan.backward(torch.from_numpy(p).cuda(), "dd").cpu().numpy()
    assert rel_l2(b, orc.dd_cone_adj(p, *vecs, geo, (n, n, nz))) <= TOL


@pytest.mark.parametrize("kind,kernel,ncand", [("dd", "dd_cone_fp", 2), ("dd", "dd_cone_adj", 6),
                                               ("siddon", "siddon_cone_bp", 3)])
def test_every_cone_launch_shape_matches_oracle(kind, kernel, ncand):
    # each launch shape of the tuned cone kernels (XCD-ordered or plain grid;
    # 1, 2 or 4 voxels per lane stacked in z, with a z extent that leaves the
    # last lane's stack partly outside the volume; the DD adjoint's row loop
    # and row-sum table), two slices of two channels and an off-centre grid,
    # against the FP64 oracle
    from reconnet_amd._lib import autotune_pin

    n, nz, nu, nv, nview = 20, 13, 36, 18, 24
    vecs = circ_cone(nview, 120.0, 200.0, th0=0.1)
    geo = dict(dx=1.0, dy=1.0, dz=1.0, du=1.3, dv=1.1, off_u=0.2, off_v=-0.3, cx=0.4, cy=-0.3, cz=0.2)
    plan = _plan(n, nz, nu, nv, 1.3, 1.1, vecs, off_u=0.2, off_v=-0.3, cx=0.4, cy=-0.3, cz=0.2)
    rng = np.random.default_rng(21)
    x = rng.uniform(0, 1, (2, n, n, nz, 2)).astype(np.float32)
    p = rng.uniform(0, 1, (2, nu, nview, nv, 2)).astype(np.float32)
    if kernel == "dd_cone_fp":
        ref, run = orc.dd_cone_fp(x, *vecs, geo, nu, nv), lambda: plan.forward(torch.from_numpy(x).cuda(), kind)
    else:
        bp = orc.dd_cone_adj if kind == "dd" else orc.siddon_cone_bp
        ref, run = bp(p, *vecs, geo, (n, n, nz)), lambda: plan.backward(torch.from_numpy(p).cuda(), kind)
    assert np.abs(ref).max() > 1.0
    try:
        for c in range(ncand):
            autotune_pin(kernel, c)
            assert rel_l2(run().cpu().numpy(), ref) <= TOL, c
        autotune_pin(kernel, ncand)
        with pytest.raises(RuntimeError, match="out of range"):
            run()
    finally:
        autotune_pin(kernel, None)


@pytest.mark.parametrize("kind", ["dd", "siddon"])
def test_cone_full_size_views_match_oracle(kind):
    # the 256^2 x 128 volume at 1 mm (seeded ellipsoid phantom), 64 x 512
    # detector of 1.0 x 1.2 mm cells, SID 595, SDD 1085.6: 6 of the 720 views
    # (every 120th, both view classes), forward and backprojection against the
    # FP64 oracle
    from reconnet_amd.phantom import phantom3d

    n, nz, nu, nv = 256, 128, 512, 64
    vecs = [a[::120].copy() for a in circ_cone(720, 595.0, 1085.6)]
    plan = _plan(n, nz, nu, nv, 1.2, 1.0, vecs)
    x = phantom3d(0, n, n, nz, 1.0, 1.0, 1.0)[None, ..., None].astype(np.float32)
    y = plan.forward(torch.from_numpy(x).cuda(), kind).cpu().numpy()
    geo = dict(dx=1.0, dy=1.0, dz=1.0, du=1.2, dv=1.0)
    fp = orc.dd_cone_fp if kind == "dd" else orc.siddon_cone_fp
    ref = fp(x, *vecs, geo, nu, nv)
    assert np.abs(ref).max() > 1.0
    assert rel_l2(y, ref) <= TOL
    p = np.random.default_rng(12).uniform(0, 1, (1, nu, 6, nv, 1)).astype(np.float32)
    b = plan.backward(torch.from_numpy(p).cuda(), kind).cpu().numpy()
    bp = orc.dd_cone_adj if kind == "dd" else orc.siddon_cone_bp
    assert rel_l2(b, bp(p, *vecs, geo, (n, n, nz))) <= TOL


@pytest.mark.parametrize("kind", ["dd", "siddon"])
def test_cone_dot_identity_full_size(kind):
    # 256^2 x 128 volume at 1 mm, 64 rows x 512 columns of 1.0 x 1.2 mm, 720 views,
    # SID 595 mm, SDD 1085.6 mm; x, y ~ U[0, 1] with seeds 1234 and 5678
    n, nz, nu, nv, nview = 256, 128, 512, 64, 720
    plan = _plan(n, nz, nu, nv, 1.2, 1.0, circ_cone(nview, 595.0, 1085.6))
    gx = torch.Generator(device="cuda").manual_seed(1234)
    gy = torch.Generator(device="cuda").manual_seed(5678)
    x = torch.rand((1, n, n, nz, 1), generator=gx, device="cuda")
    y = torch.rand((1, nu, nview, nv, 1), generator=gy, device="cuda")
    ax = plan.forward(x, kind)
    aty = plan.backward(y, kind)
    lhs = float((ax.double() * y.double()).sum())
    rhs = float((x.double() * aty.double()).sum())
    assert abs(lhs - rhs) / abs(lhs) <= TOL


def test_dd_cone_close_to_refined_siddon():
    n, nz, nu, nv, nview, ss = 64, 32, 96, 24, 36, 4
    vecs = circ_cone(nview, 300.0, 500.0, th0=0.1)
    plan = _plan(n, nz, nu, nv, 1.6, 1.6, vecs)
    fine = _plan(n, nz, nu * ss, nv * ss, 1.6 / ss, 1.6 / ss, vecs)
    r = np.arange(n) - (n - 1) / 2
    rz = np.arange(nz) - (nz - 1) / 2
    X, Y, Z = np.meshgrid(r, r, rz, indexing="ij")
    vol = np.exp(-(X ** 2 + Y ** 2) / (2 * 8.0 ** 2) - Z ** 2 / (2 * 5.0 ** 2)).astype(np.float32)[None, ..., None]
    v = torch.from_numpy(vol).cuda()
    dd = plan.forward(v, "dd").cpu().numpy()[0, :, :, :, 0]
    sf = fine.forward(v, "siddon").cpu().numpy()[0, :, :, :, 0]
    avg = sf.reshape(nu, ss, nview, nv, ss).mean(axis=(1, 4))
    assert rel_l2(dd, avg) <= 5e-3


def test_cone_plan_rejects_tilted_detector():
    vecs = circ_cone(8, 300.0, 500.0)
    vecs[2] = vecs[2] + np.float32(0.1) * vecs[1]     # detector columns no longer parallel to z
    with pytest.raises(RuntimeError, match="parallel to z"):
        _plan(16, 8, 20, 10, 1.0, 1.0, vecs)
