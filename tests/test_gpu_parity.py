"""GPU parity of the HIP kernels against the FP64 CPU oracle.

All calls go through the C-ABI of libReconNet4D_no_tf.so: the reference
exports (host pointers) via reconnet_amd.recon.ReconNet, the device-pointer
entry points via reconnet_amd.torch_ops.  Bar: relative L2 <= 1e-5 against the
FP64 oracle for every FP32 projector, filter and update; adjoint dot-product
gap <= 1e-5 for the exact transposes.
"""

import numpy as np
import pytest

import oracle as orc
from reconnet_amd.geometry import WORKLOADS, FanGeometry, fan_angles
from reconnet_amd.phantom import phantom
from tests.helpers import recon_for, rel_l2, small_fan

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _img5(a):
    return np.asarray(a, np.float32)[None, :, :, None, None]


@pytest.fixture(scope="module")
def c0():
    geo = WORKLOADS["fan256"]
    img = phantom(0, geo.nx, geo.ny, geo.dx, geo.dy)
    rng = np.random.default_rng(1234)
    sino = rng.uniform(0, 1, (1, geo.nu, geo.nview, 1, 1)).astype(np.float32)
    return geo, _img5(img), sino, fan_angles(geo.nview)


# ------------------------------------------------------------ config 0 ------

def test_dd_fp_config0(c0):
    geo, img, _, deg = c0
    out = recon_for(geo).cDDFanProjection4d(img)
    ref = orc.dd_fan_fp(img, deg, geo.oracle_geo(), geo.nu)
    assert rel_l2(out, ref) <= TOL


def test_dd_fp_batch_interleaved_config0(c0):
    # 5 slices of the config-0 geometry take the DD FP on padded quads of the
    # averaged planes (one full and one partial group of 4): slices 0 and 4
    # against the oracle
    import torch
    from reconnet_amd.torch_ops import FanPlan

    geo, _, _, deg = c0
    imgs = np.stack([phantom(40 + s, geo.nx, geo.ny, geo.dx, geo.dy) for s in range(5)]).astype(np.float32)
    out = FanPlan(geo).forward(torch.from_numpy(imgs).cuda(), "dd").cpu().numpy()
    for s in (0, 4):
        ref = orc.dd_fan_fp(_img5(imgs[s]), deg, geo.oracle_geo(), geo.nu)[0, :, :, 0, 0]
        assert rel_l2(out[s], ref) <= TOL, s


def test_siddon_fp_config0(c0):
    geo, img, _, deg = c0
    out = recon_for(geo).cSiddonFanProjection4d(img)
    assert rel_l2(out, orc.siddon_fan_fp(img, deg, geo.oracle_geo(), geo.nu)) <= TOL


@pytest.mark.parametrize("type_projector", [0, 1])
def test_dd_bp_config0(c0, type_projector):
    geo, _, sino, deg = c0
    out = recon_for(geo).cDDFanBackprojection4d(sino, type_projector=type_projector)
    ref = orc.dd_fan_bp(sino, deg, geo.oracle_geo(), (geo.nx, geo.ny, 1), type_projector=type_projector)
    assert rel_l2(out, ref) <= TOL


@pytest.mark.parametrize("type_projector", [0, 1])
def test_dd_bp_batch_interleaved(type_projector):
    # 5 slices in one call take the interleaved 4-slice DD BP (one full and
    # one padded group): each slice against the oracle
    geo = small_fan(nx=72, nu=130, nview=90)
    deg = fan_angles(geo.nview)
    sino = np.random.default_rng(8).uniform(0, 1, (5, geo.nu, geo.nview, 1, 1)).astype(np.float32)
    out = recon_for(geo).cDDFanBackprojection4d(sino, type_projector=type_projector)
    ref = orc.dd_fan_bp(sino, deg, geo.oracle_geo(), (geo.nx, geo.ny, 1), type_projector=type_projector)
    for s in range(5):
        assert rel_l2(out[s], ref[s]) <= TOL, s


def test_siddon_bp_config0(c0):
    geo, _, sino, deg = c0
    out = recon_for(geo).cSiddonFanBackprojection4d(sino)
    assert rel_l2(out, orc.siddon_fan_bp(sino, deg, geo.oracle_geo(), (geo.nx, geo.ny, 1))) <= TOL


def test_pixel_bp_config0(c0):
    geo, _, sino, deg = c0
    out = recon_for(geo).cPixelBackprojection4d(sino)
    assert rel_l2(out, orc.pixel_fan_bp(sino, deg, geo.oracle_geo(), (geo.nx, geo.ny, 1))) <= TOL


@pytest.mark.parametrize("ftype", [0, 1, 2, 3])
def test_filter_config0(c0, ftype):
    geo, img, _, deg = c0
    sino = orc.dd_fan_fp(img, deg, geo.oracle_geo(), geo.nu).astype(np.float32)
    out = recon_for(geo).cFilter4d(sino, type_filter=ftype)
    assert rel_l2(out, orc.ramp_filter(sino, geo.oracle_geo(), filter_type=ftype)) <= TOL


@pytest.mark.parametrize("nu,ftype", [(1, 0), (37, 2), (130, 0), (736, 2), (1107, 1)])
def test_filter_row_widths(nu, ftype):
    # row widths not a multiple of 4 (padded tables, scalar stores), shorter than
    # the near band, the OS-SQS bench width, and wider than one 256-lane pass
    # (1107 bins: 277 lanes of 4 outputs); smooth rows plus noise, 23 or 24 views
    # (an odd row count leaves the last row pair of a workgroup step half empty)
    nview = 23 if nu % 2 else 24
    geo = FanGeometry(nx=64, ny=64, dx=1.0, dy=1.0, nu=nu, nview=nview, da=1.0 / 1085.6, dsd=1085.6, dso=595.0)
    rng = np.random.default_rng(nu)
    u = (np.arange(nu) - (nu - 1) / 2) / max(nu / 2.2, 1.0)
    base = np.sqrt(np.clip(1 - u * u, 0, None)) * 40
    sino = (base[:, None] * rng.uniform(0.8, 1.2, (1, geo.nview)) + rng.normal(0, 0.3, (nu, geo.nview)))
    sino = sino.astype(np.float32)[None, :, :, None, None]
    out = recon_for(geo).cFilter4d(sino, type_filter=ftype)
    assert rel_l2(out, orc.ramp_filter(sino, geo.oracle_geo(), filter_type=ftype)) <= TOL


def test_parallel_filter(c0):
    geo, img, _, deg = c0
    sino = orc.dd_fan_fp(img, deg, geo.oracle_geo(), geo.nu).astype(np.float32)
    r = recon_for(geo)
    r.du = 1.2
    out = r.cFilterParallel4d(sino, type_filter=2)
    g = dict(geo.oracle_geo(), da=1.2)
    assert rel_l2(out, orc.ramp_filter(sino, g, filter_type=2, fan=False)) <= TOL


def test_fbp_config0_end_to_end(c0):
    geo, img, _, deg = c0
    r = recon_for(geo)
    sino = r.cDDFanProjection4d(img)
    rec = r.cDDFanBackprojection4d(r.cFilter4d(sino), type_projector=1)
    ref_s = orc.dd_fan_fp(img, deg, geo.oracle_geo(), geo.nu)
    ref = orc.dd_fan_bp(orc.ramp_filter(ref_s.astype(np.float32), geo.oracle_geo()).astype(np.float32), deg,
                        geo.oracle_geo(), (geo.nx, geo.ny, 1), type_projector=1)
    assert rel_l2(rec, ref) <= 1e-4
    # the reconstruction itself reproduces the phantom inside the FOV
    m = geo.fov_mask()
    assert rel_l2(rec[0, :, :, 0, 0][m], img[0, :, :, 0, 0][m]) < 0.2


def _dot_gap(fp, bp, geo, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (1, geo.nx, geo.ny, 1, 1)).astype(np.float32)
    y = rng.uniform(0, 1, (1, geo.nu, geo.nview, 1, 1)).astype(np.float32)
    ax = fp(x).astype(np.float64)
    aty = bp(y).astype(np.float64)
    lhs, rhs = np.sum(ax * y), np.sum(x * aty)
    return abs(lhs - rhs) / abs(lhs)


def test_siddon_adjoint_dot_config0(c0):
    geo = c0[0]
    r = recon_for(geo)
    assert _dot_gap(r.cSiddonFanProjection4d, r.cSiddonFanBackprojection4d, geo) <= TOL


def test_dd_exact_adjoint_dot_config0(c0):
    import torch
    from reconnet_amd.torch_ops import FanPlan

    geo = c0[0]
    plan = FanPlan(geo)

    def fp(x):
        return plan.forward(torch.from_numpy(x[..., 0, 0]).cuda(), "dd").cpu().numpy()[..., None, None]

    def bp(y):
        return plan.backward(torch.from_numpy(y[..., 0, 0]).cuda(), "dd_adjoint").cpu().numpy()[..., None, None]

    assert _dot_gap(fp, bp, geo) <= TOL


def test_dd_exact_adjoint_matches_oracle_transpose(c0):
    # the oracle's DD FP as an explicit matrix on a small grid; its transpose
    # applied to y must equal the GPU adjoint
    import torch
    from reconnet_amd.torch_ops import FanPlan

    geo = small_fan(nx=12, nu=24, nview=20, du=3.0)
    deg = fan_angles(geo.nview)
    cols = []
    for k in range(geo.nx * geo.ny):
        e = np.zeros(geo.nx * geo.ny, np.float32)
        e[k] = 1
        cols.append(orc.dd_fan_fp(e.reshape(1, geo.nx, geo.ny, 1, 1), deg, geo.oracle_geo(), geo.nu).ravel())
    A = np.stack(cols, 1)
    y = np.random.default_rng(3).uniform(0, 1, (1, geo.nu, geo.nview)).astype(np.float32)
    ref = (A.T @ y.ravel().astype(np.float64)).reshape(1, geo.nx, geo.ny)
    plan = FanPlan(geo)
    out = plan.backward(torch.from_numpy(y).cuda(), "dd_adjoint").cpu().numpy()
    assert rel_l2(out, ref) <= TOL


# ------------------------------------------------------ edge cases -----------

@pytest.mark.parametrize("case", ["off_a", "odd_sizes", "batch_channels", "multislice"])
def test_fan_edge_cases(case):
    if case == "off_a":
        geo = small_fan(nx=64, nu=131, nview=72, off_a=1.25)
        B, C = 1, 1
    elif case == "odd_sizes":
        geo = FanGeometry(nx=63, ny=50, dx=0.9, dy=1.1, nu=97, nview=45, da=1.1 / 1085.6, dsd=1085.6, dso=595.0)
        B, C = 1, 1
    elif case == "batch_channels":
        geo = small_fan(nx=48, nu=96, nview=36)
        B, C = 3, 2
    else:
        geo = FanGeometry(nx=40, ny=40, dx=1.0, dy=1.0, nu=80, nview=30, da=1.2 / 1085.6, dsd=1085.6, dso=595.0,
                          nz=4, dz=1.5, nv=3, dv=2.0, off_v=0.25)
        B, C = 2, 1
    deg = fan_angles(geo.nview)
    rng = np.random.default_rng(11)
    img = rng.uniform(0, 1, (B, geo.nx, geo.ny, geo.nz, C)).astype(np.float32)
    sino = rng.uniform(0, 1, (B, geo.nu, geo.nview, geo.nv, C)).astype(np.float32)
    r = recon_for(geo)
    g = geo.oracle_geo()
    shp = (geo.nx, geo.ny, geo.nz)
    assert rel_l2(r.cDDFanProjection4d(img), orc.dd_fan_fp(img, deg, g, geo.nu, geo.nv)) <= TOL
    assert rel_l2(r.cSiddonFanProjection4d(img), orc.siddon_fan_fp(img, deg, g, geo.nu, geo.nv)) <= TOL
    for t in (0, 1):
        assert rel_l2(r.cDDFanBackprojection4d(sino, type_projector=t),
                      orc.dd_fan_bp(sino, deg, g, shp, type_projector=t)) <= TOL
    assert rel_l2(r.cSiddonFanBackprojection4d(sino), orc.siddon_fan_bp(sino, deg, g, shp)) <= TOL
    assert rel_l2(r.cPixelBackprojection4d(sino), orc.pixel_fan_bp(sino, deg, g, shp)) <= TOL
    assert rel_l2(r.cFilter4d(sino, type_filter=1), orc.ramp_filter(sino, g, filter_type=1)) <= TOL


def test_invalid_arguments_report_errors():
    from reconnet_amd import last_error

    geo = small_fan(nx=16, nu=32, nview=8)
    r = recon_for(geo)
    r.nx = 0
    with pytest.raises(RuntimeError, match="cDistanceDrivenFanBackprojection failed"):
        r.cDDFanBackprojection4d(np.zeros((1, 32, 8, 1, 1), np.float32), img_shape=[0, 16, 1])
    assert "invalid argument" in last_error()
    # the library stays usable after an error
    r.nx = 16
    out = r.cDDFanBackprojection4d(np.ones((1, 32, 8, 1, 1), np.float32))
    assert np.isfinite(out).all() and last_error() == ""


# -------------------------------------------------------------- TV / SQS ----

def test_tv_parity():
    img = phantom(3, 128, 96, 1.0, 1.0)
    g, n, tv = recon_for(small_fan()).cTVSQS2d(img[None, :, :, None])
    go, no, tvo = orc.tv_sqs(img)
    assert rel_l2(g[0, :, :, 0], go) <= TOL
    assert rel_l2(n[0, :, :, 0], no) <= TOL
    assert abs(tv - tvo) / abs(tvo) <= TOL


def test_ossqs_parity_small():
    geo = small_fan(nx=64, nu=128, nview=96)
    deg = fan_angles(geo.nview)
    truth = _img5(phantom(5, 64, 64, 1.0, 1.0))
    b = orc.siddon_fan_fp(truth, deg, geo.oracle_geo(), geo.nu).astype(np.float32)
    w = np.random.default_rng(2).uniform(0.5, 1.0, b.shape).astype(np.float32)
    fp1 = orc.siddon_fan_fp(np.ones_like(truth), deg, geo.oracle_geo(), geo.nu)
    D = orc.siddon_fan_bp((fp1 * w).astype(np.float32), deg, geo.oracle_geo(), (64, 64, 1)).astype(np.float32)
    x0 = np.zeros_like(truth)
    r = recon_for(geo)
    out = r.cSiddonFanOSSQS4d(8, x0, b, D, w, order=None, lb=0.0, ub=1.0)
    ref = orc.os_sqs(x0, b, D, w, deg, geo.oracle_geo(), 8, lb=0.0, ub=1.0)
    assert rel_l2(out, ref) <= TOL
    # a short last subset (nview not a multiple of nSubsets)
    out = r.cSiddonFanOSSQS4d(7, x0, b, D, w)
    ref = orc.os_sqs(x0, b, D, w, deg, geo.oracle_geo(), 7)
    assert rel_l2(out, ref) <= TOL


@pytest.mark.parametrize("nz,dz,nv,dv,nC,nB", [(3, 1.5, 3, 1.0, 2, 1), (2, 3.0, 3, 1.0, 1, 2), (3, 1.0, 2, 1.5, 2, 2)])
def test_ossqs_multislice_multichannel(nz, dz, nv, dv, nC, nB):
    # cOSSQS with several image planes, detector rows and channels
    # (ossqs.cu:79-81 passes them to SiddonFan): rows map to planes through
    # the fan rays' heights (none on a plane boundary), one plane may collect
    # several rows or none, and every
    # (batch, channel) pair is an independent reconstruction
    geo = FanGeometry(nx=40, ny=36, dx=1.0, dy=1.1, nu=72, nview=48, da=1.2 / 1085.6, dsd=1085.6, dso=595.0,
                      nz=nz, dz=dz, nv=nv, dv=dv, off_v=0.2)
    deg = fan_angles(geo.nview)
    g = geo.oracle_geo()
    rng = np.random.default_rng(12)
    truth = rng.uniform(0, 0.05, (nB, geo.nx, geo.ny, nz, nC)).astype(np.float32)
    b = orc.siddon_fan_fp(truth, deg, g, geo.nu, nv).astype(np.float32)
    w = rng.uniform(0.5, 1.0, b.shape).astype(np.float32)
    # the SQS normaliser A'(w A 1) of every (batch, channel) pair
    fp1 = orc.siddon_fan_fp(np.ones_like(truth), deg, g, geo.nu, nv)
    D = orc.siddon_fan_bp((fp1 * w).astype(np.float32), deg, g, (geo.nx, geo.ny, nz)).astype(np.float32)
    x0 = rng.uniform(0, 0.02, truth.shape).astype(np.float32)
    out = recon_for(geo).cSiddonFanOSSQS4d(6, x0, b, D, w, lb=0.0, ub=1.0)
    ref = orc.os_sqs(x0, b, D, w, deg, g, 6, lb=0.0, ub=1.0)
    assert np.abs(ref - x0).max() > 1e-4   # the pass moved the image
    for bb in range(nB):
        for c in range(nC):
            assert rel_l2(out[bb, ..., c], ref[bb, ..., c]) <= TOL, (bb, c)
    # a short last subset and no crop (lb = ub)
    out = recon_for(geo).cSiddonFanOSSQS4d(5, x0, b, D, w, lb=0.0, ub=0.0)
    ref = orc.os_sqs(x0, b, D, w, deg, g, 5, lb=0.0, ub=0.0)
    assert rel_l2(out, ref) <= TOL


def _rect_fan(nx, ny, dx, dy, nu, nview):
    return FanGeometry(nx=nx, ny=ny, dx=dx, dy=dy, nu=nu, nview=nview, da=1.2 / 1085.6, dsd=1085.6, dso=595.0)


@pytest.mark.parametrize("nx,ny,dx,dy", [(64, 64, 1.0, 1.0), (72, 50, 1.1, 0.9)])
def test_ossqs_parity_batch_of_slices(nx, ny, dx, dy):
    # six slices under one engine: the interleaved 4-slice layout with a
    # padded second group, against the oracle's pass of every slice; also on
    # a rectangular, anisotropic grid (the two image orientations differ)
    geo = _rect_fan(nx, ny, dx, dy, 128, 96)
    deg = fan_angles(geo.nview)
    truth = np.concatenate([_img5(phantom(20 + s, nx, ny, dx, dy)) for s in range(6)])
    b = orc.siddon_fan_fp(truth, deg, geo.oracle_geo(), geo.nu).astype(np.float32)
    w = np.random.default_rng(3).uniform(0.5, 1.0, b.shape).astype(np.float32)
    fp1 = orc.siddon_fan_fp(np.ones_like(truth[:1]), deg, geo.oracle_geo(), geo.nu)
    D = orc.siddon_fan_bp((fp1 * w).astype(np.float32), deg, geo.oracle_geo(), (nx, ny, 1)).astype(np.float32)
    x0 = np.random.default_rng(4).uniform(0, 0.02, truth.shape).astype(np.float32)
    out = recon_for(geo).cSiddonFanOSSQS4d(8, x0, b, D, w, lb=0.0, ub=1.0)
    ref = orc.os_sqs(x0, b, D, w, deg, geo.oracle_geo(), 8, lb=0.0, ub=1.0)
    for s in range(6):
        assert rel_l2(out[s], ref[s]) <= TOL, s


@pytest.mark.parametrize("ns,rect", [(1, False), (5, False), (5, True)])
def test_sqs_tv_pass_matches_oracle_composition(ns, rect):
    # one SQS-TV pass of the engine (no reference counterpart) against its
    # definition composed from oracle pieces: per subset s,
    #   x <- clamp(x - (M A_s'W_s(A_s x - b_s) + 2 beta g_TV(x)) / (D + 2 beta n_TV(x) + 1e-6), 0)
    # with the reference TVSQSKernel's g_TV, n_TV (TotalVariation.cu:9-68)
    import torch
    from reconnet_amd.sqs import OsSqsTV

    geo = _rect_fan(56, 37, 1.2, 0.8, 96, 40) if rect else small_fan(nx=48, nu=96, nview=40)
    nsub, beta = 4, 0.2
    deg = fan_angles(geo.nview)
    rng = np.random.default_rng(5)
    g5 = geo.oracle_geo()
    # measurements of piecewise-constant slices with noise, the normaliser
    # A'(w A 1), and a start image that the pass moves by a visible amount
    truth = np.stack([phantom(40 + s, geo.nx, geo.ny, geo.dx, geo.dy) for s in range(ns)])
    b = orc.siddon_fan_fp(truth[..., None, None], deg, g5, geo.nu)[:, :, :, 0, 0]
    b = (b * rng.uniform(0.97, 1.03, b.shape)).astype(np.float32)
    w = rng.uniform(0.5, 1, (ns, geo.nu, geo.nview)).astype(np.float32)
    x = (truth + rng.uniform(0, 0.01, truth.shape)).astype(np.float32)
    fp1 = orc.siddon_fan_fp(np.ones((1, geo.nx, geo.ny, 1, 1), np.float32), deg, g5, geo.nu)[:, :, :, 0, 0]
    D = orc.siddon_fan_bp((fp1 * w)[..., None, None].astype(np.float32), deg, g5,
                          (geo.nx, geo.ny, 1))[:, :, :, 0, 0].astype(np.float32)
    eng = OsSqsTV(geo, nsub, ns, order=None)
    eng.set_data(torch.from_numpy(b).cuda(), torch.from_numpy(w).cuda())
    eng.set_norm(torch.from_numpy(D).cuda())
    eng.set_image(torch.from_numpy(x).cuda())
    eng.run_pass(beta=beta, eps=1e-6, lb=0.0)
    out = eng.image().cpu().numpy()
    eng.close()
    npers = geo.nview // nsub
    ref = x.astype(np.float64)
    for k in range(nsub):
        v = slice(k * npers, (k + 1) * npers)
        xi = ref.astype(np.float32)[:, :, :, None, None]
        fp = orc.siddon_fan_fp(xi, deg[v], g5, geo.nu)[:, :, :, 0, 0]
        r = (w[:, :, v] * (fp - b[:, :, v]))[:, :, :, None, None].astype(np.float32)
        gd = orc.siddon_fan_bp(r, deg[v], g5, (geo.nx, geo.ny, 1))[:, :, :, 0, 0]
        for s in range(ns):
            gt, nt, _ = orc.tv_sqs(ref[s].astype(np.float32))
            num = nsub * gd[s] + 2 * beta * gt
            den = D[s] + 2 * beta * nt
            ref[s] = np.maximum(ref[s] - num / (den + 1e-6), 0.0)
    for s in range(ns):
        assert rel_l2(out[s], ref[s]) <= TOL, s


def test_sqs_normaliser_fp64_matches_oracle():
    import torch
    from reconnet_amd.sqs import OsSqsTV

    geo = small_fan(nx=64, nu=128, nview=96)
    rng = np.random.default_rng(4)
    w = rng.uniform(0.2, 1.0, (2, geo.nu, geo.nview)).astype(np.float32)
    eng = OsSqsTV(geo, 8, 2, order=None)
    eng.set_data(torch.zeros(2, geo.nu, geo.nview, device="cuda"), torch.from_numpy(w).cuda())
    eng.compute_norm()
    D = eng.norm().cpu().numpy()
    deg = fan_angles(geo.nview)
    fp1 = orc.siddon_fan_fp(np.ones((1, 64, 64, 1, 1), np.float32), deg, geo.oracle_geo(), geo.nu)
    for s in range(2):
        ref = orc.siddon_fan_bp((fp1[0, :, :, 0, 0] * w[s])[None, :, :, None, None].astype(np.float32), deg,
                                geo.oracle_geo(), (64, 64, 1))[0, :, :, 0, 0]
        assert rel_l2(D[s], ref) <= TOL


# ---------------------------------------------------------------- 3D / arb --

def _cone(nview, dso=300.0, dsd=500.0, z=0.0):
    th = fan_angles(nview).astype(np.float64)
    src = np.stack([dso * np.sin(th), -dso * np.cos(th), np.full_like(th, z)], 1)
    ctr = np.stack([-(dsd - dso) * np.sin(th), (dsd - dso) * np.cos(th), np.full_like(th, z)], 1)
    du = np.stack([np.cos(th), np.sin(th), np.zeros_like(th)], 1)
    dv = np.tile([0.0, 0.0, 1.0], (nview, 1))
    return [a.astype(np.float32) for a in (ctr, du, dv, src)]


def _arb_recon(nx, ny, nz, nu, nv, du, dv, rotview):
    from reconnet_amd.recon import ReconNet

    r = ReconNet()
    r.nx, r.ny, r.nz = nx, ny, nz
    r.dx = r.dy = r.dz = 1.0
    r.nu, r.nv, r.du, r.dv, r.rotview = nu, nv, du, dv, rotview
    r.off_u, r.off_v = 0.3, -0.2
    return r


@pytest.mark.parametrize("tilt", ["none", "all", "every_other"])
def test_siddon_cone_parity(tilt):
    # "all" / "every_other": the detector of those views turned by 0.3 rad
    # about its normal, so detector rows no longer run along z (the general
    # view path of the backprojection; mixed with upright views in one call)
    C, U, V, S = _cone(24)
    if tilt != "none":
        sel = slice(None) if tilt == "all" else slice(0, None, 2)
        c, s = np.float32(np.cos(0.3)), np.float32(np.sin(0.3))
        U2, V2 = c * U + s * V, c * V - s * U
        U[sel], V[sel] = U2[sel], V2[sel]
    r = _arb_recon(24, 24, 12, 40, 16, 1.5, 1.5, 24)
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, (1, 24, 24, 12, 1)).astype(np.float32)
    y = rng.uniform(0, 1, (1, 40, 24, 16, 1)).astype(np.float32)
    g = dict(dx=1.0, dy=1.0, dz=1.0, du=1.5, dv=1.5, off_u=0.3, off_v=-0.2)
    assert rel_l2(r.cSiddonConeProjectionAbitrary4d(x, C, U, V, S), orc.siddon_cone_fp(x, C, U, V, S, g, 40, 16)) <= TOL
    assert rel_l2(r.cSiddonConeBackprojectionAbitrary4d(y, C, U, V, S),
                  orc.siddon_cone_bp(y, C, U, V, S, g, (24, 24, 12))) <= TOL


@pytest.mark.parametrize("tilted", [False, True])
def test_siddon_parallel_parity(tilted):
    # invRayDir points from the detector back through the volume towards the
    # source (SiddonParallel.cu:28-30); the tilted case has rays climbing in z
    # and a grid centre off the origin
    nview = 18
    th = fan_angles(nview).astype(np.float64)
    C = np.stack([-200 * np.sin(th), 200 * np.cos(th), np.zeros_like(th)], 1).astype(np.float32)
    U = np.stack([np.cos(th), np.sin(th), np.zeros_like(th)], 1).astype(np.float32)
    V = np.tile([0.0, 0.0, 1.0], (nview, 1)).astype(np.float32)
    D = np.stack([np.sin(th), -np.cos(th), np.full_like(th, 0.02 if tilted else 0.0)], 1)
    D = (D / np.linalg.norm(D, axis=1, keepdims=True)).astype(np.float32)
    r = _arb_recon(20, 20, 6, 36, 8, 1.0, 1.0, nview)
    g = dict(dx=1.0, dy=1.0, dz=1.0, du=1.0, dv=1.0, off_u=0.3, off_v=-0.2)
    if tilted:
        r.cx, r.cy, r.cz = 1.5, -2.0, 0.5
        g.update(cx=1.5, cy=-2.0, cz=0.5)
    rng = np.random.default_rng(6)
    x = rng.uniform(0, 1, (1, 20, 20, 6, 1)).astype(np.float32)
    y = rng.uniform(0, 1, (1, 36, nview, 8, 1)).astype(np.float32)
    ref_fp = orc.siddon_par_fp(x, C, U, V, D, g, 36, 8)
    ref_bp = orc.siddon_par_bp(y, C, U, V, D, g, (20, 20, 6))
    # most rays cross the volume: the comparison is not of two zero arrays
    assert (ref_fp > 0).mean() > 0.2 and (ref_bp > 0).mean() > 0.5
    assert rel_l2(r.cSiddonParallelProjection4d(x, C, U, V, D), ref_fp) <= TOL
    assert rel_l2(r.cSiddonParallelBackprojection4d(y, C, U, V, D), ref_bp) <= TOL


def test_dd_tomo_parity():
    nview = 9
    th = np.linspace(-0.35, 0.35, nview)
    S = np.stack([250 * np.sin(th), 10 * np.cos(th), 300 * np.cos(th)], 1).astype(np.float32)
    C = np.tile([0.0, 0.0, -60.0], (nview, 1)).astype(np.float32)
    r = _arb_recon(32, 28, 10, 48, 40, 1.2, 1.0, nview)
    rng = np.random.default_rng(7)
    x = rng.uniform(0, 1, (1, 32, 28, 10, 2)).astype(np.float32)
    y = rng.uniform(0, 1, (1, 48, nview, 40, 2)).astype(np.float32)
    g = dict(dx=1.0, dy=1.0, dz=1.0, du=1.2, dv=1.0, off_u=0.3, off_v=-0.2)
    assert rel_l2(r.cDDTomoProjection4d(x, C, S), orc.dd_tomo_fp(x, C, S, g, 48, 40)) <= TOL
    assert rel_l2(r.cDDTomoBackprojection4d(y, C, S), orc.dd_tomo_bp(y, C, S, g, (32, 28, 10))) <= TOL


@pytest.mark.parametrize("masked", [False, True])
def test_norm_images_match_oracle(masked):
    # cDDTomoNormImg4d and cSiddonConeNormImgAbitrary4d (ReconNet_no_tf.py:281-288,
    # 551-558): A^T(A(1 * mask)) * mask, against the same composition of the
    # FP64 oracle's projector pair (two channels; the mask a disc that leaves
    # the corner voxels out)
    nview = 9
    th = np.linspace(-0.35, 0.35, nview)
    S = np.stack([250 * np.sin(th), 10 * np.cos(th), 300 * np.cos(th)], 1).astype(np.float32)
    C = np.tile([0.0, 0.0, -60.0], (nview, 1)).astype(np.float32)
    r = _arb_recon(32, 28, 10, 48, 40, 1.2, 1.0, nview)
    g = dict(dx=1.0, dy=1.0, dz=1.0, du=1.2, dv=1.0, off_u=0.3, off_v=-0.2)
    shp = (1, 32, 28, 10, 2)
    yy, xx = np.meshgrid(np.arange(28) - 13.5, np.arange(32) - 15.5)
    disc = ((xx ** 2 + yy ** 2) < 14.0 ** 2).astype(np.float32)[None, :, :, None, None]
    mask = np.broadcast_to(disc, shp).astype(np.float32) if masked else None
    m = mask if masked else np.ones(shp, np.float32)
    prj = np.zeros((1, 48, nview, 40, 2), np.float32)
    out = r.cDDTomoNormImg4d(prj, C, S, mask=mask)
    ref = orc.dd_tomo_bp(orc.dd_tomo_fp(m, C, S, g, 48, 40), C, S, g, (32, 28, 10)) * m
    assert np.abs(ref).max() > 1.0
    assert rel_l2(out, ref) <= TOL

    Cc, U, V, Sc = _cone(24)
    rc = _arb_recon(24, 24, 12, 40, 16, 1.5, 1.5, 24)
    gc = dict(dx=1.0, dy=1.0, dz=1.0, du=1.5, dv=1.5, off_u=0.3, off_v=-0.2)
    shp = (1, 24, 24, 12, 2)
    yy, xx = np.meshgrid(np.arange(24) - 11.5, np.arange(24) - 11.5)
    disc = ((xx ** 2 + yy ** 2) < 10.5 ** 2).astype(np.float32)[None, :, :, None, None]
    mask = np.broadcast_to(disc, shp).astype(np.float32) if masked else None
    m = mask if masked else np.ones(shp, np.float32)
    prj = np.zeros((1, 40, 24, 16, 2), np.float32)
    out = rc.cSiddonConeNormImgAbitrary4d(prj, Cc, U, V, Sc, mask=mask)
    ref = orc.siddon_cone_bp(orc.siddon_cone_fp(m, Cc, U, V, Sc, gc, 40, 16), Cc, U, V, Sc, gc, (24, 24, 12)) * m
    assert np.abs(ref).max() > 1.0
    assert rel_l2(out, ref) <= TOL


def test_cone_ossqs_parity():
    C, U, V, S = _cone(16)
    r = _arb_recon(16, 16, 8, 32, 12, 1.5, 1.5, 16)
    rng = np.random.default_rng(8)
    truth = rng.uniform(0, 1, (1, 16, 16, 8, 1)).astype(np.float32)
    g = dict(dx=1.0, dy=1.0, dz=1.0, du=1.5, dv=1.5, off_u=0.3, off_v=-0.2)
    b = orc.siddon_cone_fp(truth, C, U, V, S, g, 32, 12).astype(np.float32)
    D = orc.siddon_cone_bp(orc.siddon_cone_fp(np.ones_like(truth), C, U, V, S, g, 32, 12).astype(np.float32),
                           C, U, V, S, g, (16, 16, 8)).astype(np.float32)
    mask = np.ones_like(truth)
    mask[:, :2] = 0
    x0 = np.full_like(truth, 0.5)
    out = r.cSiddonConeAbitraryOSSQS4d(4, x0, b, D, C, U, V, S, mask=mask, lb=0, ub=2)
    ref = orc.os_sqs_cone(x0, b, D, mask, C, U, V, S, g, 4, lb=0, ub=2)
    assert rel_l2(out, ref) <= TOL


# (mPI = 0 is left out: a lone view in the row taper has w ~ 1e-7 and adds
# w P / (w + 1e-6), whose value follows the last bits of cos near pi / 2)
@pytest.mark.parametrize("cand", [None, 0, 1, 2])
@pytest.mark.parametrize("m_pi,Q,nC", [(1, 0.5, 1), (2, 0.3, 2), (1, 0.8, 1)])
def test_helical_bp_parity(cand, m_pi, Q, nC):
    from reconnet_amd._lib import autotune_pin
    from reconnet_amd.recon import ReconNet

    npp, turns, nu, nv = 48, 3, 80, 24
    r = ReconNet()
    r.nx, r.ny, r.nz, r.dx, r.dy, r.dz = 32, 30, 13, 1.0, 1.0, 1.0
    r.du, r.dv, r.off_u, r.off_v, r.dsd, r.dso = 1.0, 1.0, 0.25, -0.5, 1000.0, 600.0
    rng = np.random.default_rng(9)
    prj = rng.uniform(0, 1, (1, nu, 2 * npp * turns, nv, nC)).astype(np.float32)
    if cand is not None:
        autotune_pin("helical_bp", cand)
    try:
        out = r.cHelicalBP4d(prj, npp, 0.1, 8.0, 12.0, mPI=m_pi, Q=Q)
    finally:
        autotune_pin("helical_bp", None)
    g = dict(dx=1.0, dy=1.0, dz=1.0, du=1.0, dv=1.0, off_u=0.25, off_v=-0.5, dsd=1000.0, dso=600.0)
    ref = orc.helical_bp(prj, g, (32, 30, 13), npp, 0.1, 8.0, 12.0, m_pi=m_pi, Q=Q)
    assert np.abs(ref).max() > 0.1
    assert rel_l2(out, ref) <= TOL


def test_helical_bp_parity_wide_fan():
    # SID 40 mm around a 32 x 30 mm volume: |u| / SID reaches 0.55, so the
    # voxels take both branches of the source-height term (the series for
    # |u| / SID <= 0.35, asinf beyond)
    from reconnet_amd.recon import ReconNet

    npp, turns, nu, nv = 48, 3, 96, 24
    r = ReconNet()
    r.nx, r.ny, r.nz, r.dx, r.dy, r.dz = 32, 30, 13, 1.0, 1.0, 1.0
    r.du, r.dv, r.off_u, r.off_v, r.dsd, r.dso = 0.5, 1.0, 0.25, -0.5, 80.0, 40.0
    rng = np.random.default_rng(10)
    prj = rng.uniform(0, 1, (1, nu, 2 * npp * turns, nv, 1)).astype(np.float32)
    out = r.cHelicalBP4d(prj, npp, 0.1, 8.0, 12.0, mPI=1, Q=0.5)
    g = dict(dx=1.0, dy=1.0, dz=1.0, du=0.5, dv=1.0, off_u=0.25, off_v=-0.5, dsd=80.0, dso=40.0)
    ref = orc.helical_bp(prj, g, (32, 30, 13), npp, 0.1, 8.0, 12.0, m_pi=1, Q=0.5)
    assert np.abs(ref).max() > 0.1
    assert rel_l2(out, ref) <= TOL


@pytest.mark.parametrize("cand", [0, 1, 2])
def test_helical_bp_truncated_detector_matches_oracle(cand):
    # 24 columns of 1 mm against a 32 x 30 mm volume: samples fall off both
    # ends of the detector, through the columns -1 / 0 and nu - 1 / nu whose
    # interpolation pair is clamped into the detector
    from reconnet_amd._lib import autotune_pin
    from reconnet_amd.recon import ReconNet

    npp, turns, nu, nv = 48, 3, 24, 24
    r = ReconNet()
    r.nx, r.ny, r.nz, r.dx, r.dy, r.dz = 32, 30, 13, 1.0, 1.0, 1.0
    r.du, r.dv, r.off_u, r.off_v, r.dsd, r.dso = 1.0, 1.0, 0.3, -0.5, 1000.0, 600.0
    prj = np.random.default_rng(23).uniform(0, 1, (1, nu, 2 * npp * turns, nv, 1)).astype(np.float32)
    autotune_pin("helical_bp", cand)
    try:
        out = r.cHelicalBP4d(prj, npp, 0.1, 8.0, 12.0, mPI=1, Q=0.5)
    finally:
        autotune_pin("helical_bp", None)
    g = dict(dx=1.0, dy=1.0, dz=1.0, du=1.0, dv=1.0, off_u=0.3, off_v=-0.5, dsd=1000.0, dso=600.0)
    ref = orc.helical_bp(prj, g, (32, 30, 13), npp, 0.1, 8.0, 12.0, m_pi=1, Q=0.5)
    assert np.abs(ref).max() > 0.1
    # voxels beyond the detector's 12 mm half-width miss it in some views
    xx, yy = np.meshgrid(np.arange(32) - 15.5, np.arange(30) - 14.5, indexing="ij")
    rad = np.hypot(xx, yy)
    assert ref[0, rad > 15, :, 0].mean() < 0.8 * ref[0, rad < 8, :, 0].mean()
    assert rel_l2(out, ref) <= TOL


def test_helical_bp_device_entry_equals_host_abi():
    import torch

    from reconnet_amd.helical import helical_backproject
    from reconnet_amd.recon import ReconNet

    npp, nu, nv = 48, 80, 24
    r = ReconNet()
    r.nx, r.ny, r.nz, r.dx, r.dy, r.dz = 32, 30, 13, 1.0, 1.0, 1.0
    r.du, r.dv, r.off_u, r.off_v, r.dsd, r.dso = 1.0, 1.0, 0.25, -0.5, 1000.0, 600.0
    prj = np.random.default_rng(19).uniform(0, 1, (2, nu, 2 * npp * 3, nv, 1)).astype(np.float32)
    host = r.cHelicalBP4d(prj, npp, 0.1, 8.0, 12.0)
    dev = helical_backproject(torch.from_numpy(prj).cuda(), npp, 0.1, 8.0, 12.0, nx=32, ny=30, nz=13, dx=1.0, dy=1.0,
                              dz=1.0, du=1.0, dv=1.0, dsd=1000.0, dso=600.0, off_u=0.25, off_v=-0.5)
    assert np.abs(host).max() > 0.1
    assert np.array_equal(dev.cpu().numpy(), host)


def test_helical_bp_rejects_a_single_detector_column():
    # the gather reads each row's two interpolation columns as one pair
    import torch

    from reconnet_amd.helical import helical_backproject

    par = torch.zeros((1, 1, 96, 8, 1), device="cuda")
    with pytest.raises(RuntimeError, match="at least 2 detector columns"):
        helical_backproject(par, 48, 0.1, 8.0, 12.0, nx=8, ny=8, nz=4, dx=1.0, dy=1.0, dz=1.0, du=1.0, dv=1.0,
                            dsd=1000.0, dso=600.0)


def test_helical_bp_full_size_slices_match_oracle():
    # 256^2 x 128 volume at 1 mm, 64 rows x 512 columns (1.0 x 1.2 mm), SID 595 mm,
    # SDD 1085.6 mm, 720 views per turn, 3 turns, pitch 1 (table feed per turn =
    # 64 * 1.0 * 595 / 1085.6 mm).  The GPU reconstructs the whole volume; the
    # FP64 oracle recomputes two of its slices (z = volZ0 + iz dz is exact here,
    # so the sub-volume with volZ0 + iz0 dz has identical voxel positions).
    from reconnet_amd.recon import ReconNet

    npp, turns, nu, nv = 360, 3, 512, 64
    zrot = np.float32(64 * 1.0 * 595.0 / 1085.6)
    r = ReconNet()
    r.nx, r.ny, r.nz, r.dx, r.dy, r.dz = 256, 256, 128, 1.0, 1.0, 1.0
    r.du, r.dv, r.off_u, r.off_v, r.dsd, r.dso = 1.2, 1.0, 0.0, 0.0, 1085.6, 595.0
    rng = np.random.default_rng(1234)
    prj = rng.uniform(0, 1, (1, nu, 2 * npp * turns, nv, 1)).astype(np.float32)
    vol_z0 = -11.0
    out = r.cHelicalBP4d(prj, npp, 0.0, vol_z0, float(zrot))
    assert np.isfinite(out).all()
    g = dict(dx=1.0, dy=1.0, dz=1.0, du=1.2, dv=1.0, off_u=0.0, off_v=0.0, dsd=1085.6, dso=595.0)
    for iz0 in (30, 77):
        ref = orc.helical_bp(prj, g, (256, 256, 2), npp, 0.0, vol_z0 + iz0, float(zrot))
        assert rel_l2(out[:, :, :, iz0:iz0 + 2], ref) <= TOL
