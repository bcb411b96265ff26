"""Helical cone-beam reconstruction on the GPU: flat-detector helical views ->
fan-to-parallel rebinning (rn_helical_rebin) -> parallel ramp filter
(cFilterParallelFilter, Filter.cu:476) -> weighted helical backprojection
(cHelicalBPFromParallelConebeamBackprojection, Recon3D/HelicalBP.cu:132).

The rebinning kernel is compared with its FP64 restatement (oracle/rebin.py,
itself pinned by directly simulated parallel-cone rays on the CPU); the whole
chain reconstructs the seeded 3D ellipsoid phantom of the 256^2 x 128 helical
workload (1 mm voxels, 64 x 512 flat detector of 1.0 x 1.2 mm cells, SID
595 mm, SDD 1085.6 mm, 720 views per turn, 3 turns at pitch 1), and input
mistakes (mirrored columns, rows upside down, a view offset) visibly do not.
"""

import numpy as np
import pytest

from oracle.rebin import rebin_fan_to_parallel as rebin_oracle
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("case", ["kscale1", "kscale_half_offsets"])
def test_rebin_matches_oracle(case):
    import torch

    from reconnet_amd.helical import rebin_fan_to_parallel

    rng = np.random.default_rng(21)
    B, C, nu_f, nview_f, nv_f = 2, 2, 40, 90, 10
    fan = rng.uniform(0, 1, (B, nu_f, nview_f, nv_f, C)).astype(np.float32)
    kw = dict(du_f=2.0, dv_f=1.5, beta0=0.3, dbeta=2 * np.pi / 60, nu_p=36, nview_p=70, nv_p=12, du_p=1.1,
              dv_p=1.4, theta0=0.6, dtheta=np.pi / 30, dsd=150.0, dso=90.0)
    if case == "kscale_half_offsets":
        kw.update(dtheta=np.pi / 60, nview_p=150, off_u_f=0.3, off_v_f=-0.4, off_u_p=-0.2, off_v_p=0.35)
    out = rebin_fan_to_parallel(torch.from_numpy(fan).cuda(), **kw).cpu().numpy()
    ref = rebin_oracle(fan, **kw)
    assert np.abs(ref).max() > 0.5 and (ref == 0).mean() < 0.5
    assert rel_l2(out, ref) <= TOL


NPP, TURNS = 360, 3
NU_F, NV_F, DU_F, DV_F = 512, 64, 1.2, 1.0
DSO, DSD = 595.0, 1085.6
ZROT = 64 * DV_F * DSO / DSD            # pitch 1: the detector height at the axis per turn
VOL_Z0 = -11.0                         # volume z from -11 to 116 mm; the source rises 0 .. 3 ZROT
R_TRUTH = 150.0                        # inside every parallel view's column range (|u| < 161 mm)


@pytest.fixture(scope="module")
def helical():
    import torch

    from reconnet_amd.helical import helical_scan_vectors, rebin_fan_to_parallel
    from reconnet_amd.phantom import phantom3d
    from reconnet_amd.recon import ReconNet
    from reconnet_amd.torch_ops import ConePlan

    nx, nz = 256, 128
    truth = phantom3d(0, nx, nx, nz, 1.0, 1.0, 1.0, device="cuda")
    r = (torch.arange(nx, device="cuda", dtype=torch.float32) - (nx - 1) / 2)
    truth = truth * ((r[:, None] ** 2 + r[None, :] ** 2) <= R_TRUTH ** 2)[:, :, None]
    nview_f = 2 * NPP * TURNS
    vecs = helical_scan_vectors(nview_f, DSO, DSD, ZROT, beta0=0.0, dbeta=np.pi / NPP, theta0=0.0)
    plan = ConePlan(nx, nx, nz, 1.0, 1.0, 1.0, NU_F, NV_F, DU_F, DV_F, *vecs, cz=VOL_Z0 + (nz - 1) / 2)
    raw = plan.forward(truth[None, ..., None], "siddon")
    torch.cuda.synchronize()
    plan.close()
    du_p = DU_F * DSO / DSD
    rb = dict(du_f=DU_F, dv_f=DV_F, beta0=0.0, dbeta=np.pi / NPP, nu_p=512, nview_p=nview_f, nv_p=NV_F, du_p=du_p,
              dv_p=DV_F, theta0=0.0, dtheta=np.pi / NPP, dsd=DSD, dso=DSO)

    rec = ReconNet()
    rec.nx, rec.ny, rec.nz, rec.dx, rec.dy, rec.dz = nx, nx, nz, 1.0, 1.0, 1.0
    rec.du, rec.dv, rec.off_u, rec.off_v, rec.dsd, rec.dso = du_p, DV_F, 0.0, 0.0, DSD, DSO

    def reconstruct(raw_views, theta_shift=0):
        par = rebin_fan_to_parallel(raw_views.contiguous(), **rb).cpu().numpy()
        filt = rec.cFilterParallel4d(par, type_filter=2)
        return rec.cHelicalBP4d(filt, NPP, theta_shift * np.pi / NPP, VOL_Z0, ZROT)[0, ..., 0]

    yield truth.cpu().numpy(), raw, reconstruct


def _fit_err(x, truth):
    # slices whose pi-windows and both neighbouring pi-segments lie inside the
    # scan (z from 25 to 72 mm), inside the reconstructed cylinder
    n = truth.shape[0]
    r = np.arange(n) - (n - 1) / 2
    m = (r[:, None] ** 2 + r[None, :] ** 2) <= 140.0 ** 2
    zs = slice(int(25 - VOL_Z0), int(72 - VOL_Z0))
    a = x[:, :, zs][m].astype(np.float64).ravel()
    b = truth[:, :, zs][m].astype(np.float64).ravel()
    s = float(a @ b / (a @ a))
    return rel_l2(s * a, b), s


def test_helical_wfbp_reconstructs_the_phantom(helical):
    truth, raw, reconstruct = helical
    x = reconstruct(raw)
    assert np.isfinite(x).all()
    err, scale = _fit_err(x, truth)
    print(f"helical WFBP: fit error {err:.4f}, scale {scale:.4f}", flush=True)
    # the parallel filter normalises by pi / (all views); the backprojection
    # sums one pi of views: the reconstruction is the phantom / (views / NPP)
    assert abs(scale * NPP / raw.shape[2] - 1.0) < 0.03, scale
    assert err < 0.09, err   # measured 0.062


@pytest.mark.parametrize("mistake", ["mirror_columns", "rows_flipped", "theta_shift"])
def test_helical_wfbp_input_mistakes_are_visible(helical, mistake):
    truth, raw, reconstruct = helical
    good, _ = _fit_err(reconstruct(raw), truth)
    if mistake == "mirror_columns":
        bad = reconstruct(raw.flip(1))
    elif mistake == "rows_flipped":
        bad = reconstruct(raw.flip(3))
    else:
        bad = reconstruct(raw, theta_shift=8)
    worse, _ = _fit_err(bad, truth)
    print(f"{mistake}: {good:.4f} -> {worse:.4f}", flush=True)
    assert worse > 1.5 * good, (mistake, good, worse)
