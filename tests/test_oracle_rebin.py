"""Pin of the rebinning restatement (oracle/rebin.py) by the geometry itself.

Flat-detector helical views of a smooth volume are simulated with the FP64
Siddon cone oracle (SiddonCone.cu:11-68) and rebinned; the same parallel-cone
rays are also simulated directly, one detector column per (t, u): source
u (sin t, -cos t) + sqrt(dso^2 - u^2) (cos t, sin t) at height
zrot (t - asin(u/dso)) / (2 pi), row v at in-plane distance dsd along
-(cos t, sin t) (the rays bpHelicalParallelConebeam assumes,
Recon3D/HelicalBP.cu:48-53).  The two agree to the interpolation error;
a rebinning with the fan angle's sign flipped, without the row scaling
1 / cos(g), or with the detector columns mirrored does not.
"""

import functools

import numpy as np
import pytest

import oracle as orc
from oracle.rebin import rebin_fan_to_parallel
from reconnet_amd.helical import helical_scan_vectors

NX, NZ, D = 128, 80, 1.0
# a wide fan (|g| up to 27 deg) makes the 1 / cos(g) row scaling matter
DSO, DSD = 150.0, 260.0
NU_F, NV_F, DU_F, DV_F = 96, 12, 3.0, 4.0
NU_P = 80
NVIEW_TURN, TURNS = 240, 2
ZROT = 0.8 * NV_F * DV_F * DSO / DSD          # pitch 0.8
# no view exactly along a voxel axis: rays along voxel rows see the voxel
# staircase differently from the slightly rotated rays they are interpolated from
TH0 = 0.05


@functools.lru_cache(maxsize=1)
def _volume():
    r = (np.arange(NX) - (NX - 1) / 2) * D
    rz = (np.arange(NZ) - (NZ - 1) / 2) * D
    X, Y, Z = np.meshgrid(r, r, rz, indexing="ij")
    # smooth and negligible (< 1e-4) at the volume's faces: no truncation edges to interpolate
    v = np.exp(-((X - 10) ** 2 + (Y + 6) ** 2) / (2 * 12.0 ** 2) - Z ** 2 / (2 * 7.0 ** 2))
    v += 0.6 * np.exp(-((X + 16) ** 2 + (Y - 12) ** 2) / (2 * 9.0 ** 2) - (Z - 6) ** 2 / (2 * 6.0 ** 2))
    # off-centre, so the outer parallel columns (largest fan angles) carry signal
    v += 0.8 * np.exp(-((X - 36) ** 2 + (Y - 30) ** 2) / (2 * 6.0 ** 2) - (Z + 4) ** 2 / (2 * 4.0 ** 2))
    return v.astype(np.float32)[None, ..., None]


def _geo():
    return dict(dx=D, dy=D, dz=D, cz=0.5 * ZROT * TURNS, du=DU_F, dv=DV_F)


@functools.lru_cache(maxsize=1)
def _scan():
    vol = _volume()
    raw = _raw(vol)
    par, a = _rebin(raw)
    ref = _direct_parallel(vol, a["nu_p"], a["du_p"], a["nview_p"], a["dtheta"], a["nv_p"], a["dv_p"])
    return raw, par, a, ref


def _raw(vol):
    nview = NVIEW_TURN * TURNS
    C, U, V, S = helical_scan_vectors(nview, DSO, DSD, ZROT, beta0=TH0, dbeta=2 * np.pi / NVIEW_TURN, theta0=TH0)
    return orc.siddon_cone_fp(vol, C, U, V, S, _geo(), NU_F, NV_F)


def _direct_parallel(vol, nu_p, du_p, nview_p, dtheta, nv_p, dv_p):
    t = TH0 + np.arange(nview_p) * dtheta
    u = (np.arange(nu_p) - (nu_p - 1) / 2) * du_p
    T, Uu = np.meshgrid(t, u, indexing="ij")          # columns ordered [view][u]
    T, Uu = T.ravel(), Uu.ravel()
    d = np.stack([np.cos(T), np.sin(T)], 1)
    n = np.stack([np.sin(T), -np.cos(T)], 1)
    S2 = Uu[:, None] * n + np.sqrt(DSO ** 2 - Uu ** 2)[:, None] * d
    zs = ZROT * (T - np.arcsin(Uu / DSO) - TH0) / (2 * np.pi)
    src = np.concatenate([S2, zs[:, None]], 1)
    ctr = np.concatenate([S2 - DSD * d, zs[:, None]], 1)
    detu = np.concatenate([n, np.zeros((len(T), 1))], 1)
    detv = np.tile([0.0, 0.0, 1.0], (len(T), 1))
    g = dict(_geo(), du=1.0, dv=dv_p)
    cols = orc.siddon_cone_fp(vol, *(a.astype(np.float32) for a in (ctr, detu, detv, src)), g, 1, nv_p)
    return cols[0, 0, :, :, 0].reshape(nview_p, nu_p, nv_p).transpose(1, 0, 2)    # [u][view][v]


def _rebin(raw, **kw):
    args = dict(du_f=DU_F, dv_f=DV_F, beta0=TH0, dbeta=2 * np.pi / NVIEW_TURN, nu_p=NU_P, nview_p=NVIEW_TURN * TURNS,
                nv_p=NV_F, du_p=DU_F * DSO / DSD, dv_p=DV_F, theta0=TH0, dtheta=2 * np.pi / NVIEW_TURN, dsd=DSD,
                dso=DSO)
    args.update(kw)
    return rebin_fan_to_parallel(raw, **args)[0, :, :, :, 0], args


def test_rebinned_views_match_directly_simulated_parallel_rays():
    raw, par, a, ref = _scan()
    # views whose fan angles all fall inside the scan, rows inside the detector
    g = np.arcsin(((np.arange(a["nu_p"]) - (a["nu_p"] - 1) / 2) * a["du_p"]) / DSO).max()
    k0 = int(np.ceil(g / a["dtheta"])) + 1
    sel = (slice(None), slice(k0, a["nview_p"] - k0), slice(1, a["nv_p"] - 1))
    err = np.abs(par[sel] - ref[sel]).max() / np.abs(ref[sel]).max()
    assert ref[sel].max() > 1.0
    assert err < 0.05, err
    mean = np.abs(par[sel] - ref[sel]).mean() / np.abs(ref[sel]).mean()
    assert mean < 0.01, mean


@pytest.mark.parametrize("mistake", ["gamma_sign", "no_row_scale", "mirror"])
def test_rebinning_mistakes_are_visible(mistake):
    raw, par, a, ref = _scan()
    if mistake == "gamma_sign":
        bad, _ = _rebin(raw, _gamma_sign=-1.0)          # b = t + g
    elif mistake == "no_row_scale":
        bad, _ = _rebin(raw, _row_scale=False)          # rows read at v instead of v / cos(g)
    else:
        bad, _ = _rebin(raw[:, ::-1])                   # detector columns mirrored
    g = np.arcsin(((np.arange(a["nu_p"]) - (a["nu_p"] - 1) / 2) * a["du_p"]) / DSO).max()
    k0 = int(np.ceil(g / a["dtheta"])) + 1
    sel = (slice(None), slice(k0, a["nview_p"] - k0), slice(1, a["nv_p"] - 1))
    good = np.abs(par[sel] - ref[sel]).mean()
    worse = np.abs(bad[sel] - ref[sel]).mean()
    assert worse > 1.8 * good, (mistake, good, worse)
