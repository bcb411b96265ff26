"""Shared helpers for the test suite (sizes, geometry records, error norms)."""

import numpy as np

from reconnet_amd.geometry import FanGeometry, fan_angles


def rel_l2(a, b):
    """Relative L2 error of a against the reference b.

    An all-zero (or non-finite) reference makes the ratio meaningless, so it
    raises instead of reporting 0 vs 0 as a pass."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.shape != b.shape:
        raise ValueError("rel_l2: shapes differ: %s vs %s" % (a.shape, b.shape))
    nb = np.linalg.norm(b)
    if not np.isfinite(nb) or nb == 0.0:
        raise ValueError("rel_l2: reference norm is %r; a parity check against it proves nothing" % nb)
    return float(np.linalg.norm(a - b) / nb)


def small_fan(nx=64, nu=128, nview=90, dx=1.0, du=1.2, off_a=0.0):
    return FanGeometry(nx=nx, ny=nx, dx=dx, dy=dx, nu=nu, nview=nview, da=du / 1085.6, dsd=1085.6, dso=595.0,
                       off_a=off_a)


def recon_for(geo):
    """reconnet_amd.recon.ReconNet configured for a FanGeometry."""
    from reconnet_amd.recon import ReconNet

    r = ReconNet()
    r.nx, r.ny, r.nz = geo.nx, geo.ny, geo.nz
    r.dx, r.dy, r.dz = geo.dx, geo.dy, geo.dz
    r.nu, r.nv, r.rotview = geo.nu, geo.nv, geo.nview
    r.da, r.dv, r.off_a, r.off_v = geo.da, geo.dv, geo.off_a, geo.off_v
    r.dsd, r.dso = geo.dsd, geo.dso
    r.UpdateAngularSampling()
    return r


def disk_image(geo, r0, cx=0.0, cy=0.0, mu=1.0, ss=8):
    sub = (np.arange(geo.nx * ss) + 0.5) / ss - geo.nx / 2
    X, Y = np.meshgrid(sub * geo.dx, sub * geo.dy, indexing="ij")
    m = ((X - cx) ** 2 + (Y - cy) ** 2 <= r0 ** 2).astype(np.float64)
    return (mu * m.reshape(geo.nx, ss, geo.ny, ss).mean(axis=(1, 3))).astype(np.float32)


def ray_geometry(geo, view_angles, iu):
    """Source point and unit direction of the ray through bin centre iu."""
    th = float(view_angles)
    a = (-(iu - (geo.nu - 1) / 2) - geo.off_a) * float(np.float32(geo.da))
    src = np.array([geo.dso * np.sin(th), -geo.dso * np.cos(th)])
    d = np.array([np.sin(a - th), np.cos(a - th)])
    return src, d


__all__ = ["rel_l2", "small_fan", "recon_for", "disk_image", "ray_geometry", "fan_angles"]
