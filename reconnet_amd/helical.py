"""Helical cone-beam scans: fan -> parallel-cone rebinning for the helical
weighted backprojection.

The reference's helical backprojection (Recon3D/HelicalBP.cu:10-83,
``cHelicalBPFromParallelConebeamBackprojection``) takes parallel-cone data:
view angle t_k = theta0 + k pi / nviewPerPI, rays of direction -(cos t, sin t)
at signed distance u = x sin t - y cos t from the axis, rows v on the cylinder
of radius dsd around the source, whose height is z_s = zrot (t - asin(u/dso)
- theta0) / (2 pi).  A scanner measures flat-detector cone-beam views;
``rebin_fan_to_parallel`` resamples them (``rn_helical_rebin``).  The raw
views follow ``helical_scan_vectors``: source angle b_m = beta0 + m dbeta at
(dso cos b, dso sin b, zrot (b - theta0) / (2 pi)), the argument layout of
cSiddonConeProjectionAbitrary (SiddonCone.cu:117), so the library's Siddon
cone projector simulates them.
"""

import ctypes

import numpy as np

from ._lib import check, lib

__all__ = ["helical_scan_vectors", "rebin_fan_to_parallel", "helical_backproject"]


def helical_scan_vectors(nview, dso, dsd, zrot, beta0=0.0, dbeta=None, theta0=0.0):
    """(detCenter, detU, detV, src), float32 [nview, 3], of a helical scan with
    a flat detector at distance dsd: view m has source angle
    b = beta0 + m dbeta (default step 2 pi / nview), source
    (dso cos b, dso sin b, zrot (b - theta0) / (2 pi)), detector columns along
    (sin b, -cos b, 0) and rows along z."""
    if dbeta is None:
        dbeta = 2 * np.pi / nview
    b = beta0 + np.arange(nview, dtype=np.float64) * dbeta
    z = zrot * (b - theta0) / (2 * np.pi)
    cb, sb = np.cos(b), np.sin(b)
    src = np.stack([dso * cb, dso * sb, z], 1)
    ctr = np.stack([(dso - dsd) * cb, (dso - dsd) * sb, z], 1)
    du = np.stack([sb, -cb, np.zeros_like(b)], 1)
    dv = np.tile([0.0, 0.0, 1.0], (nview, 1))
    return tuple(a.astype(np.float32) for a in (ctr, du, dv, src))


def _sigs():
    h = lib()
    if not getattr(h, "_rebin_sigs", False):
        ci, cf, P = ctypes.c_int, ctypes.c_float, ctypes.c_void_p
        h.rn_helical_rebin.restype = ci
        h.rn_helical_rebin.argtypes = [P, P, ci, ci, ci, ci, ci, cf, cf, cf, cf, cf, cf, ci, ci, ci, cf, cf, cf, cf,
                                       cf, cf, cf, cf, P]
        h.rn_helical_bp.restype = ci
        h.rn_helical_bp.argtypes = [P, P, ci, ci, ci, ci, ci, cf, cf, cf, ci, ci, ci, cf, cf, cf, cf, cf, cf, ci, cf,
                                    cf, cf, ci, cf, P]
        h._rebin_sigs = True
    return h


def rebin_fan_to_parallel(fan, *, du_f, dv_f, beta0, dbeta, nu_p, nview_p, nv_p, du_p, dv_p, theta0, dtheta,
                          dsd, dso, off_u_f=0.0, off_v_f=0.0, off_u_p=0.0, off_v_p=0.0, out=None):
    """fan: float32 CUDA tensor [B, nu_f, nview_f, nv_f, C] of flat-detector
    views (see helical_scan_vectors); returns the parallel-cone data
    [B, nu_p, nview_p, nv_p, C] (view theta0 + k dtheta, u and v as the
    helical backprojection reads them).  Trilinear interpolation in
    (column, view, row), zero outside the raw data."""
    import torch

    from .torch_ops import _ptr, _stream

    if fan.dim() != 5 or fan.dtype != torch.float32 or not fan.is_cuda:
        raise ValueError("fan must be a float32 CUDA tensor [B, nu, nview, nv, C]")
    fan = fan.contiguous()
    B, nu_f, nview_f, nv_f, C = fan.shape
    if out is None:
        out = torch.empty((B, nu_p, nview_p, nv_p, C), dtype=torch.float32, device=fan.device)
    check(_sigs().rn_helical_rebin(_ptr(fan), _ptr(out), B, C, nu_f, nview_f, nv_f, du_f, dv_f, off_u_f, off_v_f,
                                   beta0, dbeta, nu_p, nview_p, nv_p, du_p, dv_p, off_u_p, off_v_p, theta0, dtheta,
                                   dsd, dso, _stream()), "rn_helical_rebin")
    return out


def helical_backproject(par, nview_per_pi, theta0, vol_z0, zrot, *, nx, ny, nz, dx, dy, dz, du, dv, dsd, dso,
                        off_u=0.0, off_v=0.0, m_pi=1, Q=0.5, out=None):
    """cHelicalBPFromParallelConebeamBackprojection on the GPU: par float32
    CUDA tensor [B, nu, nview, nv, C] of parallel-cone data (normally ramp
    filtered along u); returns the volume [B, nx, ny, nz, C]."""
    import torch

    from .torch_ops import _ptr, _stream

    if par.dim() != 5 or par.dtype != torch.float32 or not par.is_cuda:
        raise ValueError("par must be a float32 CUDA tensor [B, nu, nview, nv, C]")
    par = par.contiguous()
    B, nu, nview, nv, C = par.shape
    if out is None:
        out = torch.empty((B, nx, ny, nz, C), dtype=torch.float32, device=par.device)
    check(_sigs().rn_helical_bp(_ptr(par), _ptr(out), B, C, nx, ny, nz, dx, dy, dz, nu, nview, nv, du, dv, off_u,
                                off_v, dsd, dso, nview_per_pi, theta0, vol_z0, zrot, m_pi, Q, _stream()),
          "rn_helical_bp")
    return out
