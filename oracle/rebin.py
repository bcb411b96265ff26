"""FP64 restatement of the helical fan -> parallel-cone rebinning.

TEST INFRASTRUCTURE ONLY (no reference counterpart: the reference's helical
backprojection, Recon3D/HelicalBP.cu:10-83, consumes rebinned data).  The
parallel ray (t, u) comes from the source at angle b = t - asin(u / dso); on
the flat detector of that view it meets column s = dsd tan(g) and, for the
row v of the cylinder of radius dsd, height w = v / cos(g), g = asin(u / dso).
The raw data are interpolated trilinearly at (column, view, row), samples at
integer indices, zero outside.
"""

import numpy as np

__all__ = ["rebin_fan_to_parallel"]


def rebin_fan_to_parallel(fan, du_f, dv_f, beta0, dbeta, nu_p, nview_p, nv_p, du_p, dv_p, theta0, dtheta, dsd, dso,
                          off_u_f=0.0, off_v_f=0.0, off_u_p=0.0, off_v_p=0.0, _row_scale=True, _gamma_sign=1.0):
    """_row_scale / _gamma_sign plant restatement mistakes (tests show that the
    geometric pin rejects them)."""
    fan = np.asarray(fan, np.float64)
    B, nu_f, nview_f, nv_f, C = fan.shape
    f32 = lambda x: float(np.float32(x))     # scalars as the C-ABI receives them
    du_f, dv_f, beta0, dbeta, du_p, dv_p = map(f32, (du_f, dv_f, beta0, dbeta, du_p, dv_p))
    theta0, dtheta, dsd, dso = map(f32, (theta0, dtheta, dsd, dso))
    off_u_f, off_v_f, off_u_p, off_v_p = map(f32, (off_u_f, off_v_f, off_u_p, off_v_p))
    out = np.zeros((B, nu_p, nview_p, nv_p, C))
    v = (np.arange(nv_p) - 0.5 * (nv_p - 1) - off_v_p) * dv_p
    for l in range(nu_p):
        u = (l - 0.5 * (nu_p - 1) - off_u_p) * du_p
        sg = u / dso
        if abs(sg) >= 1:
            continue
        g = np.arcsin(sg)
        fc = dsd * np.tan(g) / du_f + off_u_f + 0.5 * (nu_f - 1)
        fr = (v / np.cos(g) if _row_scale else v) / dv_f + off_v_f + 0.5 * (nv_f - 1)
        c0 = int(np.floor(fc))
        wc = fc - c0
        r0 = np.floor(fr).astype(int)
        wr = fr - r0
        for k in range(nview_p):
            fm = (theta0 + k * dtheta - _gamma_sign * g - beta0) / dbeta
            m0 = int(np.floor(fm))
            wm = fm - m0
            acc = np.zeros((B, nv_p, C))
            for dc, wcc in ((0, 1 - wc), (1, wc)):
                for dm, wmm in ((0, 1 - wm), (1, wm)):
                    cc, mm = c0 + dc, m0 + dm
                    if not (0 <= cc < nu_f and 0 <= mm < nview_f):
                        continue
                    for dr in (0, 1):
                        rr = r0 + dr
                        wrr = np.where(dr, wr, 1 - wr)
                        ok = (rr >= 0) & (rr < nv_f)
                        vals = np.zeros((B, nv_p, C))
                        vals[:, ok, :] = fan[:, cc, mm, rr[ok], :]
                        acc += (wcc * wmm * wrr)[None, :, None] * vals
            out[:, l, k] = acc
    return out
