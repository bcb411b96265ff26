"""Geometric pins of the oracle's backprojectors and arbitrary-geometry Siddon.

Each oracle function here is compared with a restatement built from the
physical geometry in numpy: source and detector positions, ray directions
from vectors, detector coordinates from atan2 or plane intersections, and
footprint integrals from direct interval overlaps (no prefix sums, no texture
emulation).  Every pin also "plants" the mistakes a restatement can make --
a mirrored detector axis, a half-bin shift, a reversed rotation, a segment
start off by one view -- into the numpy restatement and checks that the
oracle then disagrees with it by far more than the tolerance, so that an
oracle carrying such a mistake could not pass.

References: PixelDriven.cu:11-75, DistanceDriven.cu:441-614,
DistanceDriven3D.cu:409-559, Recon3D/HelicalBP.cu:10-83,
SiddonParallel.cu:11-105, ossqs.cu:163-333.
"""

import numpy as np
import pytest

import oracle as orc
from reconnet_amd.geometry import fan_angles
from tests.helpers import rel_l2, small_fan

F32_EPS6 = float(np.float32(1e-6))
# Projector.h:5 defines PI = 3.141592657f, i.e. float32(pi); the helical view
# angles and pitch are defined with it
PI_F = float(np.float32(np.pi))


def _f(x):
    return float(np.float32(x))


def _g32(g):
    """Geometry scalars as the C-ABI receives them (float arguments)."""
    return {k: _f(v) for k, v in g.items()}


# ------------------------------------------------------------ fan helpers ---

def _source(geo, th):
    return np.array([geo.dso * np.sin(th), -geo.dso * np.cos(th)])


def _det_angle(geo, th, p, mirror=False):
    """Signed angle between the central ray (source -> isocentre) and the ray
    source -> p, positive towards the rotated x axis (cos t, sin t)."""
    s = _source(geo, th)
    r = p - s
    ex = np.array([np.cos(th), np.sin(th)])
    ey = -s / np.linalg.norm(s)
    a = np.arctan2(r @ ex, r @ ey)
    return -a if mirror else a


def _bin_angle(geo, iu):
    return (-(iu - (geo.nu - 1) / 2.0) - _f(geo.off_a)) * _f(geo.da)


def _pixel_centres(geo):
    x = (np.arange(geo.nx) - (geo.nx - 1) / 2.0) * geo.dx
    y = (np.arange(geo.ny) - (geo.ny - 1) / 2.0) * geo.dy
    return x, y


def _lerp_zero_border(row, t):
    """Linear interpolation of samples row[m] placed at t = m, zero outside."""
    m0 = int(np.floor(t))
    w = t - m0
    v0 = row[m0] if 0 <= m0 < len(row) else 0.0
    v1 = row[m0 + 1] if 0 <= m0 + 1 < len(row) else 0.0
    return (1 - w) * v0 + w * v1


def _box_avg(row, lo, hi):
    """Average over [lo, hi] of the piecewise-constant function equal to
    row[m] on [m, m+1), zero outside [0, len(row))."""
    if hi < lo:
        lo, hi = hi, lo
    m = np.arange(len(row))
    ov = np.clip(np.minimum(hi, m + 1) - np.maximum(lo, m), 0, None)
    return float(row @ ov) / (hi - lo)


def pixel_bp_geometric(geo, deg, sino, mirror=False, shift=0.0):
    """Pixel-driven FBP backprojection of sino[nu, nview] (nz = nv = 1):
    the ray from the source through the pixel centre meets the detector at
    the continuous bin index whose angle formula inverts _bin_angle; the
    sinogram is interpolated linearly there (zero outside) and weighted by
    dso^2 / |p - source|^2."""
    xs, ys = _pixel_centres(geo)
    out = np.zeros((geo.nx, geo.ny))
    for k, th in enumerate(deg.astype(np.float64)):
        s = _source(geo, th)
        for i, x in enumerate(xs):
            for j, y in enumerate(ys):
                p = np.array([x, y])
                a = _det_angle(geo, th, p, mirror)
                pu = -(a / _f(geo.da) + _f(geo.off_a)) + (geo.nu - 1) / 2.0 + shift
                out[i, j] += _lerp_zero_border(sino[:, k], pu) * geo.dso ** 2 / np.sum((p - s) ** 2)
    return out


def dd_bp_geometric(geo, deg, sino, fbp, mirror=False, shift=0.0):
    """Distance-driven fan backprojection of sino[nu, nview] (nz = nv = 1).
    Per view the pixel's two edges normal to the view's main axis (x +- dx/2
    at the centre y for X views, y +- dy/2 at the centre x for Y views; a view
    is an X view iff |cos t| > |sin t| in float32) are projected from the
    source onto the detector; the pixel value is the average of the
    (pre-weighted) sinogram over that footprint, where bin m covers
    [m, m+1) of the coordinate b = -(a/da + off_a) + nu/2.  Iterative mode
    pre-weights bin m by dy/|cos(t - a_m)| (X views) or dx/|sin(t - a_m)|
    (Y views); FBP mode weights each view by dso^2 / |p - source|^2."""
    cls = orc.split_views(deg)
    xs, ys = _pixel_centres(geo)
    out = np.zeros((geo.nx, geo.ny))
    am = np.array([_bin_angle(geo, m) for m in range(geo.nu)])
    for k, th in enumerate(deg.astype(np.float64)):
        s = _source(geo, th)
        row = sino[:, k].astype(np.float64)
        if not fbp:
            row = row * (geo.dy / np.abs(np.cos(th - am)) if cls[k] == 0 else geo.dx / np.abs(np.sin(th - am)))
        for i, x in enumerate(xs):
            for j, y in enumerate(ys):
                if cls[k] == 0:
                    e = [np.array([x - geo.dx / 2, y]), np.array([x + geo.dx / 2, y])]
                else:
                    e = [np.array([x, y - geo.dy / 2]), np.array([x, y + geo.dy / 2])]
                b = [-(_det_angle(geo, th, q, mirror) / _f(geo.da) + _f(geo.off_a)) + geo.nu / 2.0 + shift for q in e]
                val = _box_avg(row, b[0], b[1])
                if fbp:
                    p = np.array([x, y])
                    val *= geo.dso ** 2 / np.sum((p - s) ** 2)
                out[i, j] += val
    return out


def _fan_case(off_a):
    geo = small_fan(nx=14, nu=40, nview=24, du=3.0, off_a=off_a)
    deg = fan_angles(geo.nview)
    sino = np.random.default_rng(21).uniform(0, 1, (geo.nu, geo.nview)).astype(np.float32)
    return geo, deg, sino


@pytest.mark.parametrize("off_a", [0.0, 1.3])
def test_pixel_bp_matches_geometric_restatement(off_a):
    geo, deg, sino = _fan_case(off_a)
    o = orc.pixel_fan_bp(sino[None, :, :, None, None], deg, geo.oracle_geo(), (geo.nx, geo.ny, 1))[0, :, :, 0, 0]
    assert rel_l2(o, pixel_bp_geometric(geo, deg, sino)) < 1e-12
    # planted mistakes: mirrored detector axis, half-bin shift
    assert rel_l2(o, pixel_bp_geometric(geo, deg, sino, mirror=True)) > 1e-2
    assert rel_l2(o, pixel_bp_geometric(geo, deg, sino, shift=0.5)) > 1e-2


@pytest.mark.parametrize("fbp", [False, True])
@pytest.mark.parametrize("off_a", [0.0, 1.3])
def test_dd_bp_matches_geometric_restatement(fbp, off_a):
    geo, deg, sino = _fan_case(off_a)
    o = orc.dd_fan_bp(sino[None, :, :, None, None], deg, geo.oracle_geo(), (geo.nx, geo.ny, 1),
                      type_projector=1 if fbp else 0)[0, :, :, 0, 0]
    assert rel_l2(o, dd_bp_geometric(geo, deg, sino, fbp)) < 1e-12
    assert rel_l2(o, dd_bp_geometric(geo, deg, sino, fbp, mirror=True)) > 1e-2
    assert rel_l2(o, dd_bp_geometric(geo, deg, sino, fbp, shift=0.5)) > 1e-2


def _ray_dist(p, s, d):
    """Distance of point p from the line s + t d (d unit)."""
    r = p - s
    return abs(r[0] * d[1] - r[1] * d[0])


@pytest.mark.parametrize("fn", ["pixel", "dd_fbp"])
def test_single_bin_backprojection_lies_on_its_ray(fn):
    # one view, one off-centre bin: the backprojected stripe follows the ray
    # from the source through that bin's centre (a mirrored or shifted
    # detector puts it on another ray)
    geo = small_fan(nx=48, nu=96, nview=36, du=2.0, off_a=0.7)
    deg = fan_angles(geo.nview)
    k, j = 5, 60
    sino = np.zeros((1, geo.nu, geo.nview, 1, 1), np.float32)
    sino[0, j, k] = 1.0
    if fn == "pixel":
        img = orc.pixel_fan_bp(sino, deg, geo.oracle_geo(), (geo.nx, geo.ny, 1))[0, :, :, 0, 0]
    else:
        img = orc.dd_fan_bp(sino, deg, geo.oracle_geo(), (geo.nx, geo.ny, 1), type_projector=1)[0, :, :, 0, 0]
    th = float(deg[k])
    s = _source(geo, th)
    a = _bin_angle(geo, j)
    # central direction rotated by a towards the rotated x axis (cos t, sin t)
    ex, ey = np.array([np.cos(th), np.sin(th)]), -s / geo.dso
    d = np.sin(a) * ex + np.cos(a) * ey
    xs, ys = _pixel_centres(geo)
    hit = img > 0.5 * img.max()
    assert hit.sum() >= geo.nx // 2
    dist = np.array([_ray_dist(np.array([xs[i], ys[jj]]), s, d) for i, jj in zip(*np.nonzero(hit))])
    # a bin is du * dso / dsd ~ 1.1 mm wide at the isocentre; pixels are 1 mm
    assert dist.max() < 1.0
    # the mirrored ray misses the stripe
    dm = np.sin(-a) * ex + np.cos(-a) * ey
    distm = np.array([_ray_dist(np.array([xs[i], ys[jj]]), s, dm) for i, jj in zip(*np.nonzero(hit))])
    assert distm.mean() > 5.0


# -------------------------------------------------------- DD tomo BP ----------

def dd_tomo_bp_geometric(prj, C, S, g, shape, nu, nv, shift=0.0, mirror=False):
    """DD tomosynthesis backprojection (detector normal to z, axes x and y).
    Detector cell (iu, iv) centre: C + (u, v, 0) with
    u = (iu - off_u - (nu-1)/2) du, v = (iv - off_v - (nv-1)/2) dv; its value
    is pre-weighted by dz |S - P| / |S.z - P.z|.  A voxel's footprint is the
    rectangle spanned by the central projections (from S onto the plane
    z = C.z) of x -+ dx/2 (for u) and y -+ dy/2 (for v); in the continuous
    cell coordinate t = u/du + off_u + (nu-1)/2 the cell iu covers
    [iu, iu+1) (likewise for v).  Voxel value = sum over views of the
    footprint average."""
    g = _g32(g)
    nx, ny, nz = shape
    nview = C.shape[0]
    out = np.zeros(shape)
    us = (np.arange(nu) - g["off_u"] - (nu - 1) / 2.0) * g["du"]
    vs = (np.arange(nv) - g["off_v"] - (nv - 1) / 2.0) * g["dv"]
    xs = (np.arange(nx) - (nx - 1) / 2.0) * g["dx"] + g.get("cx", 0.0)
    ys = (np.arange(ny) - (ny - 1) / 2.0) * g["dy"] + g.get("cy", 0.0)
    zs = (np.arange(nz) - (nz - 1) / 2.0) * g["dz"] + g.get("cz", 0.0)
    for k in range(nview):
        c, s = C[k].astype(np.float64), S[k].astype(np.float64)
        U, V = np.meshgrid(us, vs, indexing="ij")
        P = np.stack([c[0] + U, c[1] + V, np.full_like(U, c[2])], -1)
        wgt = g["dz"] * np.linalg.norm(P - s, axis=-1) / np.abs(s[2] - c[2])
        cell = prj[:, k, :].astype(np.float64) * wgt

        def hit(q):
            # intersection of the line S -> q with the plane z = C.z
            t = (c[2] - s[2]) / (q[2] - s[2])
            h = s + t * (q - s)
            return h[0] - c[0], h[1] - c[1]

        for i, x in enumerate(xs):
            for j, y in enumerate(ys):
                for l, z in enumerate(zs):
                    u1 = hit(np.array([x - g["dx"] / 2, y, z]))[0]
                    u2 = hit(np.array([x + g["dx"] / 2, y, z]))[0]
                    v1 = hit(np.array([x, y - g["dy"] / 2, z]))[1]
                    v2 = hit(np.array([x, y + g["dy"] / 2, z]))[1]
                    if mirror:
                        u1, u2 = -u1, -u2
                    t = [u / g["du"] + g["off_u"] + (nu - 1) / 2.0 + shift for u in (u1, u2)]
                    r = [v / g["dv"] + g["off_v"] + (nv - 1) / 2.0 for v in (v1, v2)]
                    mu = np.arange(nu)
                    mv = np.arange(nv)
                    ou = np.clip(np.minimum(max(t), mu + 1) - np.maximum(min(t), mu), 0, None)
                    ov = np.clip(np.minimum(max(r), mv + 1) - np.maximum(min(r), mv), 0, None)
                    out[i, j, l] += ou @ cell @ ov / ((max(t) - min(t)) * (max(r) - min(r)))
    return out


def _tomo_case():
    nview = 5
    th = np.linspace(-0.3, 0.3, nview)
    S = np.stack([220 * np.sin(th) + 7.0, 15 * np.cos(th) - 4.0, 300 * np.cos(th)], 1).astype(np.float32)
    C = np.tile([2.0, -3.0, -50.0], (nview, 1)).astype(np.float32)
    g = dict(dx=1.0, dy=1.2, dz=1.5, cx=0.5, cy=-0.7, cz=3.0, du=1.3, dv=1.1, off_u=0.4, off_v=-0.3)
    return S, C, g


def test_dd_tomo_bp_matches_geometric_restatement():
    S, C, g = _tomo_case()
    nu, nv, shape = 24, 20, (8, 7, 4)
    prj = np.random.default_rng(22).uniform(0, 1, (1, nu, C.shape[0], nv, 1)).astype(np.float32)
    o = orc.dd_tomo_bp(prj, C, S, g, shape)[0, ..., 0]
    ref = dd_tomo_bp_geometric(prj[0, :, :, :, 0], C, S, g, shape, nu, nv)
    assert rel_l2(o, ref) < 1e-10
    assert rel_l2(o, dd_tomo_bp_geometric(prj[0, :, :, :, 0], C, S, g, shape, nu, nv, shift=0.5)) > 1e-2
    assert rel_l2(o, dd_tomo_bp_geometric(prj[0, :, :, :, 0], C, S, g, shape, nu, nv, mirror=True)) > 1e-2


def test_dd_tomo_bp_of_single_cell_lies_on_its_ray():
    # one detector cell lit in one view: the voxels that receive it are
    # those on the line from the source through that cell
    S, C, g = _tomo_case()
    nu, nv, shape = 40, 36, (16, 14, 6)
    prj = np.zeros((1, nu, C.shape[0], nv, 1), np.float32)
    k, iu, iv = 2, 20, 15
    prj[0, iu, k, iv, 0] = 1.0
    o = orc.dd_tomo_bp(prj, C, S, g, shape)[0, ..., 0]
    cell = C[k].astype(np.float64) + [(iu - g["off_u"] - (nu - 1) / 2.0) * g["du"],
                                       (iv - g["off_v"] - (nv - 1) / 2.0) * g["dv"], 0.0]
    s = S[k].astype(np.float64)
    d = (cell - s) / np.linalg.norm(cell - s)
    xs = (np.arange(shape[0]) - (shape[0] - 1) / 2.0) * g["dx"] + g["cx"]
    ys = (np.arange(shape[1]) - (shape[1] - 1) / 2.0) * g["dy"] + g["cy"]
    zs = (np.arange(shape[2]) - (shape[2] - 1) / 2.0) * g["dz"] + g["cz"]
    # every z layer has a voxel lit; the lit voxels' centres lie within about
    # a voxel of the ray; cell centres are half a cell off the sampled
    # footprint coordinate (cell iu covers [iu, iu+1) of t)
    for l, z in enumerate(zs):
        lay = o[:, :, l]
        assert lay.max() > 0
        i, j = np.unravel_index(np.argmax(lay), lay.shape)
        tpar = (z - s[2]) / d[2]
        hx, hy = s[0] + tpar * d[0], s[1] + tpar * d[1]
        assert abs(xs[i] - hx) < 1.6 * g["dx"] and abs(ys[j] - hy) < 1.6 * g["dy"], (l, xs[i], hx, ys[j], hy)


# ------------------------------------------------------------ helical BP ----

def helical_bp_geometric(prj, vox, g, nview_per_pi, theta0, vol_z0, zrot, m_pi, Q,
                         mirror=False, shift=0.0, seg=0, vsign=1.0):
    """Helical BP of rebinned parallel-cone data at one voxel (ix, iy, iz).
    Parallel view k has angle t_k = theta0 + k PI / nview_per_pi (PI = float32(pi)) and rays of
    direction -(cos t, sin t) in xy; the ray through p meets the source
    circle |S| = dso upstream at S = p + s (cos t, sin t) (s > 0 from the
    quadratic), the source's helical angle b is its polar angle on the branch
    nearest t and its height zrot (b - theta0) / (2 pi).  u is the signed
    distance of the ray from the axis along (sin t, -cos t); v the height of
    p above the source, magnified by dsd / s.  Segment views i start at
    trunc((2 pi z / zrot - pi/4) nview_per_pi / pi); each i combines views
    k = i + m nview_per_pi (|m| <= m_pi) with weights w(q), q = |2 v / (nv dv)|
    (1 up to Q, cos^2 ramp to 0 at q = 1), normalised by their sum + 1e-6."""
    nu, nview, nv = prj.shape
    nx, ny = g["nx"], g["ny"]
    g = _g32(g)
    theta0, vol_z0, zrot, Q = _f(theta0), _f(vol_z0), _f(zrot), _f(Q)
    ix, iy, iz = vox
    p = np.array([(ix - (nx - 1) / 2.0) * g["dx"], (iy - (ny - 1) / 2.0) * g["dy"]])
    z = iz * g["dz"] + vol_z0
    i0 = int((z / zrot * 2 * PI_F - PI_F / 4) * nview_per_pi / PI_F) + seg
    total = 0.0
    for i in range(i0, i0 + nview_per_pi):
        val, tw = 0.0, 0.0
        for k in range(i - m_pi * nview_per_pi, i + m_pi * nview_per_pi + 1, nview_per_pi):
            if k < 0 or k >= nview:
                continue
            t = theta0 + k * PI_F / nview_per_pi
            d = np.array([np.cos(t), np.sin(t)])
            pd = p @ d
            s = -pd + np.sqrt(pd * pd - p @ p + g["dso"] ** 2)
            S = p + s * d
            b = np.arctan2(S[1], S[0])
            b += 2 * np.pi * np.round((t - b) / (2 * np.pi))
            zs = zrot * (b - theta0) / (2 * PI_F)
            u = p @ np.array([np.sin(t), -np.cos(t)])
            if mirror:
                u = -u
            v = vsign * (z - zs) * g["dsd"] / s
            q = abs(2 * v / (nv * g["dv"]))
            w = 1.0 if q <= Q else (np.cos((q - Q) / (1 - Q) * PI_F / 2) ** 2 if q < 1 else 0.0)
            tu = u / g["du"] + (nu - 1) / 2.0 + g["off_u"] + shift
            tv = v / g["dv"] + (nv - 1) / 2.0 + g["off_v"]
            # bilinear interpolation, samples at integer (tu, tv), zero outside
            a0, b0 = int(np.floor(tu)), int(np.floor(tv))
            fu, fv = tu - a0, tv - b0
            acc = 0.0
            for da_, wu in ((0, 1 - fu), (1, fu)):
                for db_, wv in ((0, 1 - fv), (1, fv)):
                    a, bb = a0 + da_, b0 + db_
                    if 0 <= a < nu and 0 <= bb < nv:
                        acc += wu * wv * prj[a, k, bb]
            val += acc * w
            tw += w
        total += val / (tw + F32_EPS6)
    return total


def _helical_case():
    npp, turns, nu, nv = 24, 2, 48, 16
    g = dict(nx=16, ny=16, dx=1.0, dy=1.0, dz=1.0, du=1.0, dv=1.0, off_u=0.3, off_v=-0.4, dsd=1000.0, dso=600.0)
    prj = np.random.default_rng(23).uniform(0, 1, (nu, 2 * npp * turns, nv)).astype(np.float32)
    return npp, g, prj, dict(theta0=0.2, vol_z0=4.1, zrot=12.0, m_pi=1, Q=0.5)


@pytest.mark.parametrize("m_pi", [1, 0])
def test_helical_bp_matches_geometric_restatement(m_pi):
    npp, g, prj, h = _helical_case()
    h = dict(h, m_pi=m_pi)
    nz = 6
    o = orc.helical_bp(prj[None, :, :, :, None], g, (g["nx"], g["ny"], nz), npp, h["theta0"], h["vol_z0"],
                       h["zrot"], h["m_pi"], h["Q"])[0, ..., 0]
    voxels = [(3, 11, 0), (8, 8, 2), (12, 2, 5), (0, 15, 3), (15, 6, 1)]
    ref = np.array([helical_bp_geometric(prj, v, g, npp, **h) for v in voxels])
    got = np.array([o[v] for v in voxels])
    assert rel_l2(got, ref) < 1e-9
    # planted mistakes: mirrored u, half-bin u shift, v sign; with m_pi = 0
    # also a segment start one view late (with m_pi = 1 the normalised
    # conjugate pairs make a one-view shift of the segment exactly invisible)
    plants = [dict(mirror=True), dict(shift=0.5), dict(vsign=-1.0)] + ([dict(seg=1)] if m_pi == 0 else [])
    for kw in plants:
        bad = np.array([helical_bp_geometric(prj, v, g, npp, **h, **kw) for v in voxels])
        assert rel_l2(got, bad) > 1e-3, kw


def test_helical_case_exercises_every_weight_branch():
    # the geometry of the restatement test reaches q <= Q, Q < q < 1 and q >= 1
    npp, g, prj, h = _helical_case()
    nu, nview, nv = prj.shape
    qs = []
    for vox in [(3, 11, 0), (8, 8, 2), (12, 2, 5)]:
        ix, iy, iz = vox
        p = np.array([(ix - 7.5) * g["dx"], (iy - 7.5) * g["dy"]])
        z = iz * g["dz"] + h["vol_z0"]
        for k in range(nview):
            t = h["theta0"] + k * np.pi / npp
            d = np.array([np.cos(t), np.sin(t)])
            s = -(p @ d) + np.sqrt((p @ d) ** 2 - p @ p + g["dso"] ** 2)
            S = p + s * d
            b = np.arctan2(S[1], S[0])
            b += 2 * np.pi * np.round((t - b) / (2 * np.pi))
            v = (z - h["zrot"] * (b - h["theta0"]) / (2 * np.pi)) * g["dsd"] / s
            qs.append(abs(2 * v / (nv * g["dv"])))
    qs = np.array(qs)
    assert (qs <= h["Q"]).any() and ((qs > h["Q"]) & (qs < 1)).any() and (qs >= 1).any()


# ------------------------------------------------------ Siddon parallel -----

def _par_geometry(nview, z_tilt=0.0, dist=40.0):
    th = fan_angles(nview).astype(np.float64) + 0.1
    C = np.stack([-dist * np.sin(th), dist * np.cos(th), np.full_like(th, 1.5)], 1)
    U = np.stack([np.cos(th), np.sin(th), np.zeros_like(th)], 1)
    V = np.tile([0.0, 0.0, 1.0], (nview, 1))
    # from the detector back through the volume; optionally tilted in z
    D = np.stack([np.sin(th), -np.cos(th), np.full_like(th, z_tilt)], 1)
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    return [a.astype(np.float32) for a in (C, U, V, D)]


def _par_rays(C, U, V, D, g, nu, nv, nx, ny, nz):
    """(point on the detector, unit direction towards the source, length of
    the traced segment) per ray.  The source is placed on the ray at
    1.1 max(|grid centre - detector centre|, grid half-diagonal) from the
    detector point (SiddonParallel.cu:11-32), so with a detector far from
    the grid the segment can end inside the volume."""
    g = _g32(g)
    ctr = np.array([g["cx"], g["cy"], g["cz"]])
    half = np.linalg.norm([nx * g["dx"], ny * g["dy"], nz * g["dz"]]) / 2
    out = []
    for k in range(C.shape[0]):
        c, uu, vv, dd = (a[k].astype(np.float64) for a in (C, U, V, D))
        m = _f(1.1) * max(np.linalg.norm(ctr - c), half)
        for iu in range(nu):
            for iv in range(nv):
                u = (iu - g["off_u"] - (nu - 1) / 2.0) * g["du"]
                v = (iv - g["off_v"] - (nv - 1) / 2.0) * g["dv"]
                p = c + u * uu + v * vv
                out.append((k, iu, iv, p, dd / np.linalg.norm(dd), m * np.linalg.norm(dd)))
    return out


def _slab(p, d, lo, hi, tmax=np.inf):
    """Length of {p + t d : 0 <= t <= tmax} inside the box [lo, hi] (Liang-Barsky)."""
    t0, t1 = 0.0, tmax
    for a in range(3):
        if abs(d[a]) < 1e-15:
            if p[a] < lo[a] or p[a] > hi[a]:
                return 0.0
            continue
        ta, tb = (lo[a] - p[a]) / d[a], (hi[a] - p[a]) / d[a]
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    return max(0.0, t1 - t0)


def _par_g(**kw):
    g = dict(dx=1.0, dy=1.25, dz=0.8, cx=1.5, cy=-2.0, cz=0.5, du=1.1, dv=0.9, off_u=0.35, off_v=-0.2)
    g.update(kw)
    return g


@pytest.mark.parametrize("z_tilt", [0.0, 0.05])
def test_siddon_parallel_fp_of_ones_is_chord_length(z_tilt):
    nx, ny, nz, nu, nv = 20, 16, 10, 40, 9
    g = _par_g()
    C, U, V, D = _par_geometry(12, z_tilt)
    p = orc.siddon_par_fp(np.ones((1, nx, ny, nz, 1), np.float32), C, U, V, D, g, nu, nv)[0, ..., 0]
    g = _g32(g)
    lo = np.array([g["cx"] - nx * g["dx"] / 2, g["cy"] - ny * g["dy"] / 2, g["cz"] - nz * g["dz"] / 2])
    hi = lo + [nx * g["dx"], ny * g["dy"], nz * g["dz"]]
    ref = np.zeros_like(p)
    for k, iu, iv, q, d, tmax in _par_rays(C, U, V, D, g, nu, nv, nx, ny, nz):
        ref[iu, k, iv] = _slab(q, d, lo, hi, tmax)
    assert (ref > 0).sum() > ref.size // 8
    assert np.max(np.abs(p - ref)) < 1e-9 * ref.max()
    # planted mistakes: the ray travelling the other way sees nothing; a
    # half-cell shift of u moves the chords
    assert np.all(orc.siddon_par_fp(np.ones((1, nx, ny, nz, 1), np.float32), C, U, V, -D, g, nu, nv) == 0)
    ps = orc.siddon_par_fp(np.ones((1, nx, ny, nz, 1), np.float32), C, U, V, D, _par_g(off_u=0.85), nu, nv)[0, ..., 0]
    assert rel_l2(ps, ref) > 1e-2


def test_siddon_parallel_fp_brute_force_voxel_clipping():
    nx, ny, nz, nu, nv = 6, 5, 4, 14, 6
    g = _par_g(du=1.4, dv=0.7)
    C, U, V, D = _par_geometry(7, 0.2, dist=12.0)
    img = np.random.default_rng(24).uniform(0, 1, (nx, ny, nz)).astype(np.float32)
    p = orc.siddon_par_fp(img[None, ..., None], C, U, V, D, g, nu, nv)[0, ..., 0]
    g = _g32(g)
    lo0 = np.array([g["cx"] - nx * g["dx"] / 2, g["cy"] - ny * g["dy"] / 2, g["cz"] - nz * g["dz"] / 2])
    step = np.array([g["dx"], g["dy"], g["dz"]])
    ref = np.zeros_like(p)
    for k, iu, iv, q, d, tmax in _par_rays(C, U, V, D, g, nu, nv, nx, ny, nz):
        s = 0.0
        for i in range(nx):
            for j in range(ny):
                for l in range(nz):
                    lo = lo0 + step * [i, j, l]
                    s += img[i, j, l] * _slab(q, d, lo, lo + step, tmax)
        ref[iu, k, iv] = s
    assert rel_l2(p, ref) < 1e-12


def test_siddon_parallel_bp_is_transpose_of_fp():
    nx, ny, nz, nu, nv = 12, 10, 5, 24, 7
    g = _par_g()
    C, U, V, D = _par_geometry(9, 0.1)
    rng = np.random.default_rng(25)
    x = rng.uniform(0, 1, (1, nx, ny, nz, 1)).astype(np.float32)
    y = rng.uniform(0, 1, (1, nu, 9, nv, 1)).astype(np.float32)
    ax = orc.siddon_par_fp(x, C, U, V, D, g, nu, nv)
    aty = orc.siddon_par_bp(y, C, U, V, D, g, (nx, ny, nz))
    lhs, rhs = float(np.sum(ax * y)), float(np.sum(x * aty))
    assert lhs > 0 and abs(lhs - rhs) / abs(lhs) < 1e-12


# ------------------------------------------------------- cone OS-SQS --------

def _cone_vecs(nview, dso=300.0, dsd=500.0):
    th = fan_angles(nview).astype(np.float64)
    S = np.stack([dso * np.sin(th), -dso * np.cos(th), np.zeros_like(th)], 1)
    C = np.stack([-(dsd - dso) * np.sin(th), (dsd - dso) * np.cos(th), np.zeros_like(th)], 1)
    U = np.stack([np.cos(th), np.sin(th), np.zeros_like(th)], 1)
    V = np.tile([0.0, 0.0, 1.0], (nview, 1))
    return [a.astype(np.float32) for a in (C, U, V, S)]


def test_cone_ossqs_fixed_point_and_mask():
    # b = A x* with x* = 0 outside the mask: a pass started at x* leaves it
    # unchanged; a pass started elsewhere zeroes the masked voxels after every
    # subset (ossqs.cu:306-308) and moves towards x*
    C, U, V, S = _cone_vecs(16)
    g = dict(dx=1.0, dy=1.0, dz=1.0, du=1.5, dv=1.5, off_u=0.3, off_v=-0.2)
    shape = (12, 12, 6)
    xs = np.random.default_rng(26).uniform(0.2, 1, (1,) + shape + (1,)).astype(np.float32)
    xs[:, :3] = 0   # x* is zero where the mask is 0, so the masked update keeps it
    b = orc.siddon_cone_fp(xs, C, U, V, S, g, 24, 10).astype(np.float32)
    D = orc.siddon_cone_bp(orc.siddon_cone_fp(np.ones_like(xs), C, U, V, S, g, 24, 10).astype(np.float32),
                           C, U, V, S, g, shape).astype(np.float32)
    mask = np.ones_like(xs)
    mask[:, :3] = 0
    out = orc.os_sqs_cone(xs, b, D, mask, C, U, V, S, g, 4, lb=0, ub=2)
    assert rel_l2(out, xs) < 1e-6
    x0 = np.clip(xs + np.random.default_rng(27).normal(0, 0.1, xs.shape), 0, 2).astype(np.float32)
    out = orc.os_sqs_cone(x0, b, D, mask, C, U, V, S, g, 4, lb=0, ub=2)
    assert np.all(out[:, :3] == 0)
    assert np.linalg.norm(out - xs) < 0.8 * np.linalg.norm(x0 - xs)
    # planted mistake: without the mask the masked voxels keep values
    out = orc.os_sqs_cone(x0, b, D, np.ones_like(xs), C, U, V, S, g, 4, lb=0, ub=2)
    assert np.abs(out[:, :3]).max() > 1e-2
