"""CPU oracle for the ReconNet projector / reconstruction path.

TEST INFRASTRUCTURE ONLY.  ``oracle.cpp`` restates the reference's CUDA kernel
bodies (wudufan/ReconNet, ``ReconNet4D_no_tf/*.cu`` and
``Recon3D/HelicalBP.cu``) in host C++ templated on the arithmetic type:
``*_f64`` functions are the parity oracle, ``*_f32`` functions the CPU timing
baseline.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
CPU-baseline leg import this package, and only as the checker.

Arrays follow the reference layouts: images ``[B, nx, ny, nz, C]`` and
sinograms ``[B, nu, nview, nv, C]`` (``ReconNet_no_tf.py:111,143``).
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

__all__ = [
    "lib", "build", "dd_fan_fp", "dd_fan_bp", "siddon_fan_fp", "siddon_fan_bp",
    "pixel_fan_bp", "ramp_kernel", "ramp_filter", "tv_sqs", "os_sqs",
    "siddon_par_fp", "siddon_par_bp", "siddon_cone_fp", "siddon_cone_bp",
    "os_sqs_cone", "dd_tomo_fp", "dd_tomo_bp", "helical_bp", "split_views",
    "set_threads", "num_threads",
]

_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_tv_sqs_f64.restype = ctypes.c_double
        _lib.orc_tv_sqs_f32.restype = ctypes.c_float
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fan_tail(nB, nC, nx, ny, nz, geo, nu, nview, nv, type_projector):
    return [ctypes.c_int(nB), ctypes.c_int(nC), ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz),
            ctypes.c_float(geo["dx"]), ctypes.c_float(geo["dy"]), ctypes.c_float(geo["dz"]),
            ctypes.c_int(nu), ctypes.c_int(nview), ctypes.c_int(nv),
            ctypes.c_float(geo["da"]), ctypes.c_float(geo["dv"]), ctypes.c_float(geo["off_a"]),
            ctypes.c_float(geo["off_v"]), ctypes.c_float(geo["dsd"]), ctypes.c_float(geo["dso"]),
            ctypes.c_int(type_projector)]


def _dt(precision):
    return np.float64 if precision == "f64" else np.float32


def set_threads(n):
    lib().orc_set_threads(ctypes.c_int(int(n)))


def num_threads():
    return int(lib().orc_num_threads())


def split_views(deg):
    """float32 X/Y split (DistanceDriven.cu:299-312): 0 = X view, 1 = Y view."""
    deg = _f32(deg)
    cls = np.zeros(deg.shape[0], np.int32)
    lib().orc_split_views(_p(cls), _p(deg), ctypes.c_int(deg.shape[0]))
    return cls


def _fan_proj(name, img, deg, geo, nu, nv, type_projector, precision):
    img = _f32(img)
    deg = _f32(deg)
    nB, nx, ny, nz, nC = img.shape
    out = np.zeros((nB, nu, deg.shape[0], nv, nC), _dt(precision))
    getattr(lib(), "orc_%s_%s" % (name, precision))(
        _p(out), _p(img), _p(deg), *_fan_tail(nB, nC, nx, ny, nz, geo, nu, deg.shape[0], nv, type_projector))
    return out


def _fan_back(name, prj, deg, geo, img_shape, type_projector, precision):
    prj = _f32(prj)
    deg = _f32(deg)
    nB, nu, nview, nv, nC = prj.shape
    nx, ny, nz = img_shape
    out = np.zeros((nB, nx, ny, nz, nC), _dt(precision))
    getattr(lib(), "orc_%s_%s" % (name, precision))(
        _p(out), _p(prj), _p(deg), *_fan_tail(nB, nC, nx, ny, nz, geo, nu, nview, nv, type_projector))
    return out


def dd_fan_fp(img, deg, geo, nu, nv=1, type_projector=0, precision="f64"):
    return _fan_proj("dd_fan_fp", img, deg, geo, nu, nv, type_projector, precision)


def dd_fan_bp(prj, deg, geo, img_shape, type_projector=0, precision="f64"):
    return _fan_back("dd_fan_bp", prj, deg, geo, img_shape, type_projector, precision)


def siddon_fan_fp(img, deg, geo, nu, nv=1, precision="f64"):
    return _fan_proj("siddon_fan_fp", img, deg, geo, nu, nv, 0, precision)


def siddon_fan_bp(prj, deg, geo, img_shape, precision="f64"):
    return _fan_back("siddon_fan_bp", prj, deg, geo, img_shape, 0, precision)


def pixel_fan_bp(prj, deg, geo, img_shape, precision="f64"):
    return _fan_back("pixel_fan_bp", prj, deg, geo, img_shape, 0, precision)


def ramp_kernel(nu, da, filter_type=0, equal_space=False):
    out = np.zeros(2 * nu - 1, np.float64)
    lib().orc_ramp_kernel_f64(_p(out), ctypes.c_int(nu), ctypes.c_float(da), ctypes.c_int(filter_type),
                              ctypes.c_int(1 if equal_space else 0))
    return out


def ramp_filter(prj, geo, filter_type=0, fan=True, precision="f64"):
    prj = _f32(prj)
    nB, nu, nview, nv, nC = prj.shape
    out = np.zeros(prj.shape, _dt(precision))
    getattr(lib(), "orc_filter_%s" % precision)(
        _p(out), _p(prj), ctypes.c_int(nB), ctypes.c_int(nC), ctypes.c_int(nu), ctypes.c_int(nview), ctypes.c_int(nv),
        ctypes.c_float(geo["da"]), ctypes.c_float(geo["dv"]), ctypes.c_float(geo["off_a"]), ctypes.c_float(geo["off_v"]),
        ctypes.c_float(geo["dsd"]), ctypes.c_float(geo["dso"]), ctypes.c_int(filter_type), ctypes.c_int(1 if fan else 0))
    return out


def tv_sqs(img2d, precision="f64"):
    img2d = _f32(img2d)
    nx, ny = img2d.shape
    g = np.zeros((nx, ny), _dt(precision))
    n = np.zeros((nx, ny), _dt(precision))
    tv = getattr(lib(), "orc_tv_sqs_%s" % precision)(_p(g), _p(n), _p(img2d), ctypes.c_int(nx), ctypes.c_int(ny))
    return g, n, float(tv)


def os_sqs(img, prj, norm_img, wprj, deg, geo, n_subsets, lb=None, ub=None, lam=1.0, precision="f64"):
    """One OS-SQS pass (ossqs.cu:12-161); nviewPerSubset = ceil(nview / nSubsets)."""
    img = _f32(img); prj = _f32(prj); norm_img = _f32(norm_img); wprj = _f32(wprj); deg = _f32(deg)
    nB, nx, ny, nz, nC = img.shape
    _, nu, nview, nv, _ = prj.shape
    if lb is None:
        lb = np.finfo(np.float32).min
    if ub is None:
        ub = np.finfo(np.float32).max
    npers = int(np.ceil(nview / float(n_subsets)))
    out = np.zeros(img.shape, _dt(precision))
    getattr(lib(), "orc_os_sqs_%s" % precision)(
        _p(out), _p(img), _p(prj), _p(norm_img), _p(wprj), _p(deg), ctypes.c_float(lb), ctypes.c_float(ub),
        ctypes.c_int(npers), ctypes.c_int(n_subsets), *_fan_tail(nB, nC, nx, ny, nz, geo, nu, nview, nv, 0),
        ctypes.c_float(lam))
    return out


def _arb_tail(nB, nC, nx, ny, nz, geo, nu, nview, nv):
    return [ctypes.c_int(nB), ctypes.c_int(nC), ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz),
            ctypes.c_float(geo["dx"]), ctypes.c_float(geo["dy"]), ctypes.c_float(geo["dz"]),
            ctypes.c_float(geo.get("cx", 0.0)), ctypes.c_float(geo.get("cy", 0.0)), ctypes.c_float(geo.get("cz", 0.0)),
            ctypes.c_int(nu), ctypes.c_int(nview), ctypes.c_int(nv),
            ctypes.c_float(geo["du"]), ctypes.c_float(geo["dv"]), ctypes.c_float(geo.get("off_u", 0.0)),
            ctypes.c_float(geo.get("off_v", 0.0))]


def _arb_fp(name, img, vecs, geo, nu, nv, precision):
    img = _f32(img)
    vecs = [_f32(v) for v in vecs]
    nB, nx, ny, nz, nC = img.shape
    nview = vecs[0].shape[0]
    out = np.zeros((nB, nu, nview, nv, nC), _dt(precision))
    getattr(lib(), "orc_%s_%s" % (name, precision))(_p(out), _p(img), *[_p(v) for v in vecs],
                                                    *_arb_tail(nB, nC, nx, ny, nz, geo, nu, nview, nv))
    return out


def _arb_bp(name, prj, vecs, geo, img_shape, precision):
    prj = _f32(prj)
    vecs = [_f32(v) for v in vecs]
    nB, nu, nview, nv, nC = prj.shape
    nx, ny, nz = img_shape
    out = np.zeros((nB, nx, ny, nz, nC), _dt(precision))
    getattr(lib(), "orc_%s_%s" % (name, precision))(_p(out), _p(prj), *[_p(v) for v in vecs],
                                                    *_arb_tail(nB, nC, nx, ny, nz, geo, nu, nview, nv))
    return out


def siddon_par_fp(img, det_center, det_u, det_v, inv_ray_dir, geo, nu, nv):
    return _arb_fp("siddon_par_fp", img, [det_center, det_u, det_v, inv_ray_dir], geo, nu, nv, "f64")


def siddon_par_bp(prj, det_center, det_u, det_v, inv_ray_dir, geo, img_shape):
    return _arb_bp("siddon_par_bp", prj, [det_center, det_u, det_v, inv_ray_dir], geo, img_shape, "f64")


def siddon_cone_fp(img, det_center, det_u, det_v, src, geo, nu, nv, precision="f64"):
    return _arb_fp("siddon_cone_fp", img, [det_center, det_u, det_v, src], geo, nu, nv, precision)


def siddon_cone_bp(prj, det_center, det_u, det_v, src, geo, img_shape, precision="f64"):
    return _arb_bp("siddon_cone_bp", prj, [det_center, det_u, det_v, src], geo, img_shape, precision)


def tv_iso_sqs(img, dims, eps=1e-8):
    """(s1, s2, var) of the isotropic TV surrogate; img [nb, nx, ny, nc] (dims 2)
    or [nb, nx, ny, nz, nc] (dims 3)."""
    img = _f32(img)
    if dims == 2:
        nb, nx, ny, nc = img.shape
        nz = 1
    else:
        nb, nx, ny, nz, nc = img.shape
    out = [np.zeros(img.shape, np.float64) for _ in range(3)]
    lib().orc_tv_iso_sqs_f64(*[_p(o) for o in out], _p(img), ctypes.c_int(nb), ctypes.c_int(nx), ctypes.c_int(ny),
                             ctypes.c_int(nz), ctypes.c_int(nc), ctypes.c_int(dims), ctypes.c_double(eps))
    return tuple(out)


def dd_cone_fp(img, det_center, det_u, det_v, src, geo, nu, nv, precision="f64"):
    return _arb_fp("dd_cone_fp", img, [det_center, det_u, det_v, src], geo, nu, nv, precision)


def dd_cone_adj(prj, det_center, det_u, det_v, src, geo, img_shape, precision="f64"):
    return _arb_bp("dd_cone_adj", prj, [det_center, det_u, det_v, src], geo, img_shape, precision)


def dd_tomo_fp(img, det_center, src, geo, nu, nv, precision="f64"):
    return _arb_fp("dd_tomo_fp", img, [det_center, src], geo, nu, nv, precision)


def dd_tomo_bp(prj, det_center, src, geo, img_shape, precision="f64"):
    return _arb_bp("dd_tomo_bp", prj, [det_center, src], geo, img_shape, precision)


def os_sqs_cone(img, prj, norm_img, mask, det_center, det_u, det_v, src, geo, n_subsets,
                lb=None, ub=None, lam=1.0):
    img = _f32(img); prj = _f32(prj); norm_img = _f32(norm_img); mask = _f32(mask)
    vecs = [_f32(v) for v in (det_center, det_u, det_v, src)]
    nB, nx, ny, nz, nC = img.shape
    _, nu, nview, nv, _ = prj.shape
    if lb is None:
        lb = np.finfo(np.float32).min
    if ub is None:
        ub = np.finfo(np.float32).max
    npers = int(np.ceil(nview / float(n_subsets)))
    out = np.zeros(img.shape, np.float64)
    lib().orc_os_sqs_cone_f64(_p(out), _p(img), _p(prj), _p(norm_img), _p(mask), *[_p(v) for v in vecs],
                              ctypes.c_float(lb), ctypes.c_float(ub), ctypes.c_int(npers), ctypes.c_int(n_subsets),
                              *_arb_tail(nB, nC, nx, ny, nz, geo, nu, nview, nv), ctypes.c_float(lam))
    return out


def helical_bp(prj, geo, img_shape, nview_per_pi, theta0, vol_z0, zrot, m_pi=1, Q=0.5):
    prj = _f32(prj)
    nB, nu, nview, nv, nC = prj.shape
    nx, ny, nz = img_shape
    out = np.zeros((nB, nx, ny, nz, nC), np.float64)
    lib().orc_helical_bp_f64(
        _p(out), _p(prj), ctypes.c_int(nB), ctypes.c_int(nC), ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz),
        ctypes.c_float(geo["dx"]), ctypes.c_float(geo["dy"]), ctypes.c_float(geo["dz"]),
        ctypes.c_int(nu), ctypes.c_int(nview), ctypes.c_int(nv), ctypes.c_float(geo["du"]), ctypes.c_float(geo["dv"]),
        ctypes.c_float(geo.get("off_u", 0.0)), ctypes.c_float(geo.get("off_v", 0.0)),
        ctypes.c_float(geo["dsd"]), ctypes.c_float(geo["dso"]), ctypes.c_int(nview_per_pi),
        ctypes.c_float(theta0), ctypes.c_float(vol_z0), ctypes.c_float(zrot), ctypes.c_int(m_pi), ctypes.c_float(Q))
    return out
