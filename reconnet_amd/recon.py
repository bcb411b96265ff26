"""Host-array front end over the C-ABI, mirroring python/ReconNet_no_tf.py.

``ReconNet`` keeps the reference class's attribute names, defaults
(ReconNet_no_tf.py:27-47), parameter-file semantics (:49-70), angle recipe
(:72-73) and the call shapes of its 4d/3d wrappers (:84-628): numpy arrays in
the reference layouts ([B, nx, ny, nz, C] images, [B, nu, nview, nv, C]
sinograms), outputs freshly allocated, every scalar passed explicitly.  The
functions it calls are the ones exported by this package's
libReconNet4D_no_tf.so, which run on the GPU.
"""

import re
from ctypes import POINTER, c_float, c_int

import numpy as np

from ._lib import last_error, lib

__all__ = ["ReconNet", "ParkerWeighting", "RiessWeighting"]


def _fp(a):
    return a.ctypes.data_as(POINTER(c_float))


def _raise_if_failed(name):
    msg = last_error()
    if msg:
        raise RuntimeError(msg if msg.startswith(name) else "%s failed: %s" % (name, msg))


class ReconNet:
    """Fan / cone geometry plus wrappers for every reference export."""

    def __init__(self):
        self.rotview = 720
        self.nu = 736
        self.nv = 1
        self.nx = 512
        self.ny = 512
        self.nz = 1
        self.dx = 1
        self.dy = 1
        self.dz = 2
        self.cx = 0
        self.cy = 0
        self.cz = 0
        self.dsd = 1085.6
        self.dso = 595
        self.du = 1.2858
        self.dv = 2
        self.da = self.du / self.dsd
        self.off_a = 0
        self.off_u = 0
        self.off_v = 0
        self.UpdateAngularSampling()

    # ----- parameters -----
    def FromFile(self, filename):
        """Read "name value" lines (ReconNet_no_tf.py:49-70): names are
        lower-cased, values float32; integer fields are cast back, off_a is
        negated, da = du / dsd and the angles are regenerated."""
        params = {}
        with open(filename, "r") as f:
            for line in f:
                fields = re.split(" |\t|\n", line)
                params[fields[0].lower()] = np.float32(fields[1])
        for k, v in params.items():
            setattr(self, k, v)
        for k in ("rotview", "nu", "nv", "nx", "ny", "nz"):
            setattr(self, k, int(getattr(self, k)))
        self.off_a = -self.off_a
        self.da = self.du / self.dsd
        self.UpdateAngularSampling()

    def UpdateAngularSampling(self):
        self.angles = (np.arange(0, self.rotview, dtype=np.float32) * 2 * np.pi / self.rotview).tolist()

    def get_det_sz(self, type_projector=0):
        return [self.da, self.dv]

    def get_det_off(self, type_projector=0):
        return [self.off_a, self.off_v]

    def get_img_shape(self):
        return [self.nx, self.ny, self.nz]

    def get_sino_shape(self):
        return [self.nu, self.rotview, self.nv]

    def cSetDevice(self, device):
        return lib().SetDevice(c_int(device))

    # ----- helpers -----
    def _angles(self, angles):
        if angles is None:
            return np.array(self.angles, np.float32)
        return np.ascontiguousarray(angles, dtype=np.float32)

    def _fan_tail(self, img_shape5, sino_shape5, type_projector):
        return (c_int(img_shape5[0]), c_int(img_shape5[4]), c_int(img_shape5[1]), c_int(img_shape5[2]),
                c_int(img_shape5[3]), c_float(self.dx), c_float(self.dy), c_float(self.dz),
                c_int(sino_shape5[1]), c_int(sino_shape5[2]), c_int(sino_shape5[3]),
                c_float(self.da), c_float(self.dv), c_float(self.off_a), c_float(self.off_v),
                c_float(self.dsd), c_float(self.dso), c_int(type_projector))

    def _arb_tail(self, img_shape5, prj_shape5):
        return (c_int(img_shape5[0]), c_int(img_shape5[-1]), c_int(img_shape5[1]), c_int(img_shape5[2]),
                c_int(img_shape5[3]), c_float(self.dx), c_float(self.dy), c_float(self.dz),
                c_float(self.cx), c_float(self.cy), c_float(self.cz),
                c_int(prj_shape5[1]), c_int(prj_shape5[2]), c_int(prj_shape5[3]),
                c_float(self.du), c_float(self.dv), c_float(self.off_u), c_float(self.off_v))

    def _fan_fp(self, name, img, angles, sino_shape, type_projector):
        img = np.ascontiguousarray(img, dtype=np.float32)
        angles = self._angles(angles)
        if sino_shape is None:
            sino_shape = [self.nu, len(angles), self.nv]
        sino = np.zeros([img.shape[0], sino_shape[0], sino_shape[1], sino_shape[2], img.shape[4]], np.float32)
        getattr(lib(), name)(_fp(sino), _fp(img), _fp(angles), *self._fan_tail(img.shape, sino.shape, type_projector))
        _raise_if_failed(name)
        return sino

    def _fan_bp(self, name, sino, angles, img_shape, type_projector):
        sino = np.ascontiguousarray(sino, dtype=np.float32)
        angles = self._angles(angles)
        if img_shape is None:
            img_shape = [self.nx, self.ny, self.nz]
        img = np.zeros([sino.shape[0], img_shape[0], img_shape[1], img_shape[2], sino.shape[4]], np.float32)
        getattr(lib(), name)(_fp(img), _fp(sino), _fp(angles), *self._fan_tail(img.shape, sino.shape, type_projector))
        _raise_if_failed(name)
        return img

    # ----- filter (Filter.cu:432) -----
    def cFilter4d(self, sino, type_filter=0, type_projector=0):
        sino = np.ascontiguousarray(sino, dtype=np.float32)
        fsino = np.zeros(sino.shape, np.float32)
        lib().cFilterFanFilter(_fp(fsino), _fp(sino), c_int(sino.shape[0]), c_int(sino.shape[4]),
                               c_int(sino.shape[1]), c_int(sino.shape[2]), c_int(sino.shape[3]),
                               c_float(self.da), c_float(self.dv), c_float(self.off_a), c_float(self.off_v),
                               c_float(self.dsd), c_float(self.dso), c_int(type_filter), c_int(type_projector))
        _raise_if_failed("cFilterFanFilter")
        return fsino

    def cFilter3d(self, sino, type_filter=0, type_projector=0):
        return self.cFilter4d(sino[..., np.newaxis, :], type_filter, type_projector)[..., 0, :]

    def cFilterParallel4d(self, sino, type_filter=0, type_projector=0):
        sino = np.ascontiguousarray(sino, dtype=np.float32)
        fsino = np.zeros(sino.shape, np.float32)
        lib().cFilterParallelFilter(_fp(fsino), _fp(sino), c_int(sino.shape[0]), c_int(sino.shape[4]),
                                    c_int(sino.shape[1]), c_int(sino.shape[2]), c_int(sino.shape[3]),
                                    c_float(self.du), c_float(self.dv), c_float(self.off_u), c_float(self.off_v),
                                    c_float(self.dsd), c_float(self.dso), c_int(type_filter), c_int(type_projector))
        _raise_if_failed("cFilterParallelFilter")
        return fsino

    # ----- pixel-driven BP (PixelDriven.cu:100) -----
    def cPixelBackprojection4d(self, sino, angles=None, img_shape=None, type_projector=0):
        return self._fan_bp("cPixelDrivenFanBackprojection", sino, angles, img_shape, type_projector)

    def cPixelBackprojection3d(self, sino, angles=None, img_shape=None, type_projector=0):
        if img_shape is None:
            img_shape = [self.nx, self.ny, 1]
        img = self.cPixelBackprojection4d(sino[..., np.newaxis, :], angles, img_shape, type_projector)
        return np.squeeze(img, -2)

    # ----- distance driven fan (DistanceDriven.cu:386,793) -----
    def cDDFanProjection4d(self, img, angles=None, sino_shape=None, type_projector=0):
        return self._fan_fp("cDistanceDrivenFanProjection", img, angles, sino_shape, type_projector)

    def cDDFanProjection3d(self, img, angles=None, sino_shape=None, type_projector=0):
        angles = self._angles(angles)
        if sino_shape is None:
            sino_shape = [self.nu, len(angles), 1]
        return np.squeeze(self.cDDFanProjection4d(img[..., np.newaxis, :], angles, sino_shape, type_projector), -2)

    def cDDFanBackprojection4d(self, sino, angles=None, img_shape=None, type_projector=0):
        return self._fan_bp("cDistanceDrivenFanBackprojection", sino, angles, img_shape, type_projector)

    def cDDFanBackprojection3d(self, sino, angles=None, img_shape=None, type_projector=0):
        if img_shape is None:
            img_shape = [self.nx, self.ny, 1]
        return self.cDDFanBackprojection4d(sino[..., np.newaxis, :], angles, img_shape, type_projector)[..., 0, :]

    def cDDFanNormImg4d(self, sino, wsino=None, angles=None, img_shape=None, type_projector=0):
        angles = self._angles(angles)
        wsino = np.ones(sino.shape, np.float32) if wsino is None else np.asarray(wsino, np.float32)
        if img_shape is None:
            img_shape = [self.nx, self.ny, self.nz]
        ones = np.ones([sino.shape[0], img_shape[0], img_shape[1], img_shape[2], sino.shape[4]], np.float32)
        fp = self.cDDFanProjection4d(ones, angles, sino.shape[1:4], type_projector)
        return self.cDDFanBackprojection4d(fp * wsino, angles, img_shape, type_projector)

    def cDDFanNormImg3d(self, sino, wsino=None, angles=None, img_shape=None, type_projector=0):
        wsino = np.ones(sino.shape, np.float32) if wsino is None else np.asarray(wsino, np.float32)
        if img_shape is None:
            img_shape = [self.nx, self.ny, 1]
        n = self.cDDFanNormImg4d(sino[..., np.newaxis, :], wsino[..., np.newaxis, :], angles, img_shape,
                                 type_projector)
        return n[..., 0, :]

    # ----- DD tomosynthesis (DistanceDriven3D.cu:346,666) -----
    def cDDTomoProjection4d(self, img, detCenter, src):
        img = np.ascontiguousarray(img, dtype=np.float32)
        detCenter = np.ascontiguousarray(detCenter, dtype=np.float32)
        src = np.ascontiguousarray(src, dtype=np.float32)
        prj = np.zeros([img.shape[0], self.nu, detCenter.shape[0], self.nv, img.shape[-1]], np.float32)
        lib().cDistanceDrivenTomoProjection(_fp(prj), _fp(img), _fp(detCenter), _fp(src),
                                            *self._arb_tail(img.shape, prj.shape))
        _raise_if_failed("cDistanceDrivenTomoProjection")
        return prj

    def cDDTomoBackprojection4d(self, prj, detCenter, src):
        prj = np.ascontiguousarray(prj, dtype=np.float32)
        detCenter = np.ascontiguousarray(detCenter, dtype=np.float32)
        src = np.ascontiguousarray(src, dtype=np.float32)
        img = np.zeros([prj.shape[0], self.nx, self.ny, self.nz, prj.shape[-1]], np.float32)
        lib().cDistanceDrivenTomoBackprojection(_fp(img), _fp(prj), _fp(detCenter), _fp(src),
                                                *self._arb_tail(img.shape, prj.shape))
        _raise_if_failed("cDistanceDrivenTomoBackprojection")
        return img

    def cDDTomoNormImg4d(self, prj, detCenter, src, mask=None):
        img = np.ones([prj.shape[0], self.nx, self.ny, self.nz, prj.shape[-1]], np.float32)
        if mask is None:
            mask = np.ones_like(img)
        fp = self.cDDTomoProjection4d(img * mask, detCenter, src)
        return self.cDDTomoBackprojection4d(fp, detCenter, src) * mask

    # ----- Siddon fan (Siddon.cu:424,481) -----
    def cSiddonFanProjection4d(self, img, angles=None, sino_shape=None, type_projector=0):
        return self._fan_fp("cSiddonFanProjection", img, angles, sino_shape, type_projector)

    def cSiddonFanProjection3d(self, img, angles=None, sino_shape=None, type_projector=0):
        angles = self._angles(angles)
        if sino_shape is None:
            sino_shape = [self.nu, len(angles), 1]
        return np.squeeze(self.cSiddonFanProjection4d(img[..., np.newaxis, :], angles, sino_shape, type_projector),
                          -2)

    def cSiddonFanBackprojection4d(self, sino, angles=None, img_shape=None, type_projector=0):
        return self._fan_bp("cSiddonFanBackprojection", sino, angles, img_shape, type_projector)

    def cSiddonFanBackprojection3d(self, sino, angles=None, img_shape=None, type_projector=0):
        if img_shape is None:
            img_shape = [self.nx, self.ny, 1]
        return self.cSiddonFanBackprojection4d(sino[..., np.newaxis, :], angles, img_shape, type_projector)[..., 0, :]

    def cSiddonFanNormImg4d(self, sino, wsino=None, angles=None, img_shape=None, type_projector=0):
        angles = self._angles(angles)
        wsino = np.ones(sino.shape, np.float32) if wsino is None else np.asarray(wsino, np.float32)
        if img_shape is None:
            img_shape = [self.nx, self.ny, self.nz]
        ones = np.ones([sino.shape[0], img_shape[0], img_shape[1], img_shape[2], sino.shape[4]], np.float32)
        fp = self.cSiddonFanProjection4d(ones, angles, sino.shape[1:4], type_projector)
        return self.cSiddonFanBackprojection4d(fp * wsino, angles, img_shape, type_projector)

    def cSiddonFanNormImg3d(self, sino, wsino=None, angles=None, img_shape=None, type_projector=0):
        wsino = np.ones(sino.shape, np.float32) if wsino is None else np.asarray(wsino, np.float32)
        if img_shape is None:
            img_shape = [self.nx, self.ny, 1]
        n = self.cSiddonFanNormImg4d(sino[..., np.newaxis, :], wsino[..., np.newaxis, :], angles, img_shape,
                                     type_projector)
        return n[..., 0, :]

    # ----- OS-SQS (ossqs.cu:12) -----
    def cSiddonFanOSSQS4d(self, nSubsets, img, sino, normImg, wsino=None, angles=None, order=None,
                          lb=None, ub=None, type_projector=0, lam=1):
        img = np.ascontiguousarray(img, dtype=np.float32)
        sino = np.ascontiguousarray(sino, dtype=np.float32)
        normImg = np.ascontiguousarray(normImg, dtype=np.float32)
        lb = np.finfo(np.float32).min if lb is None else lb
        ub = np.finfo(np.float32).max if ub is None else ub
        wsino = np.ones(sino.shape, np.float32) if wsino is None else np.ascontiguousarray(wsino, np.float32)
        angles = self._angles(angles)
        res = np.zeros(img.shape, np.float32)
        nviewPerSubset = int(np.ceil(sino.shape[2] / float(nSubsets)))
        if order is not None:
            angles = np.ascontiguousarray(angles[order])
            sino = np.copy(sino[..., order, :, :], "C")
            wsino = np.copy(wsino[..., order, :, :], "C")
        lib().cOSSQS(_fp(res), _fp(img), _fp(sino), _fp(normImg), _fp(wsino), _fp(angles),
                     c_float(lb), c_float(ub), c_int(nviewPerSubset), c_int(nSubsets),
                     *self._fan_tail(img.shape, sino.shape, type_projector)[:-1],
                     c_int(type_projector), c_float(lam))
        _raise_if_failed("cOSSQS")
        return res

    def cSiddonFanOSSQS3d(self, nSubsets, img, sino, normImg, wsino=None, angles=None, order=None,
                          lb=None, ub=None, type_projector=0, lam=1):
        wsino = np.ones(sino.shape, np.float32) if wsino is None else np.asarray(wsino, np.float32)
        res = self.cSiddonFanOSSQS4d(nSubsets, img[..., np.newaxis, :], sino[..., np.newaxis, :],
                                     normImg[..., np.newaxis, :], wsino[..., np.newaxis, :], angles, order,
                                     lb, ub, type_projector, lam)
        return res[..., 0, :]

    # ----- Siddon parallel / cone (SiddonParallel.cu:152,230; SiddonCone.cu:117,196) -----
    def _arb_fp(self, name, img, vecs, nview):
        img = np.ascontiguousarray(img, dtype=np.float32)
        vecs = [np.ascontiguousarray(v, dtype=np.float32) for v in vecs]
        prj = np.zeros([img.shape[0], self.nu, nview, self.nv, img.shape[-1]], np.float32)
        getattr(lib(), name)(_fp(prj), _fp(img), *[_fp(v) for v in vecs], *self._arb_tail(img.shape, prj.shape))
        _raise_if_failed(name)
        return prj

    def _arb_bp(self, name, prj, vecs):
        prj = np.ascontiguousarray(prj, dtype=np.float32)
        vecs = [np.ascontiguousarray(v, dtype=np.float32) for v in vecs]
        img = np.zeros([prj.shape[0], self.nx, self.ny, self.nz, prj.shape[-1]], np.float32)
        getattr(lib(), name)(_fp(img), _fp(prj), *[_fp(v) for v in vecs], *self._arb_tail(img.shape, prj.shape))
        _raise_if_failed(name)
        return img

    def cSiddonParallelProjection4d(self, img, detCenter, detU, detV, invRayDir):
        # the reference sizes the view axis by rotview (ReconNet_no_tf.py:462)
        return self._arb_fp("cSiddonParallelProjection", img, [detCenter, detU, detV, invRayDir], self.rotview)

    def cSiddonParallelBackprojection4d(self, prj, detCenter, detU, detV, invRayDir):
        return self._arb_bp("cSiddonParallelBackprojection", prj, [detCenter, detU, detV, invRayDir])

    def cSiddonConeProjectionAbitrary4d(self, img, detCenter, detU, detV, src):
        return self._arb_fp("cSiddonConeProjectionAbitrary", img, [detCenter, detU, detV, src],
                            np.asarray(detCenter).shape[0])

    def cSiddonConeBackprojectionAbitrary4d(self, prj, detCenter, detU, detV, src):
        return self._arb_bp("cSiddonConeBackprojectionAbitrary", prj, [detCenter, detU, detV, src])

    def cSiddonConeNormImgAbitrary4d(self, prj, detCenter, detU, detV, src, mask=None):
        img = np.ones([prj.shape[0], self.nx, self.ny, self.nz, prj.shape[-1]], np.float32)
        if mask is None:
            mask = np.ones_like(img)
        fp = self.cSiddonConeProjectionAbitrary4d(img * mask, detCenter, detU, detV, src)
        return self.cSiddonConeBackprojectionAbitrary4d(fp, detCenter, detU, detV, src) * mask

    def cSiddonConeAbitraryOSSQS4d(self, nSubsets, img, prj, normImg, detCenter, detU, detV, src, mask=None,
                                   order=None, lb=None, ub=None, lam=1):
        img = np.ascontiguousarray(img, dtype=np.float32)
        prj = np.ascontiguousarray(prj, dtype=np.float32)
        normImg = np.ascontiguousarray(normImg, dtype=np.float32)
        vecs = [np.ascontiguousarray(v, dtype=np.float32) for v in (detCenter, detU, detV, src)]
        mask = np.ones(img.shape, np.float32) if mask is None else np.ascontiguousarray(mask, np.float32)
        nviewPerSubset = int(np.ceil(prj.shape[2] / float(nSubsets)))
        lb = np.finfo(np.float32).min if lb is None else lb
        ub = np.finfo(np.float32).max if ub is None else ub
        if order is not None:
            prj = np.copy(prj[:, :, order, :, :], "C")
            vecs = [np.copy(v[order, ...], "C") for v in vecs]
        res = np.zeros(img.shape, np.float32)
        lib().cOSSQSConeAbitrary(_fp(res), _fp(img), _fp(prj), _fp(normImg), _fp(mask), *[_fp(v) for v in vecs],
                                 c_float(lb), c_float(ub), c_int(nviewPerSubset), c_int(nSubsets),
                                 *self._arb_tail(img.shape, prj.shape), c_int(2), c_float(lam))
        _raise_if_failed("cOSSQSConeAbitrary")
        return res

    # ----- TV (TotalVariation.cu:70) -----
    def cTVSQS2d(self, img):
        img = np.ascontiguousarray(img, dtype=np.float32)
        if img.ndim != 4 or img.shape[0] != 1 or img.shape[-1] != 1:
            raise ValueError("img.shape must be [1, _, _, 1]")
        grad = np.zeros(img.shape, np.float32)
        norm = np.zeros(img.shape, np.float32)
        tv = lib().cTVSQS2D(_fp(grad), _fp(norm), _fp(img), c_int(img.shape[1]), c_int(img.shape[2]))
        _raise_if_failed("cTVSQS2D")
        return grad, norm, tv

    # ----- helical (Recon3D/HelicalBP.cu:132) -----
    def cHelicalBP4d(self, prj, nviewPerPI, theta0, volZ0, zrot, mPI=1, Q=0.5):
        prj = np.ascontiguousarray(prj, dtype=np.float32)
        img = np.zeros([prj.shape[0], self.nx, self.ny, self.nz, prj.shape[-1]], np.float32)
        lib().cHelicalBPFromParallelConebeamBackprojection(
            _fp(img), _fp(prj), c_int(prj.shape[0]), c_int(prj.shape[-1]), c_int(self.nx), c_int(self.ny),
            c_int(self.nz), c_float(self.dx), c_float(self.dy), c_float(self.dz), c_int(prj.shape[1]),
            c_int(prj.shape[2]), c_int(prj.shape[3]), c_float(self.du), c_float(self.dv), c_float(self.off_u),
            c_float(self.off_v), c_float(self.dsd), c_float(self.dso), c_int(nviewPerPI), c_float(theta0),
            c_float(volZ0), c_float(zrot), c_int(mPI), c_float(Q))
        _raise_if_failed("cHelicalBPFromParallelConebeamBackprojection")
        return img


def _short_scan_setup(recon):
    angles = np.asarray(recon.angles, dtype=np.float64)
    gamma = (angles[-1] - np.pi) / 2
    alphas = ((np.arange(recon.nu) - (recon.nu - 1) / 2) - recon.off_a) * recon.da
    return angles, gamma, alphas


def ParkerWeighting(recon):
    """Parker short-scan weights [1, nu, nview, 1] (ReconNet_no_tf.py:640-659)."""
    angles, gamma, alphas = _short_scan_setup(recon)
    cols = []
    for beta in recon.angles:
        rise = np.sin(np.pi / 4 * beta / (gamma - alphas)) ** 2
        fall = np.sin(np.pi / 4 * (np.pi + 2 * gamma - beta) / (gamma + alphas)) ** 2
        w = fall.copy()
        w[beta <= np.pi - 2 * alphas] = 1
        first = beta < 2 * gamma - 2 * alphas
        w[first] = rise[first]
        cols.append(w)
    return np.transpose(np.array(cols))[np.newaxis, ..., np.newaxis]


def RiessWeighting(recon, sigma=30, theta=10 * np.pi / 180):
    """Riess short-scan weights with Gaussian smoothing (ReconNet_no_tf.py:665-693)."""
    import scipy.ndimage

    angles, gamma, alphas = _short_scan_setup(recon)
    cols = []
    for beta in recon.angles:
        rise = np.sin(np.pi / 4 * beta / (gamma - alphas)) ** 2
        fall = np.sin(np.pi / 4 * (np.pi + 2 * gamma - beta) / (gamma + alphas)) ** 2
        w = np.ones([recon.nu], np.float32)
        sel4 = beta > np.pi + 2 * gamma + 2 * alphas
        sel3 = beta > np.pi - 2 * alphas
        sel2 = beta < -2 * gamma + 2 * alphas
        sel1 = beta < 2 * gamma - 2 * alphas
        w[sel4] = 2 - fall[sel4]
        w[sel3] = fall[sel3]
        w[sel2] = 2 - rise[sel2]
        w[sel1] = rise[sel1]
        smooth = scipy.ndimage.gaussian_filter1d(w, sigma, mode="nearest")
        # ReconNet_no_tf.py:687-688 indexes with np.where of a 0-d boolean; under
        # the numpy 1.x semantics it was written for that selects element 0 only
        # (numpy 2 rejects the expression).
        if beta > theta and beta < np.pi + 2 * gamma - theta:
            smooth[0] = w[0]
        cols.append(smooth)
    return np.transpose(np.array(cols))[np.newaxis, ..., np.newaxis]
