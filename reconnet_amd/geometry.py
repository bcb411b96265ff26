"""Fan-beam geometry records and the workloads used by the tests and benchmark.

Conventions follow the reference projector (DistanceDriven.cu:14-20,81-99;
Siddon.cu:253-263): source at (dso sin t, -dso cos t), equiangular detector arc
of nu bins, bin angle a = (-(iu - (nu-1)/2) - off_a) da, da = du / dsd
(ReconNet_no_tf.py:44,69), angles t_k = float32(k * 2 pi / N)
(ReconNet_no_tf.py:72-73).
"""

from dataclasses import asdict, dataclass

import numpy as np

__all__ = ["FanGeometry", "fan_angles", "WORKLOADS", "interleaved_order", "HU_WATER", "circular_cone_vectors",
           "CONE_WORKLOAD", "uneven_fan_batch"]

# mu of water used by the reference's HU <-> mu map (ProjectorTester.cu:27-28)
HU_WATER = 0.019


@dataclass
class FanGeometry:
    nx: int
    ny: int
    dx: float
    dy: float
    nu: int
    nview: int
    da: float
    dsd: float
    dso: float
    off_a: float = 0.0
    nz: int = 1
    dz: float = 1.0
    nv: int = 1
    dv: float = 1.0
    off_v: float = 0.0

    def as_dict(self):
        return asdict(self)

    def oracle_geo(self):
        return dict(dx=self.dx, dy=self.dy, dz=self.dz, da=self.da, dv=self.dv, off_a=self.off_a,
                    off_v=self.off_v, dsd=self.dsd, dso=self.dso)

    def angles(self):
        return fan_angles(self.nview)

    def fov_radius(self):
        """Radius of the fully sampled field of view (mm)."""
        r_det = self.dso * np.sin(self.nu * self.da / 2)
        return float(min(r_det, 0.5 * self.nx * self.dx, 0.5 * self.ny * self.dy))

    def fov_mask(self):
        x = (np.arange(self.nx) - (self.nx - 1) / 2) * self.dx
        y = (np.arange(self.ny) - (self.ny - 1) / 2) * self.dy
        X, Y = np.meshgrid(x, y, indexing="ij")
        return (X ** 2 + Y ** 2 <= self.fov_radius() ** 2)


def fan_angles(nview):
    """float32(k * 2 pi / N), the reference's angle recipe."""
    return (np.arange(0, nview, dtype=np.float32) * 2 * np.pi / nview).astype(np.float32)


def interleaved_order(nview, nsubsets):
    """View permutation that makes contiguous blocks of ceil(nview/M) views
    interleaved subsets {k, k+M, k+2M, ...} (the `order` argument of the
    reference's OS-SQS wrapper, ReconNet_no_tf.py:419-422)."""
    return np.concatenate([np.arange(k, nview, nsubsets) for k in range(nsubsets)])


WORKLOADS = {
    # configs of the benchmark definition: plain FP/BP/FBP at 256^2
    "fan256": FanGeometry(nx=256, ny=256, dx=1.0, dy=1.0, nu=512, nview=360, da=1.2 / 1085.6,
                          dsd=1085.6, dso=595.0),
    # OS-SQS-TV and learned unrolled slices at 512^2
    "fan512": FanGeometry(nx=512, ny=512, dx=0.7, dy=0.7, nu=736, nview=1152, da=1.0 / 1085.6,
                          dsd=1085.6, dso=595.0),
}


def circular_cone_vectors(nview, dso=595.0, dsd=1085.6, th0=0.0):
    """Per-view (detector centre, detector u axis, detector v axis, source),
    float32 [nview, 3] each, of a circular cone-beam scan whose source follows
    the fan convention (dso sin t, -dso cos t, 0) with t the float32 angle
    recipe (+ th0), a flat detector at distance dsd from the source, columns
    along (cos t, sin t, 0) and rows along z (the argument order of
    cSiddonConeProjectionAbitrary, SiddonCone.cu:117)."""
    th = fan_angles(nview).astype(np.float64) + th0
    src = np.stack([dso * np.sin(th), -dso * np.cos(th), np.zeros_like(th)], 1)
    ctr = np.stack([-(dsd - dso) * np.sin(th), (dsd - dso) * np.cos(th), np.zeros_like(th)], 1)
    du = np.stack([np.cos(th), np.sin(th), np.zeros_like(th)], 1)
    dv = np.tile([0.0, 0.0, 1.0], (nview, 1))
    return tuple(a.astype(np.float32) for a in (ctr, du, dv, src))


# 3D cone-beam workload: 256^2 x 128 voxels of 1 mm, 64 rows x 512 columns of
# 1.0 x 1.2 mm, 720 views over 2 pi, SID 595 mm, SDD 1085.6 mm
CONE_WORKLOAD = dict(nx=256, ny=256, nz=128, dx=1.0, dy=1.0, dz=1.0, nu=512, nv=64, du=1.2, dv=1.0, nview=720,
                     dso=595.0, dsd=1085.6)


def uneven_fan_batch():
    """The 16-slice uneven batch: six 256^2 slices with 360 views, six 384^2
    with 720 and four 512^2 with 1152, 736 bins of 1.0 mm, each slice covering
    about the same 358 mm field."""
    geos = []
    for n, nx, nview, dx in [(6, 256, 360, 1.4), (6, 384, 720, 0.93), (4, 512, 1152, 0.7)]:
        geos += [FanGeometry(nx=nx, ny=nx, dx=dx, dy=dx, nu=736, nview=nview, da=1.0 / 1085.6, dsd=1085.6,
                             dso=595.0)] * n
    return geos
