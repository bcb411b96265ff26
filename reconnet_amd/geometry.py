"""Fan-beam geometry records and the workloads used by the tests and benchmark.

Conventions follow the reference projector (DistanceDriven.cu:14-20,81-99;
Siddon.cu:253-263): source at (dso sin t, -dso cos t), equiangular detector arc
of nu bins, bin angle a = (-(iu - (nu-1)/2) - off_a) da, da = du / dsd
(ReconNet_no_tf.py:44,69), angles t_k = float32(k * 2 pi / N)
(ReconNet_no_tf.py:72-73).
"""

from dataclasses import asdict, dataclass

import numpy as np

__all__ = ["FanGeometry", "fan_angles", "WORKLOADS", "interleaved_order", "HU_WATER"]

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
