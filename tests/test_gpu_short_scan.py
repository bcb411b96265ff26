"""Short-scan FBP on the GPU: Parker / Riess redundancy weights
(ReconNet_no_tf.py:640-693, host numpy as in the reference) -> the GPU ramp
filter (cFilterFanFilter) -> the GPU FBP-weighted DD backprojection
(cDistanceDrivenFanBackprojection, typeProjector = 1), each stage against the
FP64 oracle on the same input, and the reconstruction against the phantom."""

import numpy as np
import pytest

import oracle as orc
from reconnet_amd.geometry import FanGeometry
from reconnet_amd.phantom import phantom
from reconnet_amd.recon import ParkerWeighting, RiessWeighting
from tests.helpers import recon_for, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("kind", ["parker", "riess"])
def test_short_scan_fbp_chain_matches_oracle(kind):
    # 256^2 at 1 mm, 512 bins of 1.2 mm: the fan spans +-0.283 rad; a short
    # scan of pi + 2 gamma with gamma a little wider, at the 1-degree step
    geo0 = FanGeometry(nx=256, ny=256, dx=1.0, dy=1.0, nu=512, nview=214, da=1.2 / 1085.6, dsd=1085.6,
                       dso=595.0)
    deg = (np.arange(geo0.nview, dtype=np.float32) * np.float32(2 * np.pi / 360)).astype(np.float32)
    r = recon_for(geo0)
    r.angles = list(deg)
    g5 = geo0.oracle_geo()
    img = phantom(3, 256, 256, 1.0, 1.0)
    sino = orc.dd_fan_fp(img[None, :, :, None, None], deg, g5, geo0.nu).astype(np.float32)
    w = ParkerWeighting(r) if kind == "parker" else RiessWeighting(r)
    weighted = (sino * w[..., None]).astype(np.float32)
    filt = r.cFilter4d(weighted, type_filter=2)
    assert rel_l2(filt, orc.ramp_filter(weighted, g5, filter_type=2)) <= TOL
    rec = r.cDDFanBackprojection4d(filt, angles=deg, type_projector=1)
    ref_bp = orc.dd_fan_bp(filt, deg, g5, (256, 256, 1), type_projector=1)
    assert rel_l2(rec, ref_bp) <= TOL
    # end to end against the oracle chain
    ref = orc.dd_fan_bp(orc.ramp_filter(weighted, g5, filter_type=2).astype(np.float32), deg, g5, (256, 256, 1),
                        type_projector=1)
    assert rel_l2(rec, ref) <= 1e-4
    # the short-scan reconstruction reproduces the phantom inside the field of
    # view up to the global scale of a half-scan filter normalisation (pi/nview
    # with nview the short scan's view count)
    m = geo0.fov_mask()

    def fit_err(x):
        a, b = x[0, :, :, 0, 0][m].astype(np.float64), img[m].astype(np.float64)
        return rel_l2(float(a @ b / (a @ a)) * a, b)

    assert fit_err(rec) < 0.08
    # without the redundancy weights the doubly measured rays show (~15 %)
    unweighted = r.cDDFanBackprojection4d(r.cFilter4d(sino, type_filter=2), angles=deg, type_projector=1)
    assert fit_err(unweighted) > 0.1
