"""Failure handling at the C-ABI: a failing call never aborts the process or
throws across extern "C"; it prints "<function> failed: <reason>", leaves the
message in rn_last_error() and the Python mirror raises.  On a machine without
a GPU every call fails that way (CPU test); on the GPU the failure is forced
from the host with the allocation fault hook (RECONNET_FAULT=alloc), after
which the library keeps working."""

import os

import numpy as np
import pytest
import torch


def _recon(n=16, nu=24, nview=8):
    from reconnet_amd.recon import ReconNet

    r = ReconNet()
    r.nx = r.ny = n
    r.nu, r.rotview = nu, nview
    r.UpdateAngularSampling()
    return r


@pytest.mark.skipif(torch.cuda.is_available(), reason="needs a machine without a GPU")
def test_calls_fail_cleanly_without_a_device():
    from reconnet_amd._lib import last_error

    r = _recon()
    with pytest.raises(RuntimeError, match="cDistanceDrivenFanProjection failed"):
        r.cDDFanProjection4d(np.ones((1, 16, 16, 1, 1), np.float32))
    assert "cDistanceDrivenFanProjection failed" in last_error()
    with pytest.raises(RuntimeError, match="failed"):
        r.cSiddonFanBackprojection4d(np.ones((1, 24, 8, 1, 1), np.float32))


@pytest.mark.gpu
def test_allocation_failure_is_reported_and_recovered():
    from reconnet_amd._lib import last_error
    from reconnet_amd.torch_ops import FanPlan
    from tests.helpers import small_fan

    geo = small_fan(nx=40, nu=70, nview=30)
    os.environ["RECONNET_FAULT"] = "alloc"
    try:
        with pytest.raises(RuntimeError, match="fault injection"):
            FanPlan(geo)
        r = _recon(n=44, nu=66, nview=20)        # sizes no earlier call cached
        with pytest.raises(RuntimeError, match="fault injection"):
            r.cDDFanProjection4d(np.ones((1, 44, 44, 1, 1), np.float32))
        assert "fault injection" in last_error()
    finally:
        del os.environ["RECONNET_FAULT"]
    plan = FanPlan(geo)
    y = plan.forward(torch.ones(1, 40, 40, device="cuda"), "siddon")
    assert torch.isfinite(y).all() and float(y.max()) > 0
    assert last_error() == ""
