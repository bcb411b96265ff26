"""The launch-shape autotuner: tuned launches give the same bits as the
default shapes, and every tuned key records a timing per candidate."""

import numpy as np
import pytest
import torch

from tests.helpers import small_fan

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fp,bp", [("dd", "dd_adjoint"), ("siddon", "siddon"), ("dd", "dd"), ("dd", "dd_fbp")])
def test_tuned_shapes_give_identical_results(fp, bp):
    from reconnet_amd._lib import autotune
    from reconnet_amd.torch_ops import FanPlan

    geo = small_fan(nx=200, nu=300, nview=120, dx=1.1)
    plan = FanPlan(geo)
    rng = np.random.default_rng(7)
    x = torch.from_numpy(rng.uniform(0, 1, (2, geo.nx, geo.ny)).astype(np.float32)).cuda()
    try:
        autotune(enable=False, reset=True)
        y0 = plan.forward(x, fp)
        b0 = plan.backward(y0, bp)
        autotune(enable=True, reset=True)
        y1 = plan.forward(x, fp)
        b1 = plan.backward(y0, bp)
        rep = autotune()
    finally:
        autotune(enable=True)
    assert torch.equal(y0, y1)
    assert torch.equal(b0, b1)
    assert rep, "no launch was tuned"
    for key, (choice, ms) in rep.items():
        assert 0 <= choice < len(ms) and len(ms) >= 3
        assert all(m > 0 for m in ms)
        # the first candidate within 3 % of the fastest (compared in float32,
        # as the tuner does)
        lim = np.float32(1.03) * np.float32(min(ms))
        assert np.float32(ms[choice]) <= lim
        assert all(np.float32(m) > lim for m in ms[:choice])


@pytest.mark.parametrize("ns", [1, 3, 11])
def test_every_launch_shape_agrees(ns):
    # every candidate of the Siddon FP / BP and DD FP / BP / FBP launches
    # (block shapes, XCD order, slices per lane S = 1, 2, 4, 8, groups of 4
    # interleaved slices and line-range parts) on slice counts that leave
    # partial groups, against candidate 0
    from reconnet_amd._lib import autotune, autotune_pin
    from reconnet_amd.torch_ops import FanPlan

    geo = small_fan(nx=96, nu=140, nview=60, dx=1.3)
    plan = FanPlan(geo)
    rng = np.random.default_rng(11 + ns)
    x = torch.from_numpy(rng.uniform(0, 1, (ns, geo.nx, geo.ny)).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.uniform(0, 1, (ns, geo.nu, geo.nview)).astype(np.float32)).cuda()
    counts = {}
    try:
        autotune(enable=True, reset=True)
        plan.forward(x, "siddon")
        plan.forward(x, "dd")
        plan.backward(y, "siddon")
        plan.backward(y, "dd")
        plan.backward(y, "dd_fbp")
        plan.backward(y, "dd_adjoint")
        plan.backward(y, "pixel")
        for (kernel, *_), (_, ms) in autotune().items():
            counts[kernel] = len(ms)
        assert counts.get(1, 0) >= 13 and counts.get(2, 0) >= 13
        # 4 or more slices take the DD FP and BP on the interleaved 4-slice layout
        dd_fp, dd_bp, dd_fbp = ((("dd_fp", 3), ("dd_bp", 5), ("dd_fbp", 7)) if ns < 4 else
                                (("dd_fp_q", 12), ("dd_bp_q", 10), ("dd_fbp_q", 11)))
        assert counts.get(dd_bp[1], 0) >= 5 and counts.get(dd_fbp[1], 0) >= 5, counts
        assert counts.get(dd_fp[1], 0) >= 8 and counts.get(4, 0) >= 8 and counts.get(17, 0) >= 8, counts
        refs = {}
        for kname, kid, run in (("siddon_fp", 1, lambda: plan.forward(x, "siddon")),
                                (dd_fp[0], dd_fp[1], lambda: plan.forward(x, "dd")),
                                ("dd_adjoint", 4, lambda: plan.backward(y, "dd_adjoint")),
                                ("siddon_bp", 2, lambda: plan.backward(y, "siddon")),
                                (dd_bp[0], dd_bp[1], lambda: plan.backward(y, "dd")),
                                (dd_fbp[0], dd_fbp[1], lambda: plan.backward(y, "dd_fbp")),
                                ("pixel_bp", 17, lambda: plan.backward(y, "pixel"))):
            for c in range(counts[kid]):
                autotune_pin(kname, c)
                out = run()
                torch.cuda.synchronize()
                if c == 0:
                    refs[kname] = out
                    continue
                ref = refs[kname]
                err = float((out - ref).abs().max() / ref.abs().max())
                assert err <= 1e-6, (kname, c, err)
            autotune_pin(kname, None)
        autotune_pin("siddon_fp", counts[1])
        with pytest.raises(RuntimeError, match="out of range"):
            plan.forward(x, "siddon")
    finally:
        for k in ("siddon_fp", "siddon_bp", "dd_fp", "dd_fp_q", "dd_adjoint", "dd_bp", "dd_fbp", "dd_bp_q",
                  "dd_fbp_q", "pixel_bp"):
            autotune_pin(k, None)
        autotune(enable=True, reset=True)


def _engine(ns, seed=3, rect=False):
    import torch
    from reconnet_amd.geometry import FanGeometry
    from reconnet_amd.sqs import OsSqsTV

    if rect:
        geo = FanGeometry(nx=90, ny=61, dx=1.3, dy=0.95, nu=140, nview=60, da=1.2 / 1085.6, dsd=1085.6, dso=595.0)
    else:
        geo = small_fan(nx=96, nu=140, nview=60, dx=1.3)
    rng = np.random.default_rng(seed)
    b = torch.from_numpy(rng.uniform(0, 2, (ns, geo.nu, geo.nview)).astype(np.float32)).cuda()
    w = torch.from_numpy(rng.uniform(0.5, 1, (ns, geo.nu, geo.nview)).astype(np.float32)).cuda()
    x0 = torch.from_numpy(rng.uniform(0, 0.05, (ns, geo.nx, geo.ny)).astype(np.float32)).cuda()
    eng = OsSqsTV(geo, 6, ns)
    eng.set_data(b, w)
    eng.compute_norm()
    return eng, x0


def test_every_interleaved_launch_shape_agrees():
    # the OS-SQS engine's 4-slice interleaved FP / BP: every candidate gives
    # the pass of candidate 0 (6 slices: one full and one padded group)
    from reconnet_amd._lib import autotune, autotune_pin

    eng, x0 = _engine(6)
    try:
        autotune(enable=True, reset=True)
        eng.set_image(x0)
        eng.run_pass(beta=1e-3, eps=1e-4)
        counts = {k[0]: len(ms) for k, (_, ms) in autotune().items()}
        assert counts.get(8, 0) >= 7 and counts.get(9, 0) >= 8, counts
        for kname, kid in (("siddon_fp_q", 8), ("siddon_bp_q", 9)):
            ref = None
            for c in range(counts[kid]):
                autotune_pin(kname, c)
                eng.set_image(x0)
                eng.run_pass(beta=1e-3, eps=1e-4)
                out = eng.image()
                if ref is None:
                    ref = out
                    continue
                err = float((out - ref).abs().max() / ref.abs().max())
                assert err <= 1e-6, (kname, c, err)
            autotune_pin(kname, None)
    finally:
        autotune_pin("siddon_fp_q", None)
        autotune_pin("siddon_bp_q", None)
        autotune(enable=True, reset=True)
        eng.close()


@pytest.mark.parametrize("beta", [0.0, 1e-3])
def test_fused_update_matches_separate_update(beta):
    # 32 slices (8 groups of 4 per lane): two OS-SQS passes with the update in
    # the staged backprojection's epilogue (shape 0; image and TV neighbours read
    # from the ping-pong interleaved copies) against two passes with a plain
    # gather shape (index 7: no staging) and the separate update kernel; the
    # second pass reads the copies the first one wrote
    from reconnet_amd._lib import autotune_pin

    eng, x0 = _engine(32)
    try:
        outs = []
        for c in (0, 7):
            autotune_pin("siddon_bp_q", c)
            eng.set_image(x0)
            for _ in range(2):
                eng.run_pass(beta=beta, eps=1e-4, lb=0.0)
            outs.append(eng.image().clone())
        err = float((outs[0] - outs[1]).abs().max() / outs[1].abs().max())
        assert err <= 1e-6, err
        assert float(outs[1].abs().max()) > 0
    finally:
        autotune_pin("siddon_bp_q", None)
        eng.close()


@pytest.mark.parametrize("ns,rect,pin", [(4, False, None), (7, False, None), (6, True, None), (6, True, 0)])
def test_interleaved_engine_matches_plain_layout(ns, rect, pin, monkeypatch):
    # the same passes with the engine's plain one-slice layout
    # (RECONNET_SQS_INTERLEAVE=0) and with the interleaved 4-slice layout,
    # also on a rectangular grid whose two orientations have different pitches;
    # pin = 0 forces the staged backprojection shape, whose epilogue carries the
    # update, with a partly filled 4-slice group (6 slices) on nx != ny
    from reconnet_amd._lib import autotune_pin

    outs = []
    try:
        if pin is not None:
            autotune_pin("siddon_bp_q", pin)
        for flag in ("0", "1"):
            monkeypatch.setenv("RECONNET_SQS_INTERLEAVE", flag)
            eng, x0 = _engine(ns, rect=rect)
            eng.set_image(x0)
            for _ in range(2):
                eng.run_pass(beta=1e-3, eps=1e-4, lb=0.0)
            outs.append(eng.image())
            eng.close()
    finally:
        autotune_pin("siddon_bp_q", None)
    err = float((outs[0] - outs[1]).abs().max() / outs[0].abs().max())
    assert err <= 1e-6, err


@pytest.mark.parametrize("rect", [False, True])
def test_fp_per_class_launches_match_one_launch(rect, monkeypatch):
    # the interleaved engine's forward projection as one launch per subset
    # (default) and as one launch per view class (RECONNET_SQS_FP_SPLIT=1)
    outs = []
    for flag in ("0", "1"):
        monkeypatch.setenv("RECONNET_SQS_FP_SPLIT", flag)
        eng, x0 = _engine(6, rect=rect)
        eng.set_image(x0)
        for _ in range(2):
            eng.run_pass(beta=1e-3, eps=1e-4, lb=0.0)
        outs.append(eng.image())
        eng.close()
    err = float((outs[0] - outs[1]).abs().max() / outs[0].abs().max())
    assert err <= 1e-6, err


def _outside_fan_geometry():
    # a 128 x 128 image of 1 mm pixels under a 64-bin fan that covers a disc
    # of about 21 mm: most 8 x 8 pixel waves meet no bin in any view
    from reconnet_amd.geometry import FanGeometry

    return FanGeometry(nx=128, ny=128, dx=1.0, dy=1.0, nu=64, nview=48, da=1.2 / 1085.6, dsd=1085.6, dso=595.0)


def test_staged_siddon_bp_reads_nothing_outside_its_window(monkeypatch):
    # The LDS-staged BP (candidate 0 of the interleaved BP) after a launch that
    # left non-finite values in LDS: waves with an empty bin window (and lanes
    # past their last bin) must read nothing of it.  Checked against the plain
    # one-slice layout, whose gathers do not use LDS.
    import torch
    from reconnet_amd._lib import autotune_pin
    from reconnet_amd.sqs import OsSqsTV

    geo = _outside_fan_geometry()
    ns = 8
    rng = np.random.default_rng(21)
    b = torch.from_numpy(rng.uniform(0, 2, (ns, geo.nu, geo.nview)).astype(np.float32)).cuda()
    w = torch.from_numpy(rng.uniform(0.5, 1, (ns, geo.nu, geo.nview)).astype(np.float32)).cuda()
    x0 = torch.from_numpy(rng.uniform(0, 0.05, (ns, geo.nx, geo.ny)).astype(np.float32)).cuda()
    outs = []
    try:
        autotune_pin("siddon_bp_q", 0)
        for flag in ("0", "1"):
            monkeypatch.setenv("RECONNET_SQS_INTERLEAVE", flag)
            if flag == "1":
                # fill the LDS of every CU with the staged windows of a NaN residual
                bad = OsSqsTV(geo, 2, ns)
                bad.set_data(torch.full_like(b, float("nan")), w)
                bad.set_norm(torch.ones_like(x0))
                bad.set_image(x0)
                bad.run_pass(beta=0.0, eps=1e-4, lb=0.0)
                torch.cuda.synchronize()
                bad.close()
            eng = OsSqsTV(geo, 2, ns)
            eng.set_data(b, w)
            eng.set_norm(torch.ones_like(x0) * 50.0)
            eng.set_image(x0)
            eng.run_pass(beta=0.0, eps=1e-4, lb=0.0)
            outs.append(eng.image())
            eng.close()
    finally:
        autotune_pin("siddon_bp_q", None)
    assert bool(torch.isfinite(outs[1]).all())
    err = float((outs[0] - outs[1]).abs().max() / outs[0].abs().max())
    assert err <= 1e-6, err


def test_staged_dd_fbp_reads_nothing_outside_its_window():
    # the LDS-staged DD FBP backprojection (candidate 0) on 8 slices after a
    # launch on a NaN sinogram, against the oracle
    import oracle as orc
    from reconnet_amd._lib import autotune_pin
    from reconnet_amd.geometry import fan_angles
    from tests.helpers import recon_for, rel_l2

    geo = _outside_fan_geometry()
    deg = fan_angles(geo.nview)
    sino = np.random.default_rng(22).uniform(0, 1, (8, geo.nu, geo.nview, 1, 1)).astype(np.float32)
    r = recon_for(geo)
    try:
        autotune_pin("dd_fbp_q", 0)
        r.cDDFanBackprojection4d(np.full_like(sino, np.nan), type_projector=1)
        out = r.cDDFanBackprojection4d(sino, type_projector=1)
    finally:
        autotune_pin("dd_fbp_q", None)
    ref = orc.dd_fan_bp(sino, deg, geo.oracle_geo(), (geo.nx, geo.ny, 1), type_projector=1)
    assert np.isfinite(out).all()
    for s in range(8):
        assert rel_l2(out[s], ref[s]) <= 1e-5, s
