"""Host-side logic that needs no GPU: parameter files, angle and subset
recipes, phantoms, noise, short-scan weights and argument plumbing."""

import numpy as np
import pytest

from reconnet_amd.geometry import WORKLOADS, fan_angles, interleaved_order
from reconnet_amd.phantom import add_poisson_noise, ellipse_params, phantom, rasterize
from reconnet_amd.recon import ParkerWeighting, ReconNet, RiessWeighting


def test_defaults_match_reference_class():
    r = ReconNet()
    assert (r.rotview, r.nu, r.nv, r.nx, r.ny, r.nz) == (720, 736, 1, 512, 512, 1)
    assert (r.dsd, r.dso, r.du) == (1085.6, 595, 1.2858)
    assert r.da == pytest.approx(1.2858 / 1085.6)
    assert np.array_equal(np.asarray(r.angles, np.float32), fan_angles(720))


def test_fromfile_semantics(tmp_path):
    p = tmp_path / "param.txt"
    p.write_text("NX 256\nny\t128\nDU 1.0\ndsd 1000\nROTVIEW 90\noff_a 1.5\ndx 0.7\n")
    r = ReconNet()
    r.FromFile(str(p))
    assert r.nx == 256 and r.ny == 128 and isinstance(r.nx, int)
    assert r.rotview == 90 and len(r.angles) == 90
    assert r.off_a == pytest.approx(-1.5)            # negated
    assert r.da == pytest.approx(1.0 / 1000.0)        # recomputed
    assert r.dx == pytest.approx(0.7) and isinstance(r.dx, np.float32)


def test_tvsqs_shape_check():
    with pytest.raises(ValueError):
        ReconNet().cTVSQS2d(np.zeros((2, 8, 8, 1), np.float32))


def test_interleaved_order_is_permutation():
    for n, m in [(1152, 12), (360, 7), (100, 3)]:
        o = interleaved_order(n, m)
        assert np.array_equal(np.sort(o), np.arange(n))


def test_phantom_is_seeded_and_in_range():
    a = phantom(7, 64, 64, 1.0, 1.0)
    b = phantom(7, 64, 64, 1.0, 1.0)
    assert np.array_equal(a, b) and a.dtype == np.float32
    p = ellipse_params(7)
    assert p.shape == (10, 6)
    assert np.all(np.hypot(p[:, 0], p[:, 1]) <= 110 + 1e-9)
    assert np.all((p[:, 2] >= 5) & (p[:, 2] <= 80) & (p[:, 5] >= 0) & (p[:, 5] <= 0.04))
    assert not np.array_equal(a, phantom(8, 64, 64, 1.0, 1.0))


def test_phantom_area_matches_ellipse_sum():
    p = ellipse_params(3, n_ellipses=3, axis_range=(5, 20), r_max=40)
    img = rasterize(p, 256, 256, 0.5, 0.5)
    # cut off (r_max=40 inside a 128 mm image): integral = sum of mu * pi a b
    total = float(img.sum()) * 0.25
    ref = float(np.sum(p[:, 5] * np.pi * p[:, 2] * p[:, 3]))
    assert total == pytest.approx(ref, rel=2e-3)


def test_poisson_noise_statistics():
    p = np.full((200000,), 2.0, np.float32)
    noisy, w = add_poisson_noise(p, 1e5, seed=1)
    counts = w * 1e5
    assert abs(counts.mean() - 1e5 * np.exp(-2.0)) < 3 * np.sqrt(1e5 * np.exp(-2.0) / 200000) * 3
    assert noisy.mean() == pytest.approx(2.0, abs=2e-3)
    n2, _ = add_poisson_noise(p, 1e5, seed=1)
    assert np.array_equal(noisy, n2)


def _short_scan(nu, m, gam_bins, off_a):
    # da equal to the view step pi/m, and the scan pi + gam_bins*da, so that the
    # conjugate of every ray lands on a sample of the same sinogram
    r = ReconNet()
    r.nu, r.da, r.off_a = nu, np.pi / m, off_a
    r.rotview = m + gam_bins + 1
    r.angles = list((np.arange(r.rotview) * r.da).astype(np.float32))
    return r


@pytest.mark.parametrize("kind", ["parker", "riess"])
@pytest.mark.parametrize("gam_bins,off_a", [(64, 0.0), (40, 0.0), (64, 1.5)])
def test_short_scan_weights_sum_to_one_on_conjugate_rays(kind, gam_bins, off_a):
    # the ray (beta, alpha) and its conjugate (beta + pi + 2 alpha, -alpha) see
    # the same line; redundancy weights of the pair add up to 1.  Checked on the
    # whole weight map (every pair inside the scan), for a scan covering the
    # full fan (gam_bins = nu) and a shorter one, and with a detector offset.
    nu, m = 64, 200
    r = _short_scan(nu, m, gam_bins, off_a)
    if kind == "parker":
        w = ParkerWeighting(r)[0, :, :, 0].astype(np.float64)
    else:
        # sigma below one bin: the Gaussian smoothing is the identity
        w = RiessWeighting(r, sigma=1e-3)[0, :, :, 0].astype(np.float64)
    assert w.shape == (nu, r.rotview)
    sums = []
    for iu in range(nu):
        two_alpha = 2 * (iu - (nu - 1) / 2 - off_a)            # in units of da
        ju = nu - 1 - iu + int(round(2 * off_a))               # bin of -alpha
        for k in range(r.rotview):
            k2 = k + m + int(round(two_alpha))
            if 0 <= k2 < r.rotview and 0 <= ju < nu:
                sums.append(w[iu, k] + w[ju, k2])
    sums = np.array(sums)
    assert len(sums) > 2000
    assert np.abs(sums - 1).max() < 2e-5                  # float32 angles
    alpha = np.arange(nu) - (nu - 1) / 2 - off_a
    inside = np.abs(alpha) < gam_bins / 2                  # rays within the scan's fan
    assert w[inside, 0].max() < 1e-6                       # rise starts at 0
    mid = r.rotview // 2
    if gam_bins == 64 and off_a == 0:
        assert np.allclose(w[:, mid], 1.0)                 # centre of the scan: no redundancy
    if kind == "riess" and gam_bins == 40:
        assert w.max() > 1.5                               # Riess extends past [0, 1]


def test_riess_smoothing_keeps_only_element_zero_in_the_central_range():
    # ReconNet_no_tf.py:685-688: every view is smoothed along the detector; for
    # views in (theta, pi + 2 gamma - theta) the reference restores the
    # unsmoothed value, which its index expression limits to element 0
    r = _short_scan(64, 200, 40, 0.0)
    raw = RiessWeighting(r, sigma=1e-3)[0, :, :, 0]
    sm = RiessWeighting(r)[0, :, :, 0]
    betas = np.asarray(r.angles, np.float64)
    gamma = (betas[-1] - np.pi) / 2
    theta = 10 * np.pi / 180
    central = (betas > theta) & (betas < np.pi + 2 * gamma - theta)
    assert central.any() and (~central).any()
    assert np.array_equal(sm[0, central], raw[0, central])
    assert not np.allclose(sm[1:, central], raw[1:, central])
    assert not np.allclose(sm[0, ~central], raw[0, ~central])
    # smoothing a column that is constant leaves it unchanged
    const = np.all(raw == raw[:1], axis=0)
    assert np.allclose(sm[:, const], raw[:, const], atol=1e-6)


def test_workloads_match_benchmark_configs():
    c0, c1 = WORKLOADS["fan256"], WORKLOADS["fan512"]
    assert (c0.nx, c0.nu, c0.nview, c0.dx) == (256, 512, 360, 1.0)
    assert c0.da == pytest.approx(1.2 / 1085.6)
    assert (c1.nx, c1.nu, c1.nview, c1.dx) == (512, 736, 1152, 0.7)
    assert c1.da == pytest.approx(1.0 / 1085.6)
    assert c1.fov_radius() == pytest.approx(179.2, abs=0.1)


def test_convergence_rule():
    from reconnet_amd.metrics import convergence_pass

    # RMSE reaches 1% of its final value at pass 3, the update drops below 1e-4 at pass 5
    r = [10.0, 5.0, 4.04, 4.02, 4.01, 4.0]
    u = [1e-2, 1e-3, 5e-4, 2e-4, 9e-5, 5e-5]
    assert convergence_pass(r, u) == 5
    # never converged: k_max
    assert convergence_pass(r, [1.0] * 6) == 6
    # a dip below the final RMSE counts only with a small update
    assert convergence_pass([9.0, 3.0, 4.0], [1e-5, 1e-5, 1e-5]) == 2
    with pytest.raises(ValueError):
        convergence_pass([1.0], [1.0, 2.0])


def test_rmse_hu_scale():
    from reconnet_amd.metrics import rmse_hu

    t = np.zeros((2, 4, 4))
    x = t.copy()
    x[1] += 0.019      # one mu_water everywhere = 1000 HU
    m = np.ones((4, 4))
    assert np.allclose(rmse_hu(x, t, m), [0.0, 1000.0])


def test_custom_ops_registered():
    import torch

    from reconnet_amd import ops

    for name in ("fan_forward", "fan_backward", "cone_forward", "cone_backward"):
        assert hasattr(torch.ops.reconnet, name)
    with pytest.raises(TypeError):
        ops.register_plan(object())
    with pytest.raises(ValueError):
        ops.fan_backproject(torch.zeros(1, 2, 3), 1, "nonsense")
