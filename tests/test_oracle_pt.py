"""Delta-tracking path tracer of the oracle (core/render.cuh:60-103, 399-434, 535-597)
pinned by closed forms and statistics."""
import numpy as np
import pytest


def test_empty_volume_is_black(orc):
    N = 16
    sc = orc.Scene(np.zeros(N ** 3, np.float32), N)
    sc.update_tr(np.array([0, 1, 0], np.float32), 64.0)
    v = orc.orbit_view(8, 8)
    mean, m2, taps = sc.pt_render(v, 8, 16)
    assert np.all(mean == 0) and np.all(m2 == 0)


def test_transmittance_estimator_homogeneous(orc):
    # ratio tracking to 0.5 then delta tracking: unbiased for exp(-sigma d)
    N = 32
    g = np.ones(N ** 3, np.float32)
    for alpha, d in [(2.0, 0.3), (8.0, 0.25), (20.0, 0.2)]:
        est = orc.tr_stochastic(g, N, 1.0, [0.0, 0.0, -0.2], [0.0, 0.0, 1.0], d, alpha, 200000)
        want = np.exp(-alpha * d)
        assert abs(est.mean() - want) < 4 * est.std() / np.sqrt(est.size) + 1e-4


def test_single_scatter_closed_form(orc):
    """max_bounce = 1, homogeneous cube: E[L] = int_0^D sigma e^{-sigma t} * albedo * HG(cos) *
    exp(-sigma * l(t)) dt, l(t) = distance to the exit along the light (numerical quadrature)."""
    N, alpha, g, albedo = 32, 3.0, 0.5, 0.9
    sc = orc.Scene(np.ones(N ** 3, np.float32), N)
    L = np.array([0.0, 1.0, 0.0], np.float32)
    sc.update_tr(L, alpha)
    v = orc.orbit_view(1, 1, dist=2.0, azimuth=0.0, elevation=0.0, fov_deg=10.0, light=L, alpha=alpha, g=g,
                       albedo=albedo, jitter=0.5)
    mean, m2, _ = sc.pt_render(v, 40000, 1, nee_baked=0)
    # camera on +x looking at the origin: the ray enters at x = 0.4999 and travels along -x
    ts = np.linspace(0, 0.9999, 20001)
    pts_y = np.zeros_like(ts)  # ray along the x axis through the centre: y = 0 -> light path 0.5
    lpath = 0.5 - pts_y
    cosang = 0.0  # ray direction (-x) . light (+y)
    hg = 0.25 * (1 - g * g) / (1 + g * g - 2 * g * cosang) ** 1.5
    f = alpha * np.exp(-alpha * ts) * albedo * hg * np.exp(-alpha * lpath)
    want = np.trapezoid(f, ts)
    se = np.sqrt(m2[0, 0] / (40000 * 39999))
    assert abs(mean[0, 0] - want) < 4 * se + 1e-4 * want
