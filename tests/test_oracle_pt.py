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


def _fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    ph = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], axis=1)


def _exit_dist(p, d):
    """Distance from points p (inside the unit box) to its boundary along unit directions d."""
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(d > 0, (0.5 - p) / d, np.where(d < 0, (-0.5 - p) / d, np.inf))
    return t.min(axis=-1)


def _edge(u, N):
    """Per-axis factor of a constant grid's zero-border trilinear density at coordinate u: 1
    inside, falling linearly to 1/2 at the faces over the last half texel."""
    return np.clip((0.5 + 0.5 / N - np.abs(u)) * N, 0.0, 1.0)


def _up_depth(y, N):
    """int_y^0.5 of the y factor (light along +y from height y to the top face)."""
    h = 0.5 / N
    y = np.asarray(y, np.float64)
    inner = np.clip(0.5 - h - y, 0.0, None)
    lo = np.maximum(y, 0.5 - h)  # start of the ramp part
    a, b = 0.5 + h - lo, h  # the ramp integrand N*s over s from b to a
    return inner + N * (a * a - b * b) / 2


def test_point_sample_empty_is_black_and_furnace(orc):
    """GetSample (core/render.cuh:615-654): an empty volume without a sky gives 0; a white
    sky with albedo 1.001 * scatter = 1 and no sun gives exactly the sky (every path escapes
    with throughput 1, up to float rounding of the per-vertex factor)."""
    N = 16
    n = 6
    rng = np.random.default_rng(0)
    ori = rng.uniform(-0.3, 0.3, (n, 3)).astype(np.float32)
    dirs = _fib_sphere(n).astype(np.float32)
    lights = np.tile(np.array([0, 1, 0], np.float32), (n, 1))
    empty = orc.Scene(np.zeros(N ** 3, np.float32), N)
    mean, m2 = empty.point_radiance(ori, dirs, lights, [8.0] * n, [0.857] * n, [0.99] * n, [1, 1, 1], 64, 64, 3)
    assert np.all(mean == 0) and np.all(m2 == 0)
    full = orc.Scene(np.full(N ** 3, 0.5, np.float32), N)
    orc.set_env(np.ones((8, 16, 4), np.float32), 1.0)
    try:
        sc = np.float32(1.0) / np.float32(1.001)
        mean, m2 = full.point_radiance(ori, dirs, lights, [4.0] * n, [0.6] * n, [sc] * n, [0, 0, 0], 512, 400, 5)
    finally:
        orc.set_env(None)
    assert np.all(np.abs(mean - 1.0) < 1e-4), mean


def test_point_sample_single_scatter_closed_form(orc):
    """max_bounce 1, g = 0 (uniform first direction, phase 0.25), constant cube, point at the
    centre, sun along +y: E[L] = mean over directions w of int_0^D(w) sigma(t) e^{-tau(t)} * a *
    0.25 * exp(-tau_L(t)) dt, a = 1.001 * scatter, with the grid's zero-border density (sigma
    falls to half at the faces over the last half texel) on the view and the light paths."""
    N, alpha, scatter = 32, 3.0, 0.8
    sc = orc.Scene(np.ones(N ** 3, np.float32), N)
    ori = np.zeros((1, 3), np.float32)
    mean, m2 = sc.point_radiance(ori, [[1, 0, 0]], [[0, 1, 0]], [alpha], [0.0], [scatter], [1, 1, 1], 1, 200000, 9)
    w = _fib_sphere(4000)
    D = _exit_dist(np.zeros((1, 3)), w)
    a = 1.001 * scatter
    tot = 0.0
    for k in range(w.shape[0]):
        ts = np.linspace(0.0, D[k], 2000)
        p = ts[:, None] * w[k][None, :]
        sig = alpha * _edge(p[:, 0], N) * _edge(p[:, 1], N) * _edge(p[:, 2], N)
        tau = np.concatenate([[0.0], np.cumsum(0.5 * (sig[1:] + sig[:-1]) * np.diff(ts))])
        tau_l = alpha * _edge(p[:, 0], N) * _edge(p[:, 2], N) * _up_depth(p[:, 1], N)
        f = sig * np.exp(-tau) * a * 0.25 * np.exp(-tau_l)
        tot += np.trapezoid(f, ts)
    want = tot / w.shape[0]
    se = np.sqrt(m2[0, 0] / (200000 * 199999))
    assert abs(mean[0, 0] - want) < 4 * se + 2e-3 * want, (mean[0, 0], want, se)


def test_cached_bake_homogeneous_and_empty(orc):
    """Fill_BakeTR (core/volume.cu:1562-1587) on a constant cube: every texel estimates
    exp(-alpha c I) with I the density integral from its centre to the exit along the light
    (the zero-border ramp of the last half texel included); the mean over each y layer
    (16 x 16 texels, 5 estimates each) matches within 5 standard errors.  An empty volume
    bakes exactly 1."""
    N, res, alpha, c = 32, 16, 6.0, 0.5
    sc = orc.Scene(np.full(N ** 3, c, np.float32), N)
    t = np.arange(res ** 3)
    got = sc.bake_cached_tr([0, 1, 0], alpha, 1.0, 4, res, t).reshape(res, res, res)  # [x, y, z]
    for y in range(res):
        yc = (y + 0.5) / res - 0.5
        want = np.exp(-alpha * c * _up_depth(yc, N))
        layer = got[:, y, :].ravel()
        assert abs(layer.mean() - want) < 5 * layer.std() / np.sqrt(layer.size) + 1e-6, (y, layer.mean(), want)
    empty = orc.Scene(np.zeros(N ** 3, np.float32), N)
    assert np.all(empty.bake_cached_tr([0.3, 0.8, 0.2], alpha, 1.0, 4, res, t) == 1.0)


def test_tr_samples_homogeneous(orc):
    """ShadowTerm_TR(s) (core/render.cuh:656-680): mean Tr * HG = exp(-alpha c I) * HG(cos, g),
    I the density integral to the exit (the zero-border ramp included)."""
    N, alpha, c, g = 32, 5.0, 0.8, 0.6
    sc = orc.Scene(np.full(N ** 3, c, np.float32), N)
    ori = np.array([[0.0, -0.2, 0.1], [0.2, 0.1, -0.3]], np.float32)
    dirs = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], np.float32)
    spp = 1000000
    got = sc.tr_samples(ori, dirs, [0, 1, 0], alpha, g, spp, 2)
    for i in range(2):
        cs = dirs[i] @ np.array([0, 1, 0])
        hg = 0.25 * (1 - g * g) / (1 + g * g - 2 * g * cs) ** 1.5
        T = np.exp(-alpha * c * _up_depth(ori[i, 1], N))
        # each estimate lies in [0, 1], so its variance is at most T: 4 standard errors
        assert abs(got[i] - T * hg) < 4 * hg * np.sqrt(T / spp), (i, got[i], T * hg)
