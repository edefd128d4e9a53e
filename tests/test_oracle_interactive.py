"""Oracle of the interactive progressive path (core/volume.cu:57-372, core/render.cuh:1144-1236),
pinned independently of its code:

  * ACES: an float64 pipeline whose colour matrices are derived here from the published
    primaries and white points (ACES AP0 / AP1, Rec.709, Bradford adaptation D60 <-> D65) and
    whose tone curve is the published rational fit; Gamma is x^(1/2.2) in closed form;
  * the histogram distance Compare: 0 between proportional histograms, symmetric, and the
    closed form (n + m) / 2 for two histograms with disjoint single-bin channels;
  * Denoise: identical histograms give the 11x11 box mean with clamped borders (numpy
    edge padding), dissimilar ones give the pixel itself, quiet pixels skip the filter, frame 0
    fills the checkerboard's other half with the mean of four clamped neighbours;
  * the progressive blend: after K frames the target is the mean of the frames that rendered
    the pixel, the histogram counts and the running luminance moments are closed forms.
"""
import math

import numpy as np
import pytest


def _rgb2xyz(prim, white):
    xy = np.array(prim, float)
    P = np.stack([xy[:, 0] / xy[:, 1], np.ones(3), (1 - xy[:, 0] - xy[:, 1]) / xy[:, 1]])
    W = np.array([white[0] / white[1], 1.0, (1 - white[0] - white[1]) / white[1]])
    return P * np.linalg.solve(P, W)


def _bradford(ws, wd):
    B = np.array([[0.8951, 0.2664, -0.1614], [-0.7502, 1.7135, 0.0367], [0.0389, -0.0685, 1.0296]])
    xyz = lambda w: np.array([w[0] / w[1], 1, (1 - w[0] - w[1]) / w[1]])
    return np.linalg.inv(B) @ np.diag((B @ xyz(wd)) / (B @ xyz(ws))) @ B


D60, D65 = (0.32168, 0.33767), (0.3127, 0.3290)
AP0 = _rgb2xyz([(0.7347, 0.2653), (0.0, 1.0), (0.0001, -0.0770)], D60)
AP1 = _rgb2xyz([(0.713, 0.293), (0.165, 0.830), (0.128, 0.044)], D60)
R709 = _rgb2xyz([(0.64, 0.33), (0.30, 0.60), (0.15, 0.06)], D65)


def aces_f64(c):
    """The fitted ACES RRT + ODT (Unity's ACES tonemapper form) with derived matrices."""
    srgb_2_ap0 = np.linalg.inv(AP0) @ _bradford(D65, D60) @ R709
    x = np.linalg.inv(AP1) @ AP0 @ (srgb_2_ap0 @ c)
    y_ap1 = AP1[1]  # luminance weights of AP1
    x = y_ap1 @ x + 0.96 * (x - y_ap1 @ x)
    a, b, cc, d, e = 278.5085, 10.7772, 293.6045, 88.7122, 80.6889
    x = (x * (x * a + b)) / (x * (x * cc + d) + e)
    XYZ = AP1 @ x
    s = max(XYZ.sum(), 1e-4)
    xx, yy, Y = XYZ[0] / s, XYZ[1] / s, min(max(XYZ[1], 0.0), 65504.0) ** 0.9811
    XYZ = np.array([xx * Y / max(yy, 1e-4), Y, (1 - xx - yy) * Y / max(yy, 1e-4)])
    x = np.linalg.inv(AP1) @ XYZ
    x = y_ap1 @ x + 0.93 * (x - y_ap1 @ x)
    lin = np.linalg.inv(R709) @ _bradford(D60, D65) @ AP1 @ x
    return np.sign(lin) * np.abs(lin) ** (1 / 2.2)


def test_aces_against_derived_matrices(orc):
    rng = np.random.default_rng(1)
    rgb = np.concatenate([rng.uniform(0, 4, (200, 3)), np.array([[0.18] * 3, [1.0] * 3, [0.02, 0.5, 3.0]])])
    _, got = orc.tonemap(rgb, 1.0, 2)
    want = np.array([aces_f64(c) for c in rgb.astype(np.float64)])
    ok = want > 1e-3  # negative channels: pow gives NaN on both sides and saturates to 0
    np.testing.assert_allclose(got[ok], want[ok], rtol=2e-3, atol=2e-4)
    # neutral stays neutral (to the rounding of the published matrices)
    g = got[200]
    assert abs(g[0] - g[2]) < 5e-4 and 0.30 < g[1] < 0.42


def test_gamma_and_packing(orc):
    rgb = np.array([[0.0, 0.25, 1.0], [2.0, 0.5, 0.01], [-1.0, 0.7, 0.999]], np.float32)
    rgba, mapped = orc.tonemap(rgb, 0.8, 1)
    want = np.where(rgb * 0.8 >= 0, (np.maximum(rgb * 0.8, 0)) ** (1 / 2.2), np.nan)
    np.testing.assert_allclose(mapped[~np.isnan(want)], want[~np.isnan(want)], rtol=2e-7)
    sat = np.nan_to_num(np.clip(want, 0, 1), nan=0.0)
    exp = 0xff000000 | (np.floor(255 * sat[:, 0]).astype(np.uint32) << 16) | \
        (np.floor(255 * sat[:, 1]).astype(np.uint32) << 8) | np.floor(255 * sat[:, 2]).astype(np.uint32)
    assert np.array_equal(rgba, exp.astype(np.uint32))
    rgba0, _ = orc.tonemap(rgb, 1.0, 0)
    assert rgba0[0] == (0xff000000 | (63 << 8) | 255)


def _histo(bins, total, x=0.0, x2=0.0):
    h = np.zeros(33, np.float32)
    h[:30] = bins
    h[30], h[31], h[32] = total, x2, x
    return h


def test_compare_closed_forms(orc):
    rng = np.random.default_rng(2)
    b = rng.integers(0, 9, 30).astype(np.float32)
    x = _histo(b, 10)
    assert orc.histo_compare(x, x) == 0.0
    assert orc.histo_compare(x, _histo(3 * b, 30)) == pytest.approx(0.0, abs=1e-5)
    y = _histo(rng.integers(0, 9, 30), 7)
    assert orc.histo_compare(x, y) == pytest.approx(orc.histo_compare(y, x), rel=1e-6)
    # n samples all in bin 2 of each channel against m samples all in bin 7: (n + m) / 2
    n, m = 12, 5
    bx, by = np.zeros(30), np.zeros(30)
    bx[[2, 12, 22]] = n
    by[[7, 17, 27]] = m
    assert orc.histo_compare(_histo(bx, n), _histo(by, m)) == pytest.approx((n + m) / 2, rel=1e-6)
    assert orc.histo_compare(_histo(np.zeros(30), 0), _histo(np.zeros(30), 0)) == 0.0


def _box_clamped(img, W, H, r=5):
    pad = np.pad(img.reshape(H, W, 3).astype(np.float64), ((r, r), (r, r), (0, 0)), mode="edge")
    out = np.zeros((H, W, 3))
    for di in range(2 * r + 1):
        for dj in range(2 * r + 1):
            out += pad[di:di + H, dj:dj + W]
    return (out / (2 * r + 1) ** 2).reshape(-1, 3)


def test_denoise_filter_closed_forms(orc):
    W, H = 23, 17
    rng = np.random.default_rng(3)
    target = rng.uniform(0, 1, (W * H, 3)).astype(np.float32)
    # identical histograms with luminance variance 0.05: the weighted mean is the box mean
    histo = np.tile(_histo(rng.integers(0, 5, 30), 9, x=0.5, x2=0.3), (W * H, 1)).astype(np.float32)
    got = orc.denoise(target, histo, W, H, 3, 0, 1, 1.0, 0)
    want, _ = orc.tonemap(_box_clamped(target, W, H).astype(np.float32), 1.0, 0)
    diff = np.abs(got.view(np.uint8).reshape(-1, 4).astype(int) - want.view(np.uint8).reshape(-1, 4).astype(int))
    assert diff.max() <= 1 and np.mean(diff == 0) > 0.99
    # every pixel its own triple of bins (40 samples each): Compare = k (40 + 40) / (3 + k) >= 20
    # for k >= 1 differing channels, so every neighbour weighs 0 and the pixel stays itself
    h2 = np.zeros((W * H, 33), np.float32)
    for p in range(W * H):
        bins = np.zeros(30)
        for ch in range(3):
            bins[10 * ch + (p // 10 ** ch) % 10] = 40
        h2[p] = _histo(bins, 40, x=0.5, x2=0.3)
    got = orc.denoise(target, h2, W, H, 3, 0, 1, 1.0, 0)
    plain, _ = orc.tonemap(target, 1.0, 0)
    assert np.array_equal(got, plain)
    # quiet pixels (variance <= 0.01) and denoise off: the pixel itself
    quiet = histo.copy()
    quiet[:, 31], quiet[:, 32] = 0.26, 0.5
    assert np.array_equal(orc.denoise(target, quiet, W, H, 3, 0, 1, 1.0, 0), plain)
    assert np.array_equal(orc.denoise(target, histo, W, H, 3, 0, 0, 1.0, 0), plain)


@pytest.mark.parametrize("flip", [0, 1])
def test_denoise_frame0_checkerboard_fill(orc, flip):
    W, H = 9, 7
    rng = np.random.default_rng(4)
    target = rng.uniform(0, 1, (W * H, 3)).astype(np.float32)
    histo = np.zeros((W * H, 33), np.float32)
    got = orc.denoise(target, histo, W, H, 0, flip, 1, 1.0, 0)
    t = target.reshape(H, W, 3).astype(np.float64)
    pad = np.pad(t, ((1, 1), (1, 1), (0, 0)), mode="edge")
    avg = (pad[:-2, 1:-1] + pad[2:, 1:-1] + pad[1:-1, :-2] + pad[1:-1, 2:]) / 4
    r, c = np.divmod(np.arange(W * H), W)
    keep = (((r + flip) % 2) + c) % 2 == 0
    want = np.where(keep[:, None], target, avg.reshape(-1, 3)).astype(np.float32)
    exp, _ = orc.tonemap(want, 1.0, 0)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("checker", [0, 1])
def test_progressive_blend_and_histogram(orc, checker):
    W, H, K = 8, 6, 9
    rng = np.random.default_rng(5)
    target = np.zeros((W * H, 3), np.float32)
    histo = np.zeros((W * H, 33), np.float32)
    frames = []
    r, c = np.divmod(np.arange(W * H), W)
    for f in range(K):
        flip = f % 2
        rad = np.zeros((W * H, 4), np.float32)
        rad[:, :3] = rng.uniform(-0.2, 3.0, (W * H, 3))
        fnum = f // 2 if checker else f
        orc.pt_accumulate(target, histo, rad, W, H, checker, flip, fnum)
        s = (((r + flip) % 2) + c) % 2 == 0 if checker else np.ones(W * H, bool)
        frames.append((s, np.maximum(rad[:, :3], 0).astype(np.float64)))
    cnt = np.sum([s for s, _ in frames], axis=0)
    mean = np.sum([np.where(s[:, None], v, 0) for s, v in frames], axis=0) / cnt[:, None]
    np.testing.assert_allclose(target, mean, rtol=2e-6, atol=1e-6)
    assert np.array_equal(histo[:, 30], cnt.astype(np.float32))
    tm = [(s, v / (v + 1)) for s, v in frames]
    lum = np.sum([np.where(s, t.sum(1), 0) for s, t in tm], axis=0) / cnt
    lum2 = np.sum([np.where(s, t.sum(1) ** 2, 0) for s, t in tm], axis=0) / cnt
    np.testing.assert_allclose(histo[:, 32], lum, rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(histo[:, 31], lum2, rtol=2e-6, atol=1e-6)
    bins = np.zeros((W * H, 30))
    for s, t in tm:
        b = np.floor(np.minimum(t, 0.999) * 10).astype(int)
        for ch in range(3):
            bins[np.nonzero(s)[0], ch * 10 + b[s, ch]] += 1
    assert np.array_equal(histo[:, :30], bins.astype(np.float32))


def test_reprojection_and_commit(orc):
    W, H = 5, 4
    rng = np.random.default_rng(6)
    target = rng.uniform(0, 1, (W * H, 3)).astype(np.float32)
    rad = rng.uniform(-1, 2, (W * H, 4)).astype(np.float32)
    t1 = target.copy()
    orc.predict_commit(t1, rad, W, H, 1, 1)
    r, c = np.divmod(np.arange(W * H), W)
    s = (((r + 1) % 2) + c) % 2 == 0
    assert np.array_equal(t1[s], np.maximum(rad[s, :3], 0)) and np.array_equal(t1[~s], target[~s])
    histo = np.zeros((W * H, 33), np.float32)
    rgba = orc.reprojection_denoise(t1, histo, 1.3, 2)
    assert np.array_equal(histo[:, 30], t1[:, 0]) and np.array_equal(histo[:, 32], t1[:, 1])
    assert np.array_equal(histo[:, 31], t1[:, 2]) and np.all(histo[:, :30] == 0)
    assert np.array_equal(rgba, orc.tonemap(t1, 1.3, 2)[0])
