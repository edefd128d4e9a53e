"""Oracle of the tile feature gather, pinned by closed forms and independent root finding:

  * the tau-threshold march (DeterminateEstimatedHit float4 version with CheckExitCondition,
    core/render.cuh:199-397, via GetSampleT render.cuh:464-486): on strictly positive fields
    the march is the composite trapezoid rule over samples one step apart, starting from a
    zero sample one step before the entry point; the hits are where that optical depth
    reaches -ln 0.75 and -ln 0.25, found here by bisection on the piecewise-linear density
    in float64 (no quadratic formula), and the opacity comes from the depth before the second
    crossing;
  * a constant medium: tau(t) = c alpha (t + 1e-4) past the first step (the entry plane is
    the 0.4999 box of RayBoxOffset, vector.cu:45-63), so the hits are closed forms;
  * the 773-float row (GatherSamples volume.cu:257-283, NNGatherSamples render.cuh:1049-1142,
    SetupInputData2 radiancePredict.cu:487-565, layout mrpnn_helpers.py:188-208) on a constant
    medium: every sub-ray's t grid, the nearest / farthest choice of each 2x2 group, the
    sample features, the gamma terms and the t differences against float64 closed forms, and
    the stencil features against the predictor's gather of the same point and basis;
  * empty volumes and rays that miss the box.
"""
import math

import numpy as np
import pytest

TAU1 = -math.log(0.75)
TAU2 = -math.log(0.25)
TAU_STEP_MAX = -math.log(0.76)


def tri(grid, N, pts):
    """Zero-border trilinear of a cube grid (x slowest) at points of [-0.5, 0.5]^3, float64."""
    g = grid.reshape(N, N, N).astype(np.float64)
    p = (np.asarray(pts, np.float64) + 0.5) * N - 0.5
    i0 = np.floor(p).astype(np.int64)
    w = p - i0
    acc = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                ix, iy, iz = i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz
                ok = (ix >= 0) & (ix < N) & (iy >= 0) & (iy < N) & (iz >= 0) & (iz < N)
                v = np.where(ok, g[np.clip(ix, 0, N - 1), np.clip(iy, 0, N - 1), np.clip(iz, 0, N - 1)], 0.0)
                wt = (w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) * (w[:, 2] if dz else 1 - w[:, 2])
                acc += v * wt
    return acc


def slab(ori, d, half):
    """Entry and exit distances of a ray through [-half, half]^3 (float64); None if it misses."""
    ori, d = np.asarray(ori, np.float64), np.asarray(d, np.float64)
    with np.errstate(divide="ignore"):
        inv = 1.0 / d
    t0, t1 = (-half - ori) * inv, (half - ori) * inv
    tmi, tma = np.minimum(t0, t1).max(), np.maximum(t0, t1).min()
    return (max(tmi, 0.0), tma) if tma >= tmi else None


def expected_ts(grid, N, max_density, ori, d, alpha):
    """GetSampleT on a field that is positive everywhere inside the box: trapezoid optical
    depth over samples t_k (float32 accumulation of the step, as the march does), hits by
    bisection of the linear density between samples."""
    ori, d = np.asarray(ori, np.float64), np.asarray(d, np.float64)
    hit = slab(ori, d, 0.4999)
    if hit is None:
        return np.zeros(4)
    dis = hit[0]
    sp = ori + d * dis
    max_dis = slab(sp, d, 0.5)[1]
    smax = max_density * alpha
    h = np.float32(1.0 / N)
    if smax * float(h) > TAU_STEP_MAX:
        h = np.float32(np.float32(TAU_STEP_MAX) / np.float32(smax))
    ts = [np.float32(0.0)]
    while float(ts[-1]) + float(h) <= max_dis:
        ts.append(np.float32(ts[-1] + h))
    t = np.array(ts, np.float64)
    S = alpha * tri(grid, N, sp[None, :] + t[:, None] * d[None, :])
    assert S.min() > 0
    Sp = np.concatenate([[0.0], S])
    tau = np.cumsum((Sp[1:] + Sp[:-1]) * 0.5 * float(h))
    tau_before = np.concatenate([[0.0], tau[:-1]])

    def root(k, target):
        a, b = 0.0, 1.0
        s0, s1 = Sp[k], Sp[k + 1]
        for _ in range(80):
            m = 0.5 * (a + b)
            if tau_before[k] + float(h) * (s0 * m + 0.5 * (s1 - s0) * m * m) < target:
                a = m
            else:
                b = m
        return t[k] - float(h) * (1.0 - 0.5 * (a + b))

    out = np.zeros(4)
    k1 = np.nonzero(tau >= TAU1)[0]
    k2 = np.nonzero(tau >= TAU2)[0]
    if len(k1):
        assert k1[0] >= 1, "crossing inside the phantom first segment: pick another case"
        out[0] = dis + root(k1[0], TAU1)
        out[2] = dis
    if len(k2):
        out[1] = dis + root(k2[0], TAU2)
        tau_exit = tau_before[k2[0]]
    else:
        tau_exit = tau[-1]
    out[3] = min(max(1.0 - math.exp(-tau_exit), 0.0) / 0.75, 1.0)
    return out


def smooth_field(N, seed):
    rng = np.random.default_rng(seed)
    x = (np.arange(N) + 0.5) / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = np.full((N, N, N), 0.5)
    for _ in range(3):
        k = rng.uniform(1.0, 4.0, 3)
        ph = rng.uniform(0, 2 * np.pi)
        f += 0.1 * np.sin(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z) + ph)
    return f.astype(np.float32).ravel()


@pytest.mark.parametrize("alpha", [4.0, 16.0])
def test_sample_t_trapezoid_bisection(orc, alpha):
    """Smooth positive field, random rays; alpha 16 takes the reduced-step branch
    (max density x alpha / N above -ln 0.76)."""
    N = 32
    grid = smooth_field(N, 3)
    md = float(grid.max())
    assert (md * alpha / N > TAU_STEP_MAX) == (alpha == 16.0)
    rng = np.random.default_rng(4)
    n_x = n_both = 0
    for _ in range(24):
        o = rng.normal(size=3)
        o = 2.0 * o / np.linalg.norm(o)
        target = rng.uniform(-0.3, 0.3, 3)
        d = target - o
        d = (d / np.linalg.norm(d)).astype(np.float32)
        o = o.astype(np.float32)
        want = expected_ts(grid, N, md, o, d, alpha)
        got = orc.sample_t(grid, N, md, o, d, alpha)
        np.testing.assert_allclose(got[:3], want[:3], rtol=0, atol=2e-5)
        np.testing.assert_allclose(got[3], want[3], rtol=1e-5, atol=1e-6)
        n_both += want[1] > 0
        n_x += want[0] > 0
    assert n_x == 24 and n_both >= 6


def test_sample_t_low_density_exits_without_hits(orc):
    """Optical depth below -ln 0.75 through the whole box: no hit, t values stay 0 and the
    opacity is that of the whole path."""
    N = 32
    grid = np.full(N ** 3, 0.05, np.float32)
    o = np.array([2.0, 0.1, -0.05], np.float32)
    d = np.array([-1.0, 0.0, 0.0], np.float32)
    got = orc.sample_t(grid, N, 0.05, o, d, 1.0)
    want = expected_ts(grid, N, 0.05, o, d, 1.0)
    assert got[0] == 0 and got[1] == 0 and got[2] == 0
    assert 0 < got[3] < 1.0 / 0.75 * (1 - math.exp(-0.06))
    np.testing.assert_allclose(got[3], want[3], rtol=1e-5)


def test_sample_t_constant_medium_closed_form(orc):
    N, c, alpha = 64, 1.0, 8.0
    grid = np.full(N ** 3, c, np.float32)
    rng = np.random.default_rng(5)
    h = 1.0 / N
    for _ in range(16):
        o = np.array([2.5, rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1)], np.float32)
        tgt = np.array([0.0, rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2)])
        d = (tgt - o) / np.linalg.norm(tgt - o)
        d = d.astype(np.float32)
        dis = slab(o, d, 0.4999)[0]
        k2 = math.ceil(TAU2 / (c * alpha * h) - 1e-4 * N)
        tau_exit = c * alpha * h * (k2 - 1 + 1e-4 * N)
        want = [dis + TAU1 / (c * alpha) - 1e-4, dis + TAU2 / (c * alpha) - 1e-4, dis,
                min((1 - math.exp(-tau_exit)) / 0.75, 1.0)]
        got = orc.sample_t(grid, N, c, o, d, alpha)
        np.testing.assert_allclose(got, want, rtol=0, atol=3e-6)


def test_sample_t_miss_and_empty(orc):
    N = 16
    grid = np.ones(N ** 3, np.float32)
    o = np.array([2.0, 2.0, 0.0], np.float32)
    assert np.all(orc.sample_t(grid, N, 1.0, o, np.array([0, 1, 0], np.float32), 4.0) == 0)
    empty = np.zeros(N ** 3, np.float32)
    d = np.array([-1, 0, 0], np.float32)
    assert np.all(orc.sample_t(empty, N, 0.0, np.array([2, 0.1, 0.1], np.float32), d, 4.0) == 0)


def _tile_view(orc, W, H, ori, light, alpha, g, tan_half):
    v = orc.orbit_view(W, H, light=light, alpha=alpha, g=g, albedo=1.0)
    ori = np.asarray(ori, np.float32)
    fwd = -ori / np.linalg.norm(ori)
    right = np.cross(np.array([0, 1, 0], np.float32), fwd)
    right /= np.linalg.norm(right)
    up = np.cross(fwd, right)
    v.ori[:] = ori.tolist()
    v.right[:] = right.astype(np.float32).tolist()
    v.up[:] = up.astype(np.float32).tolist()
    v.tan_half_fov = tan_half
    return v


def _subrays(v, row, col, off_x, off_y):
    """The 64 sub-ray directions of a pixel (float64), index = 8 * row8 + col8."""
    W, H = v.width, v.height
    ori = np.array(v.ori[:], np.float64)
    right, up = np.array(v.right[:], np.float64), np.array(v.up[:], np.float64)
    sy = v.tan_half_fov
    sx = sy * W / H
    u = 1 - (col + off_x / 8 + 0.5) / W
    vv = 1 - (row + off_y / 8 + 0.5) / H
    fwd = -ori / np.linalg.norm(ori)
    dr = fwd + right * ((2 * u - 1) * sx) + up * ((2 * vv - 1) * sy)
    ro, uo = -right * sx / (4 * W), -up * sy / (4 * H)
    out = np.zeros((64, 3))
    for r8 in range(8):
        for c8 in range(8):
            s = dr + ro * (c8 - 3.5) + uo * (r8 - 3.5)
            out[8 * r8 + c8] = s / np.linalg.norm(s)
    return dr, ro, uo, out


def hg(cs, g):
    return (1 - g * g) / (1 + g * g - 2 * g * cs) ** 1.5 * 0.25


def test_tile_row_constant_medium(orc):
    N, c, alpha, g = 64, 1.0, 8.0, 0.857
    grid = np.full(N ** 3, c, np.float32)
    sc = orc.Scene(grid, N, c)
    L = np.array([0.3, 0.8, -0.5])
    L /= np.linalg.norm(L)
    sc.update_tr(L.astype(np.float32), alpha)
    sc.set_lut(orc.hg_lut(g, 1))
    v = _tile_view(orc, 5, 4, (2.5, 0.07, -0.04), L, alpha, g, 0.15)
    off_x, off_y = 3, 5
    rows, active = sc.tile_gather(v, off_x, off_y)
    assert active.all()
    ori = np.array(v.ori[:], np.float64)
    h = 1.0 / N
    k2 = math.ceil(TAU2 / (c * alpha * h) - 1e-4 * N)
    w_exp = min((1 - math.exp(-c * alpha * h * (k2 - 1 + 1e-4 * N))) / 0.75, 1.0)
    for p in range(v.width * v.height):
        row, col = divmod(p, v.width)
        r = rows[p]
        dr, ro, uo, sub = _subrays(v, row, col, off_x, off_y)
        assert np.all(np.abs(sub[:, 0]) > 0.9)
        dis = np.array([slab(ori, s, 0.4999)[0] for s in sub])
        tx = dis + TAU1 / (c * alpha) - 1e-4
        ty = dis + TAU2 / (c * alpha) - 1e-4
        # t grid: (mid, range) of [first t, max(tx, ty)] / 3 per sub-ray, then opacities
        idx = 581 + 2 * np.arange(64)
        np.testing.assert_allclose(r[idx], (dis + ty) / 6, rtol=0, atol=2e-6)
        np.testing.assert_allclose(r[idx + 1], (ty - dis) / 3, rtol=0, atol=2e-6)
        np.testing.assert_allclose(r[709:773], w_exp, rtol=4e-6)
        # groups: nearest 0.75-survival hit, or farthest 0.25-survival hit on the checkerboard
        chosen = np.zeros(16, np.int64)
        wsel = np.zeros(16)
        hor = vert = 0.0
        for sid in range(16):
            gx, gy = sid % 4, sid // 4
            cand = [(2 * gy + dy) * 8 + 2 * gx + dx for dy in (0, 1) for dx in (0, 1)]
            do25 = (sid + (gy & 1)) & 1
            k = cand[int(np.argmax(ty[cand]))] if do25 else cand[int(np.argmin(tx[cand]))]
            fill = sid // 2 + 8 * do25
            chosen[fill], wsel[fill] = k, (ty[k] if do25 else tx[k])
            s = float(np.sum(tx[cand] + ty[cand]))
            hor += s if gx >= 2 else -s
            vert += s if gy >= 2 else -s
        np.testing.assert_allclose(r[579:581], [hor, vert], rtol=2e-5, atol=2e-4)
        # basis from the average direction (every sub-ray reaches the 0.25 depth: weight 2)
        X = sub.sum(0) * 2
        X /= np.linalg.norm(X)
        Y = np.cross(X, [0.0, 0.0, 1.0])
        Y /= np.linalg.norm(Y)
        Z = np.cross(Y, X)
        bl = dr + ro * -4 + uo * -4
        rc, uc = dr + ro * 4, dr + uo * 4
        ang = [math.acos(np.clip(np.dot(q / np.linalg.norm(q), L), -1, 1)) for q in (X, bl, rc, uc)]
        np.testing.assert_allclose(r[576:579], [ang[0], ang[2] - ang[1], ang[3] - ang[1]], rtol=0, atol=3e-6)
        # the 16 samples: density log(1 + alpha c / 64), phase of the chosen sub-ray
        np.testing.assert_allclose(r[0:16], math.log(1 + alpha * c / 64), rtol=1e-6)
        np.testing.assert_allclose(r[192:208], np.log(hg(sub[chosen] @ L, g) + 1), rtol=2e-5)
        # stencil points 16..191 around the samples in round-robin order: the predictor's gather
        # of the same point and basis
        pos = ori[None, :] + sub[chosen] * wsel[:, None]
        for k in range(16):
            f = sc.gather_state(1, pos[k], X, Y, Z, L, g, alpha)
            ii = np.arange(16 + k, 192, 16)
            for base in (0, 192, 384):
                np.testing.assert_allclose(r[base + ii], f[base + ii], rtol=2e-4, atol=2e-6)


def test_tile_row_empty_volume(orc):
    N = 32
    sc = orc.Scene(np.zeros(N ** 3, np.float32), N, 0.0)
    L = np.array([0.0, 1.0, 0.0], np.float32)
    sc.update_tr(L, 4.0)
    sc.set_lut(orc.hg_lut(0.857, 1))
    v = _tile_view(orc, 4, 3, (0.67085, -0.03808, -0.04856), L, 4.0, 0.857, 1.0)
    rows, active = sc.tile_gather(v, 0, 0)
    assert not active.any()
    assert np.all(rows[:, 0] == -1000.0) and np.all(rows[:, 1:] == 0.0)


def test_tile_pixel_offset_shifts_subrays(orc):
    """A pixel offset of 8 eighths equals the next pixel (GatherSamples' sub-pixel offset)."""
    N = 32
    grid = smooth_field(N, 6)
    sc = orc.Scene(grid, N, float(grid.max()))
    L = np.array([0.34281, 0.70711, 0.61845])
    L /= np.linalg.norm(L)
    sc.update_tr(L.astype(np.float32), 4.0)
    sc.set_lut(orc.hg_lut(0.857, 1))
    v = _tile_view(orc, 6, 6, (2.0, 0.3, 0.5), L, 4.0, 0.857, 0.25)
    a, _ = sc.tile_gather(v, 8, 0)
    b, _ = sc.tile_gather(v, 0, 0)
    a, b = a.reshape(6, 6, -1), b.reshape(6, 6, -1)
    np.testing.assert_array_equal(a[:, :-1], b[:, 1:])
    assert not np.array_equal(a[:, -1], b[:, -1])
