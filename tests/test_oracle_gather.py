"""The oracle's MRPNN feature gather (SetupInputData, core/radiancePredict.cu:427-485, with
GetSamples23_ / SphereRandom3, core/vector.cu:265-405, MipDensityDynamic / ShadowTerm_TRTex,
core/renderResources.cu:32-102, HenyeyGreenstein, core/vector.cu:210-214) pinned by closed forms.

The oracle's `gather_features` runs on hand-built pyramids whose texture lookups have float64
closed forms, so each of the 577 outputs is predicted independently:

* density mip m constant c_m (distinct per level): a zero-border trilinear lookup of a constant
  is c_m times the product over axes of the weight that falls on in-grid texels, so X_Val
  reveals the ring's mip, the alpha/64 scale and the border semantics;
* density mips linear in x, y and z (distinct slopes): trilinear filtering reproduces a linear
  field exactly away from the border, so X_Val reveals every stencil position
  pos + (X L.x + Y L.y + Z L.z) offset and pos + LightDir offset;
* TR mip k constant 0.1 (k + 1), or linear: a clamp-addressed lookup of a linear field is the
  field at the per-axis clamped coordinate, so X_Val_Sub reveals the TR mip max(mip - 1, 0),
  the lookup position, and for points outside the box the entry point along the light
  (slab intersection with the 0.4999 box of RayBoxOffset) or 1 when the light ray misses;
* a LUT linear in u and in v with different slopes: X_Val_Hg reveals u = cos * 0.5 + 0.5 for
  XMain and LXMain against the direction to the point and v = atan(0.5 2^mip / 256 / offset) / (pi/3);
* X_Val_Hg[0] = ln(HG(XMain . L, g) + 1), gamma = acos(XMain . LXMain);
* through the oracle's own pyramid and TR bake: a constant 256^3 volume (X_Val closed form on
  every ring), an empty volume (X_Val = 0, X_Val_Sub = 1), an inactive pixel (all zero).

A wrong mip or TR level, a swapped LUT axis, a ring offset, sign or basis mix-up changes a
prediction by far more than the float32 tolerance.  The stencil directions are the data table
generated from core/vector.cu:82-177 (checked against that file by tests/test_assets.py); the
ring schedule below restates core/vector.cu:304-405.
"""
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT

PC = 192
ALPHA = 48.0
G = 0.857
RTOL = 3e-6


def stencil_dirs():
    txt = open(os.path.join(ROOT, "mrpnn_amd", "csrc", "ref_tables.h")).read()
    body = txt[txt.index("kStencilDirs[160][3]"):]
    body = body[:body.index("};")]
    vals = [float(x) for x in re.findall(r"(-?\d+\.\d+)f", body)]
    assert len(vals) == 480
    return np.array(vals, np.float64).reshape(160, 3)


def schedule():
    """GetSamples23_ (core/vector.cu:304-405): density mip, offset (cube units), sphere flag."""
    mip = np.zeros(PC, int)
    off = np.zeros(PC)
    sphere = np.zeros(PC, bool)
    rings = [(0, 8, 0, 1.0 * 1.0), (8, 16, 1, 2.5 * 0.92), (16, 32, 2, 5.5 * 0.84), (32, 48, 3, 11.5 * 0.76),
             (48, 64, 4, 23.5 * 0.68), (64, 96, 5, 47.5 * 0.6), (96, 128, 6, 95.5 * 0.6), (128, 160, 7, 191.5 * 0.6)]
    for a, b, m, o in rings:
        mip[a:b], off[a:b], sphere[a:b] = m, o / 256.0, True
    for k in range(8):
        mip[160 + k], off[160 + k] = 0, (k + 1) / 256.0
        mip[168 + k], off[168 + k] = 1, (9.5 + 2 * k) / 256.0
        mip[176 + k], off[176 + k] = 2, (26.5 + 4 * k) / 256.0
        mip[184 + k], off[184 + k] = 3, (60.5 + 8 * k) / 256.0
    return mip, off, sphere


def positions(pos, X, Y, Z, L):
    D = stencil_dirs()
    mip, off, sphere = schedule()
    p = np.zeros((PC, 3))
    d = D @ np.stack([X, Y, Z]).astype(np.float64)  # X L.x + Y L.y + Z L.z
    p[:160] = pos + d * off[:160, None]
    p[160:] = pos + np.asarray(L, np.float64)[None, :] * off[160:, None]
    return p, mip, off, sphere


def border_weight(uv, r):
    """Per axis, the share of linear-filter weight on in-grid texels of a resolution-r axis."""
    u = uv * r - 0.5
    x0 = np.floor(u)
    w = u - x0
    return (1 - w) * ((x0 >= 0) & (x0 <= r - 1)) + w * ((x0 + 1 >= 0) & (x0 + 1 <= r - 1))


def const_mip_value(c, p, m):
    r = 256 >> m
    return c * np.prod(border_weight(p + 0.5, r), axis=-1)


def interior(p, r):
    """Every texel of the linear-filter footprint of p is in-grid (resolution r)."""
    u = (p + 0.5) * r - 0.5
    return np.all((np.floor(u) >= 0) & (np.floor(u) + 1 <= r - 1), axis=-1)


def box_entry(p, d, h=0.4999):
    """RayBoxOffset (core/vector.cu:46-64): distance along d to the 0.4999 box, -1 on a miss."""
    with np.errstate(divide="ignore"):
        inv = 1.0 / d
    t0, t1 = (-h - p) * inv, (h - p) * inv
    tmi = np.max(np.minimum(t0, t1))
    tma = np.min(np.maximum(t0, t1))
    return max(tmi, 0.0) if tma >= tmi else -1.0


def hg(c, g):
    return 0.25 * (1 - g * g) / (1 + g * g - 2 * g * c) ** 1.5


def cube_texel_centres(r):
    c = (np.arange(r) + 0.5) / r - 0.5
    return np.meshgrid(c, c, c, indexing="ij")  # index (x, y, z) <-> position (x, y, z)


def concat_mips(levels, n):
    return np.concatenate([np.asarray(levels[m], np.float32).ravel() for m in range(n)])


def lut_linear(a, b, c):
    t = (np.arange(512) + 0.5) / 512
    return (a + b * t[None, :] + c * t[:, None]).astype(np.float32)  # lut[j*512 + i]: i <-> u, j <-> v


def lut_value(a, b, c, u, v):
    lo, hi = 0.5 / 512, 511.5 / 512
    return a + b * np.clip(u, lo, hi) + c * np.clip(v, lo, hi)


def frame():
    X = np.array([0.3, -0.2, 0.93])
    X /= np.linalg.norm(X)
    Y = np.cross(X, np.array([0.1, 1.0, 0.05]))
    Y /= np.linalg.norm(Y)
    Z = np.cross(Y, X)
    return X.astype(np.float32), Y.astype(np.float32), Z.astype(np.float32)


def run(orc, sc, active, pos, X, Y, Z, L, g=G, alpha=ALPHA):
    f = sc.gather_state(active, pos, X, Y, Z, L, g, alpha).astype(np.float64)
    return f[:PC], f[PC:2 * PC], f[2 * PC:3 * PC], f[3 * PC]


def expected_hg(p, pos, X, LX, mip, off, a, b, c):
    ms = p - pos
    ms[0] = X  # the centre point (index 0) takes the HG branch; keep its row finite
    ms /= np.linalg.norm(ms, axis=1, keepdims=True)
    u0 = (ms @ X.astype(np.float64)) * 0.5 + 0.5
    u1 = (ms @ LX.astype(np.float64)) * 0.5 + 0.5
    v = np.arctan(0.5 * (2.0 ** mip) / 256.0 / off) / (math.pi / 3)
    return np.log(lut_value(a, b, c, u0, v) * lut_value(a, b, c, u1, v) + 1.0)


@pytest.fixture(scope="module")
def synthetic(orc):
    """Hand-built pyramids: per-level constants (case 'const') and per-level linear fields
    (case 'linear'), TR levels likewise, and a LUT linear in (u, v)."""
    out = {}
    cm = [(m + 1) * 0.01 for m in range(9)]
    const_mips = [np.full((256 >> m,) * 3, cm[m]) for m in range(9)]
    const_tr = [np.full((128 >> k,) * 3, 0.1 * (k + 1)) for k in range(8)]
    slope = np.array([0.3, -0.2, 0.1])
    lin_mips = []
    for m in range(9):
        x, y, z = cube_texel_centres(256 >> m)
        lin_mips.append(0.5 + 0.05 * m + slope[0] * x + slope[1] * y + slope[2] * z)
    tq = np.array([-0.25, 0.15, 0.35])
    lin_tr = []
    for k in range(8):
        x, y, z = cube_texel_centres(128 >> k)
        lin_tr.append(0.4 + 0.07 * k + tq[0] * x + tq[1] * y + tq[2] * z)
    lut_abc = (0.2, 0.9, -0.6)
    lut = lut_linear(*lut_abc)
    grid = np.zeros(8 ** 3, np.float32)  # the gather reads mips only
    out["const"] = orc.Scene.from_arrays(grid, 8, 1.0, concat_mips(const_mips, 9), concat_mips(const_tr, 8), lut)
    out["linear"] = orc.Scene.from_arrays(grid, 8, 1.0, concat_mips(lin_mips, 9), concat_mips(lin_tr, 8), lut)
    out["cm"], out["slope"], out["tq"], out["lut_abc"] = cm, slope, tq, lut_abc
    return out


def test_schedule_matches_stencil_accessor(orc):
    mip, off, sphere = schedule()
    for i in range(PC):
        o, m, t, li = orc.stencil(i)
        assert m == mip[i] and (t >= 4) == sphere[i] and (li == 0) == (i == 0), i
        assert abs(o - off[i]) <= 1e-7 * off[i], i


def test_const_levels_pin_ring_mip_tr_mip_and_lut(orc, synthetic):
    """Per-level constants: X_Val names the density mip of every point (zero border included),
    X_Val_Sub the TR mip max(mip - 1, 0); the linear LUT pins u and v of X_Val_Hg."""
    sc = synthetic["const"]
    X, Y, Z = frame()
    L = np.array([0.5, 0.8, -0.33])
    L = (L / np.linalg.norm(L)).astype(np.float32)
    pos = np.array([0.01, -0.02, 0.015], np.float32)
    xv, xh, xs, gamma = run(orc, sc, 1, pos, X, Y, Z, L)
    p, mip, off, sphere = positions(pos.astype(np.float64), X, Y, Z, L)
    assert np.all(np.abs(p) < 0.5 - 1e-3)  # every point inside the box: TR looked up in place
    cm = synthetic["cm"]
    rho = np.array([const_mip_value(cm[mip[i]], p[i], mip[i]) for i in range(PC)])
    np.testing.assert_allclose(xv, np.log1p(ALPHA * rho / 64.0), rtol=RTOL, atol=1e-7)
    assert np.any(rho[128:160] < np.array(cm)[7] * 0.999)  # the border weight is exercised on ring 7
    np.testing.assert_allclose(xs, 0.1 * (np.maximum(mip - 1, 0) + 1), rtol=RTOL)
    exp_h = expected_hg(p, pos.astype(np.float64), X, L, mip, off, *synthetic["lut_abc"])
    exp_h[0] = math.log(hg(float(X.astype(np.float64) @ L.astype(np.float64)), G) + 1.0)
    np.testing.assert_allclose(xh, exp_h, rtol=RTOL, atol=2e-7)
    np.testing.assert_allclose(gamma, math.acos(float(X.astype(np.float64) @ L.astype(np.float64))), rtol=1e-6)


def test_linear_levels_pin_stencil_positions(orc, synthetic):
    """Fields linear in x, y, z with distinct slopes: X_Val and X_Val_Sub are the fields at
    each stencil point, which fixes the sphere positions pos + (X L.x + Y L.y + Z L.z) offset
    and the light points pos + LightDir offset."""
    sc = synthetic["linear"]
    X, Y, Z = frame()
    L = np.array([-0.4, 0.7, 0.59])
    L = (L / np.linalg.norm(L)).astype(np.float32)
    pos = np.array([-0.03, 0.025, 0.04], np.float32)
    xv, xh, xs, _ = run(orc, sc, 1, pos, X, Y, Z, L)
    p, mip, off, sphere = positions(pos.astype(np.float64), X, Y, Z, L)
    s, tq = synthetic["slope"], synthetic["tq"]
    rho = 0.5 + 0.05 * mip + p @ s
    ok = np.array([interior(p[i], 256 >> mip[i]) for i in range(PC)])
    assert ok[:128].all() and ok[160:].all()  # rings 0-6 and the light rings are away from the border
    np.testing.assert_allclose(xv[ok], np.log1p(ALPHA * rho[ok] / 64.0), rtol=RTOL)
    k = np.maximum(mip - 1, 0)
    tr = np.array([0.4 + 0.07 * k[i] + tq @ np.clip(p[i] + 0.5, 0.5 / (128 >> k[i]), 1 - 0.5 / (128 >> k[i])) - tq @ np.full(3, 0.5)
                   for i in range(PC)])
    np.testing.assert_allclose(xs, tr, rtol=RTOL, atol=1e-7)
    # the LUT term depends on the direction to each point only
    exp_h = expected_hg(p, pos.astype(np.float64), X, L, mip, off, *synthetic["lut_abc"])
    np.testing.assert_allclose(xh[1:], exp_h[1:], rtol=RTOL, atol=2e-7)


def test_ring7_border_against_brute_force(orc, synthetic):
    """Ring 7 samples the 2^3 mip 7 beyond its texel centres: zero-border trilinear by brute force."""
    from test_oracle_volume import trilinear_zero_border_f64
    sc = synthetic["linear"]
    X, Y, Z = frame()
    L = np.array([0.2, 0.9, 0.1], np.float32)
    L /= np.linalg.norm(L)
    pos = np.array([0.0, 0.01, -0.01], np.float32)
    xv, _, _, _ = run(orc, sc, 1, pos, X, Y, Z, L)
    p, mip, _, _ = positions(pos.astype(np.float64), X, Y, Z, L)
    x, y, z = cube_texel_centres(2)
    m7 = 0.5 + 0.05 * 7 + synthetic["slope"] @ np.stack([x, y, z]).reshape(3, -1)
    m7 = m7.reshape(2, 2, 2)
    rho = np.array([trilinear_zero_border_f64(m7, p[i] + 0.5) for i in range(128, 160)])
    np.testing.assert_allclose(xv[128:160], np.log1p(ALPHA * rho / 64.0), rtol=RTOL, atol=1e-7)


def test_points_outside_the_box_use_the_light_entry(orc, synthetic):
    """ShadowTerm_TRTex (core/renderResources.cu:84-102): a stencil point outside the cube reads
    the TR mip where the light ray from it enters the 0.4999 box, or 1 when that ray misses."""
    sc = synthetic["linear"]
    X, Y, Z = frame()
    L = np.array([-0.7, 0.5, 0.2])
    L = (L / np.linalg.norm(L)).astype(np.float32)
    pos = np.array([0.43, 0.18, -0.05], np.float32)
    xv, _, xs, _ = run(orc, sc, 1, pos, X, Y, Z, L)
    p, mip, _, _ = positions(pos.astype(np.float64), X, Y, Z, L)
    tq = synthetic["tq"]
    k = np.maximum(mip - 1, 0)
    n_entry = n_miss = 0
    reach = np.max(np.abs(p), axis=1)
    assert np.all(np.abs(reach - 0.5) > 1e-4)  # no point on a face: inside / outside is unambiguous
    for i in range(PC):
        r = 128 >> k[i]
        if reach[i] < 0.5:
            exp = 0.4 + 0.07 * k[i] + tq @ (np.clip(p[i] + 0.5, 0.5 / r, 1 - 0.5 / r) - 0.5)
            assert abs(xs[i] - exp) <= 2e-6 * max(1.0, abs(exp)), (i, xs[i], exp)
            continue
        t = box_entry(p[i], L.astype(np.float64))
        if t < 0:
            exp = 1.0
            n_miss += 1
        else:
            e = p[i] + L.astype(np.float64) * t
            exp = 0.4 + 0.07 * k[i] + tq @ (np.clip(e + 0.5, 0.5 / r, 1 - 0.5 / r) - 0.5)
            n_entry += 1
        assert abs(xs[i] - exp) <= 2e-6 * max(1.0, abs(exp)), (i, xs[i], exp)
    assert n_entry >= 10 and n_miss >= 3, (n_entry, n_miss)
    # points far outside the density mips read zero density
    far = np.array([np.any(np.abs(p[i]) > 0.5 + 0.5 / (256 >> mip[i]) + 1e-6) for i in range(PC)])
    assert far.sum() >= 5 and np.all(xv[far] == 0.0)


def test_constant_volume_through_the_pipeline(orc):
    """A constant 256^3 grid: every pyramid level is the constant, so X_Val is
    ln(1 + alpha c H / 64) with H the zero-border weight product of the point's mip."""
    N, c = 256, 0.7
    sc = orc.Scene(np.full(N ** 3, c, np.float32), N)
    L = np.array([0.3, 0.9, 0.3], np.float32)
    L /= np.linalg.norm(L)
    sc.update_tr(L, ALPHA)
    sc.set_lut(lut_linear(0.5, 0.1, 0.2))
    X, Y, Z = frame()
    pos = np.array([0.1, -0.12, 0.05], np.float32)
    xv, _, xs, _ = run(orc, sc, 1, pos, X, Y, Z, L)
    p, mip, _, _ = positions(pos.astype(np.float64), X, Y, Z, L)
    rho = np.array([const_mip_value(c, p[i], mip[i]) for i in range(PC)])
    np.testing.assert_allclose(xv, np.log1p(ALPHA * rho / 64.0), rtol=RTOL, atol=1e-7)
    inner = np.array([interior(p[i], 256 >> mip[i]) for i in range(PC)])
    assert inner.sum() > 150
    np.testing.assert_allclose(xv[inner], math.log1p(ALPHA * c / 64.0), rtol=RTOL)
    assert np.all((xs > 0) & (xs <= 1))


def test_empty_volume_and_inactive_pixel(orc):
    """Empty grid: no density anywhere (X_Val = 0) and full transmittance (X_Val_Sub = 1, in the
    box and through the entry fallback).  Inactive pixel: all 576 features zero."""
    N = 32
    sc = orc.Scene(np.zeros(N ** 3, np.float32), N)
    L = np.array([-0.7, 0.5, 0.2], np.float32)
    L /= np.linalg.norm(L)
    sc.update_tr(L, ALPHA)
    sc.set_lut(lut_linear(0.5, 0.1, 0.2))
    X, Y, Z = frame()
    for pos in (np.array([0.0, 0.0, 0.0], np.float32), np.array([0.43, 0.18, -0.05], np.float32)):
        xv, xh, xs, gamma = run(orc, sc, 1, pos, X, Y, Z, L)
        assert np.all(xv == 0.0) and np.all(xs == 1.0)
        assert np.all(np.isfinite(xh)) and np.all(xh > 0)
    xv, xh, xs, gamma = run(orc, sc, 0, pos, X, Y, Z, L)
    assert np.all(xv == 0.0) and np.all(xh == 0.0) and np.all(xs == 0.0)
    np.testing.assert_allclose(gamma, math.acos(float(X.astype(np.float64) @ L.astype(np.float64))), rtol=1e-6)
