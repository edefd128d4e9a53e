"""CPU model of the LDS-staged level-0 TR bake (k_tr_bake_l0, mrpnn_amd/csrc/volume_kernels.hip).

The kernel stages, per tile of light-sheared coordinates, one window of density mip 1 per padded
slab and reads every trilinear tap of the tile's rays from those windows.  Its correctness rests
on three facts, checked here with a float32 emulation of the kernel's own arithmetic (running
position sum, padded coordinates, tile test, window origins) for several light directions:
  * the tile test partitions the 128^3 texels (each texel in exactly one tile, and among that
    tile's candidates);
  * every tap of every sample lies inside its slab's window, both taps of each axis;
  * every tap's slab lies in the slab range the workgroup loads (the tile's layers towards the
    exit face, cut to the slabs where the tile's window meets the padded level), and no tile
    holds more texels than the kernel's list.
The constants mirror the kernel's (tile 4, window 10, 130 slabs, list 128 * 5 * 5).
"""
import numpy as np
import pytest

T, W, SL, LIST = 4, 10, 130, 128 * 25
f32 = np.float32


def fma32(a, b, c):  # a*b + c rounded once (exact product of float32 operands in float64)
    return (np.float64(a) * np.float64(b) + np.float64(c)).astype(np.float32)


def model(L, stride, slope_margin=True):
    L = np.asarray(L, np.float32)
    A = int(np.argmax(np.abs(L)))
    U = 1 if A == 0 else 0
    V = 1 if A == 2 else 2
    ru, rv = f32(L[U] / L[A]), f32(L[V] / L[A])
    lo, hi = [1e30, 1e30], [-1e30, -1e30]
    for ca in (f32(1), f32(128)):
        for cw in (f32(1), f32(128)):
            for j, r in enumerate((ru, rv)):
                w = f32(cw - f32(r * ca))
                lo[j], hi[j] = min(lo[j], w), max(hi[j], w)
    umin, vmin = f32(np.floor(lo[0]) - 1), f32(np.floor(lo[1]) - 1)
    ntu = int((f32(hi[0] + 1) - umin) / T) + 1
    ntv = int((f32(hi[1] + 1) - vmin) / T) + 1
    idx = np.arange(128)
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], -1)
    ca = (P[:, A] + 1).astype(np.float32)
    up = fma32(-ru, ca, (P[:, U] + 1).astype(np.float32))
    vp = fma32(-rv, ca, (P[:, V] + 1).astype(np.float32))
    tu = np.floor((up - umin) * f32(1.0 / T)).astype(int)
    tv = np.floor((vp - vmin) * f32(1.0 / T)).astype(int)
    assert tu.min() >= 0 and tu.max() < ntu and tv.min() >= 0 and tv.max() < ntv
    ulo = (umin + f32(T) * tu.astype(np.float32)).astype(np.float32)
    vlo = (vmin + f32(T) * tv.astype(np.float32)).astype(np.float32)
    du = P[:, U] - (np.floor(fma32(ru, ca, ulo)).astype(int) - 2)
    dv = P[:, V] - (np.floor(fma32(rv, ca, vlo)).astype(int) - 2)
    assert du.min() >= 0 and du.max() < T + 3 and dv.min() >= 0 and dv.max() < T + 3  # among the candidates
    tid = tu * ntv + tv
    assert np.bincount(tid, minlength=ntu * ntv).max() <= LIST
    lay = P[:, A]
    lo_l = np.full(ntu * ntv, 1 << 30)
    hi_l = np.full(ntu * ntv, -1)
    np.minimum.at(lo_l, tid, lay)
    np.maximum.at(hi_l, tid, lay)
    upward = L[A] > 0
    s_lo = np.where(upward, np.maximum(lo_l[tid] - 1, 0), 0)
    s_hi = np.where(upward, SL - 1, np.minimum(hi_l[tid] + 3, SL - 1))
    sl = np.arange(SL).astype(np.float32)

    def origin_table(lo_, r):
        m = f32(abs(r)) if slope_margin else f32(0)
        return np.floor(((lo_[:, None] + (r * sl)[None, :]).astype(np.float32) - m).astype(np.float32)
                        - f32(0.01)).astype(int)
    OU = origin_table((umin + f32(T) * np.arange(ntu).astype(np.float32)).astype(np.float32), ru)
    OV = origin_table((vmin + f32(T) * np.arange(ntv).astype(np.float32)).astype(np.float32), rv)
    meet = ((OU + W > 0) & (OU < 131))[:, None, :] & ((OV + W > 0) & (OV < 131))[None, :, :]
    first = np.where(meet.any(2), meet.argmax(2), SL).reshape(-1)
    last = np.where(meet.any(2), SL - 1 - meet[:, :, ::-1].argmax(2), -1).reshape(-1)
    s_lo = np.maximum(s_lo, first[tid])
    s_hi = np.minimum(s_hi, last[tid])
    # the samples of a subset of texels
    sel = np.arange(0, len(P), stride)
    P, tuS, tvS, s_lo, s_hi = P[sel], tu[sel], tv[sel], s_lo[sel], s_hi[sel]
    uv = ((P.astype(np.float32) + f32(0.5)) / f32(128)).astype(np.float32)
    ori = (uv - f32(0.5)).astype(np.float32)
    inv = np.where(L == 0, f32(999999999.0), f32(1) / np.where(L == 0, f32(1), L)).astype(np.float32)
    to = ((-ori) * inv).astype(np.float32)
    a2f = (f32(0.5) * inv).astype(np.float32)
    dis = np.min(np.maximum((to + a2f).astype(np.float32), (to - a2f).astype(np.float32)), axis=1).astype(np.float32)
    st = (dis / f32(128)).astype(np.float32)
    Ls = (L[None, :] * st[:, None]).astype(np.float32)
    lp = ori.copy()
    worst = [99, -99]
    for _ in range(128):
        lp = (lp + Ls).astype(np.float32)
        c = fma32(lp, f32(128), f32(64.5))
        assert c.min() >= 0
        f = c.astype(int)
        s0 = f[:, A]
        assert (s0 >= s_lo).all() and (s0 + 1 <= s_hi).all()
        for d in (0, 1):
            ou = OU[tuS, s0 + d]
            ov = OV[tvS, s0 + d]
            ou_off, ov_off = f[:, U] - ou, f[:, V] - ov
            worst = [min(worst[0], ou_off.min(), ov_off.min()), max(worst[1], ou_off.max(), ov_off.max())]
    assert worst[0] >= 0 and worst[1] <= W - 2, worst  # both taps of each axis inside the window


LIGHTS = [(0.671, 0.206, 0.712), (-0.116, 0.68, 0.724), (0.577, 0.577, 0.577), (0.0, 1.0, 0.0),
          (0.992, 0.067, -0.105), (-0.588, 0.725, -0.359), (0.3, -0.9, 0.2)]


@pytest.mark.parametrize("light", LIGHTS)
def test_tr_bake_l0_windows_cover_every_tap(light):
    L = np.asarray(light, np.float64)
    model(L / np.linalg.norm(L), stride=7)


def test_tr_bake_l0_model_detects_a_window_without_the_slope_margin():
    # the window origin must reach back by the ray's slope over the slab pair a blend touches;
    # without that term taps fall before the window's first column
    L = np.array([0.671, 0.206, 0.712])
    with pytest.raises(AssertionError):
        model(L / np.linalg.norm(L), stride=7, slope_margin=False)
