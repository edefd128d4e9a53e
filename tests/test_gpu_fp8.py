"""FP8 with per-layer max-abs activation scales on the GPU.

* Calibration statistic: mrp_fp8_calibrate (BF16 network) records per layer the largest |input|;
  it agrees with the oracle's FP64 network (predict_amax_f64, pinned against the numpy
  restatement) within 3 % per layer, the BF16 network's activation rounding.
* Exponents: each installed exponent is the largest e with max * 2^e <= 448.
* Accumulation: maxima recorded over two calls without a reset equal the elementwise maximum
  of the two calls made separately, bit for bit.
* Error: FP8 with the shipped exponents and with exponents calibrated on the rows themselves,
  against the FP64 network (mean relative error bound stated in the test).
* Clamp: exponents far above the calibrated ones saturate every input at 448 instead of
  producing NaN; the outputs stay finite.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 0.857
SR = [0.99 / 1.001] * 3


def exp_ok(m, e):
    """e is the largest integer with m * 2^e <= 448 (exact float arithmetic)."""
    if m == 0.0:
        return e == 0
    return math.ldexp(m, e) <= 448.0 < math.ldexp(m, e + 1)


@pytest.fixture(scope="module")
def rows(ctx):
    import torch

    import mrpnn_amd as M
    vol = M.Volume(ctx, 64)
    vol.fill_fbm(5, 4.0, 3)
    vol.update()
    v = M.orbit_view(96, 80, azimuth=0.7, elevation=0.3, light=(0.5, 0.8, 0.3), alpha=64.0, g=G)
    ctx.prepare_views([vol], [v])
    vol.update_tr(np.asarray(v.light_dir[:], np.float32), v.alpha)
    feats, active, _, _ = vol.gather_features(v)
    f = feats[active > 0].contiguous()
    torch.cuda.synchronize()
    assert f.shape[0] > 500
    saved = ctx.fp8_exponents()
    yield f
    ctx.set_fp8_exponents(saved)
    vol.close()


def test_calibration_maxima_match_fp64(ctx, orc, rows):
    m = ctx.fp8_calibrate(rows, SR, G, reset=True)
    ref = orc.predict_amax_f64(rows.cpu().numpy().astype(np.float64), np.asarray(SR, np.float64), G)
    assert (ref > 0).all()
    rel = np.abs(m - ref) / ref
    assert rel.max() < 0.03, (int(rel.argmax()), float(rel.max()))
    exps = ctx.fp8_exponents()
    assert all(exp_ok(float(x), int(e)) for x, e in zip(m, exps))


def test_calibration_accumulates(ctx, rows):
    half = rows.shape[0] // 2
    a = ctx.fp8_calibrate(rows[:half].contiguous(), SR, G, reset=True)
    b = ctx.fp8_calibrate(rows[half:].contiguous(), SR, G, reset=True)
    ctx.fp8_calibrate(rows[:half].contiguous(), SR, G, reset=True)
    both = ctx.fp8_calibrate(rows[half:].contiguous(), SR, G, reset=False)
    assert np.array_equal(both, np.maximum(a, b))
    again = ctx.fp8_calibrate(None, SR, G, reset=False)  # no rows: re-derives the same exponents
    assert np.array_equal(again, both)


def test_set_get_and_refusals(ctx, rows):
    import mrpnn_amd as M
    saved = ctx.fp8_exponents()
    e = np.arange(60, dtype=np.int32) % 7 - 3
    ctx.set_fp8_exponents(e)
    assert np.array_equal(ctx.fp8_exponents(), e)
    with pytest.raises(M.MrpError):
        ctx.set_fp8_exponents(e[:59])
    bad = e.copy()
    bad[5] = 101
    with pytest.raises(M.MrpError):
        ctx.set_fp8_exponents(bad)
    assert np.array_equal(ctx.fp8_exponents(), e)  # refused calls change nothing
    ctx.set_fp8_exponents(saved)


def _fp8_error(ctx, orc, rows, n=2048):
    import mrpnn_amd as M
    f = rows[:n].contiguous()
    out = ctx.predict_features(f, SR, G, M.FP8).cpu().numpy()[:, 0].astype(np.float64)
    ref = orc.predict_f64(f.cpu().numpy().astype(np.float64), np.asarray(SR, np.float64), G)[:, 0]
    assert np.all(np.isfinite(out))
    return np.abs(out - ref) / np.maximum(np.abs(ref), 1e-12)


def test_fp8_error_shipped_and_calibrated(ctx, orc, rows):
    """Mean relative error vs the FP64 network <= 15 % with the exponents that ship with the
    library and with exponents calibrated on these rows."""
    saved = ctx.fp8_exponents()
    import mrpnn_amd as M
    c0 = M.Context()  # a fresh context holds the shipped exponents
    try:
        shipped = c0.fp8_exponents()
    finally:
        c0.close()
    ctx.set_fp8_exponents(shipped)
    e_ship = _fp8_error(ctx, orc, rows)
    ctx.fp8_calibrate(rows, SR, G, reset=True)
    e_cal = _fp8_error(ctx, orc, rows)
    ctx.set_fp8_exponents(saved)
    assert e_ship.mean() < 0.15 and e_cal.mean() < 0.15, (e_ship.mean(), e_cal.mean())


def test_fp8_clamp_keeps_outputs_finite(ctx, orc, rows):
    import mrpnn_amd as M
    saved = ctx.fp8_exponents()
    try:
        ctx.set_fp8_exponents(np.full(60, 40, np.int32))  # clamp at 448 * 2^-40: every input saturates
        out = ctx.predict_features(rows[:1024].contiguous(), SR, G, M.FP8).cpu().numpy()
        assert np.all(np.isfinite(out)) and np.all(out >= 0)
    finally:
        ctx.set_fp8_exponents(saved)
