"""The oracle's RadiancePredictInternal (FP32 and FP64) pinned against an independent
numpy restatement and against invariants of the network."""
import numpy as np
import pytest

import mlp_numpy

G = 0.857


@pytest.fixture(scope="module")
def feats(orc, small_scene):
    sc, L = small_scene
    sc.set_lut(orc.hg_lut(G, 1))
    v = orc.orbit_view(24, 20)
    r = sc.render_predict(v, want_feats=True)
    f = r["feats"][r["active"] > 0]
    assert f.shape[0] > 50
    rng = np.random.default_rng(3)
    # plus random rows in the feature ranges: X_Val in [0,1.2], X_Val_Hg in [0,3], X_Val_Sub in [0,1]
    rnd = np.concatenate([rng.uniform(0, 1.2, (40, 192)), rng.uniform(0, 3, (40, 192)), rng.uniform(0, 1, (40, 192)),
                          rng.uniform(0, np.pi, (40, 1))], axis=1).astype(np.float32)
    return np.concatenate([f, rnd])


@pytest.fixture(scope="module")
def w():
    return mlp_numpy.load_weights()


def test_fp64_matches_numpy_restatement(orc, feats, w):
    sr = np.array([0.99 / 1.001] * 3)
    a = orc.predict_f64(feats.astype(np.float64), sr, G)
    b = mlp_numpy.predict(w, feats, G, sr)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-14)


def test_distinct_channels_match_numpy(orc, feats, w):
    sr = np.array([0.97, 0.93, 0.99]) / 1.001
    a = orc.predict_f64(feats[:30].astype(np.float64), sr, G)
    b = mlp_numpy.predict(w, feats[:30], G, sr)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-14)
    assert not np.allclose(a[:, 0], a[:, 1])


def test_fp32_close_to_fp64(orc, feats):
    sr = np.array([0.99 / 1.001] * 3, np.float32)
    a = orc.predict_f32(feats, sr, G).astype(np.float64)
    b = orc.predict_f64(feats.astype(np.float64), sr.astype(np.float64), G)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-12)
    assert rel.max() < 2e-4 and np.median(rel) < 1e-5


def test_equal_channels_and_nonnegative(orc, feats):
    out = orc.predict_f32(feats, np.array([0.95] * 3, np.float32), G)
    assert (out >= 0).all()
    assert np.array_equal(out[:, 0], out[:, 1]) and np.array_equal(out[:, 0], out[:, 2])
    # z equal to y but not x: Z is taken from Y (radiancePredict.cu:417-419)
    out2 = orc.predict_f32(feats[:10], np.array([0.9, 0.95, 0.95], np.float32), G)
    assert np.array_equal(out2[:, 1], out2[:, 2]) and not np.array_equal(out2[:, 0], out2[:, 1])


def test_zero_scatter_rate_gives_zero(orc, feats):
    out = orc.predict_f64(feats[:10].astype(np.float64), np.zeros(3), G)
    assert np.all(out == 0.0)


def test_inactive_rows_constant(orc):
    # all-zero features (inactive lane) -> same output for the same gamma
    z = np.zeros((4, 577), np.float32)
    z[:, 576] = 1.1
    out = orc.predict_f32(z, np.array([0.99 / 1.001] * 3, np.float32), G)
    assert np.all(out == out[0])


def fp8_layer_names():
    """The FP8 layers of the library's C ABI in its order (MRP_FP8_LAYERS, mrp_c_api.h)."""
    names = ["%s%d1W" % (b, k) for b in ("L", "L_Tr", "L_Hg") for k in range(9)]
    names += ["%s%d1W" % (b, k) for b in ("LD", "LD_Tr", "LD_Hg") for k in range(4)]
    names += ["LD41W", "LD_Tr41W", "LD_Hg41W", "LSEFin2W", "LC0W", "LC1W", "LC2W", "LXW"]
    return names + ["LX%dW" % k for k in range(13)]


@pytest.mark.parametrize("sr", [[0.99 / 1.001] * 3, [0.97 / 1.001, 0.93 / 1.001, 0.99 / 1.001]])
def test_layer_input_maxima_match_numpy(orc, feats, w, sr):
    """The FP8 calibration statistic (largest |input| of every matrix-product layer) of the
    oracle's FP64 network equals the numpy restatement's, layer by layer."""
    sr = np.asarray(sr, np.float64)
    a = orc.predict_amax_f64(feats.astype(np.float64), sr, G)
    d = {}
    mlp_numpy.predict(w, feats, G, sr, amax=d)
    names = fp8_layer_names()
    assert len(names) == 60 and len(set(names)) == 60
    b = np.array([d[n] for n in names])
    np.testing.assert_allclose(a, b, rtol=1e-12)
    assert (a > 0).all()
