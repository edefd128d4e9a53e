"""exp_f32 (mrpnn_amd/csrc/fexp.h), the single-precision exp of the FP32 networks on the device and
in the oracle, against a float64 exp: within 2 ulp of the correctly rounded value on the range
the networks use and beyond, the limits right, and the value at integers and zero exact or
within an ulp (Cody-Waite reduction with a degree-7 Taylor polynomial; float operations only)."""
import numpy as np


def _ulps(got, x):
    want = np.exp(x.astype(np.float64))
    w32 = want.astype(np.float32)
    ulp = (np.nextafter(w32, np.float32(np.inf)) - w32).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / ulp


def test_exp_f32_within_two_ulp(orc):
    x = np.linspace(-87.0, 88.0, 2000001, dtype=np.float32)
    rng = np.random.default_rng(3)
    x = np.concatenate([x, rng.uniform(-20, 20, 500000).astype(np.float32),
                        rng.normal(0, 1e-3, 100000).astype(np.float32)])
    e = _ulps(orc.exp_f32(x), x)
    assert e.max() <= 2.0, (e.max(), x[np.argmax(e)])
    # almost always correctly rounded
    assert np.mean(orc.exp_f32(x) == np.exp(x.astype(np.float64)).astype(np.float32)) > 0.9


def test_exp_f32_limits_and_exact_points(orc):
    x = np.array([0.0, -0.0, 1.0, 89.0, 1e30, -104.0, -1e30, np.inf, -np.inf], np.float32)
    y = orc.exp_f32(x)
    assert y[0] == 1.0 and y[1] == 1.0
    assert abs(y[2] - np.float32(np.e)) <= np.spacing(np.float32(np.e))
    assert np.isinf(y[3]) and np.isinf(y[4]) and np.isinf(y[7])
    assert y[5] == 0.0 and y[6] == 0.0 and y[8] == 0.0
    # gradual underflow: subnormal results within one subnormal step
    xs = np.array([-88.0, -95.0, -100.0, -103.0], np.float32)
    got = orc.exp_f32(xs).astype(np.float64)
    assert np.all(np.abs(got - np.exp(xs.astype(np.float64))) <= 2.0 ** -149 + 2.0 ** -126 * 2e-7)
    # monotone on a fine grid around 0 and across reduction boundaries
    g = np.linspace(-1.5, 1.5, 300001, dtype=np.float32)
    assert np.all(np.diff(orc.exp_f32(g)) >= 0)
