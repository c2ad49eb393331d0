"""Pins of the oracle's powf (the selection test pow(e/emax, alpha) >= u,
kdgrid.cu:326, and the sRGB curve, cuda/helpers.h:38).

The reference's device powf is not reproducible (fast math); the oracle defines
powf as exp(a ln x) in double rounded once to float.  These checks tie it to
libm's double pow (an independent implementation), and to closed forms:
pow(x, 1) = x, pow(x, 2) = fl(x*x), pow(x, 1/2) = sqrtf(x).
"""
import numpy as np
import pytest


def _f32_ulp_diff(a, b):
    ia = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


@pytest.mark.parametrize("alpha", [0.25, 1 / 64, 1 / 32, 1 / 16, 0.5, 1 / 2.4, 0.3, 1.0, 2.0, 3.7])
def test_pow_matches_libm_double(orc, alpha):
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.random(4000, dtype=np.float32),
                         np.float32(10.0) ** rng.uniform(-30, 0, 2000).astype(np.float32),
                         np.array([1e-38, 1e-45, 0.0031308, 0.5, 0.999999], np.float32)])
    xs = xs[xs > 0]
    a32 = np.float32(alpha)
    got = np.array([orc.powf(float(x), float(a32)) for x in xs], np.float32)
    ref = np.power(xs.astype(np.float64), np.float64(a32)).astype(np.float32)
    d = _f32_ulp_diff(got, ref)
    assert d.max() <= 1
    assert (d == 0).mean() > 0.999


def test_pow_closed_forms(orc):
    rng = np.random.default_rng(3)
    xs = rng.random(3000, dtype=np.float32) + np.float32(1e-6)
    for x in xs[:1000]:
        assert orc.powf(float(x), 1.0) == x
    sq = np.array([orc.powf(float(x), 2.0) for x in xs], np.float32)
    assert np.array_equal(sq, xs * xs)
    rt = np.array([orc.powf(float(x), 0.5) for x in xs], np.float32)
    assert np.array_equal(rt, np.sqrt(xs))


def test_pow_special_values(orc):
    assert orc.powf(0.0, 0.25) == 0.0
    assert orc.powf(0.0, 0.0) == 1.0
    assert orc.powf(0.7, 0.0) == 1.0
    assert orc.powf(1.0, 0.3) == 1.0
    assert np.isnan(orc.powf(float("nan"), 0.25))
    assert np.isnan(orc.powf(-0.5, 0.25))
    assert orc.powf(float("inf"), 0.25) == float("inf")
    # err / gmax with gmax = 0 (no leaves with error) -> inf -> always selected
    assert orc.powf(float("inf"), 0.25) >= 0.999
