"""Pins the oracle's FP32 elementary functions against libm in double
precision (numpy float64), on the domains the path tracer uses."""
import numpy as np
import pytest

import oracle


def ulp_err(got, ref64):
    got64 = got.astype(np.float64)
    spacing = np.spacing(np.abs(ref64).astype(np.float32)).astype(np.float64)
    return np.abs(got64 - ref64) / spacing


@pytest.mark.parametrize("fn,ref,lo,hi,tol", [
    (0, np.sin, -4 * np.pi, 4 * np.pi, 2.0),
    (1, np.cos, -4 * np.pi, 4 * np.pi, 2.0),
    (3, np.exp, -87.0, 88.0, 2.0),
    (6, np.exp2, -125.0, 127.0, 2.0),
])
def test_against_libm(fn, ref, lo, hi, tol):
    rng = np.random.default_rng(fn)
    x = rng.uniform(lo, hi, 400000).astype(np.float32)
    got = oracle.math_eval(fn, x)
    r = ref(x.astype(np.float64))
    err = ulp_err(got, r)
    if fn in (0, 1):
        # absolute error near the zeros of sin/cos (the reduction is exact to ~1e-15)
        ok = (err <= tol) | (np.abs(got.astype(np.float64) - r) < 2e-8)
        assert ok.all(), "max ulp %g" % err[~ok].max()
    else:
        assert err.max() <= tol, "max ulp %g at x=%r" % (err.max(), x[np.argmax(err)])


def test_log_against_libm():
    rng = np.random.default_rng(7)
    x = np.exp(rng.uniform(np.log(1e-38), np.log(3e38), 400000)).astype(np.float32)
    x = np.concatenate([x, rng.uniform(0.5, 2.0, 100000).astype(np.float32),
                        np.array([1.0, 2.0, 0.5, 1e-45, 1.17549435e-38], np.float32)])
    got = oracle.math_eval(2, x)
    r = np.log(x.astype(np.float64))
    err = np.abs(got.astype(np.float64) - r)
    rel = err / np.maximum(np.abs(r), 1e-30)
    ulp = ulp_err(got, r)
    assert ((ulp <= 2.0) | (err < 1e-7)).all(), "max ulp %g" % ulp.max()
    assert rel[np.abs(r) > 1e-3].max() < 3e-7


def test_pow_against_libm():
    rng = np.random.default_rng(8)
    x = rng.uniform(1e-4, 10.0, 200000).astype(np.float32)
    y = rng.uniform(-3.0, 3.0, 200000).astype(np.float32)
    got = oracle.math_eval(5, x, y)
    r = np.power(x.astype(np.float64), y.astype(np.float64))
    rel = np.abs(got.astype(np.float64) - r) / r
    # exp(y log x): relative error grows with |y log x| (at most ~20 here)
    assert rel.max() < 3e-6
    assert oracle.math_eval(5, np.array([0.0, 0.0, 0.0, -1.0], np.float32),
                            np.array([2.0, 0.0, -1.0, 2.0], np.float32)).tolist()[:3] == [0.0, 1.0, np.inf]


def test_special_values():
    nan, inf = np.nan, np.inf
    x = np.array([0.0, -0.0, inf, -1.0, nan], np.float32)
    lg = oracle.math_eval(2, x)
    assert lg[0] == -inf and lg[1] == -inf and lg[2] == inf and np.isnan(lg[3]) and np.isnan(lg[4])
    ex = oracle.math_eval(3, np.array([-inf, inf, nan, 0.0, 200.0, -200.0], np.float32))
    assert ex[0] == 0 and ex[1] == inf and np.isnan(ex[2]) and ex[3] == 1.0 and ex[4] == inf and ex[5] == 0.0
    sc = oracle.math_eval(0, np.array([inf, nan, 0.0], np.float32))
    assert np.isnan(sc[0]) and np.isnan(sc[1]) and sc[2] == 0.0
    assert oracle.math_eval(1, np.array([0.0], np.float32))[0] == 1.0


def test_exact_points():
    # exp2 of integers is exact; exp(0) = 1; log(1) = 0
    k = np.arange(-126, 128, dtype=np.float32)
    assert np.array_equal(oracle.math_eval(6, k), np.exp2(k.astype(np.float64)).astype(np.float32))
    assert oracle.math_eval(2, np.array([1.0], np.float32))[0] == 0.0
