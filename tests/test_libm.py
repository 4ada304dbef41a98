"""CPU: accuracy of the shared FP64 library (pisco_libm.h) that both the
oracle and the kernels evaluate, against mpmath; special values against
the C library."""
import math

import numpy as np
import pytest

mpmath = pytest.importorskip("mpmath")


def ulp_err(got, ref):
    out = []
    for g, r in zip(got, ref):
        sp = float(np.spacing(abs(float(r)))) or 5e-324
        out.append(float(abs(mpmath.mpf(float(g)) - r) / sp))
    return np.array(out)


@pytest.mark.parametrize("name,gen", [
    ("sin", lambda r: r.uniform(-13, 13, 3000)),
    ("cos", lambda r: r.uniform(-13, 13, 3000)),
    ("acos", lambda r: np.concatenate([r.uniform(-1, 1, 2000), 1 - 10 ** r.uniform(-16, 0, 1000)])),
    ("asin", lambda r: np.concatenate([r.uniform(-1, 1, 2000), 1 - 10 ** r.uniform(-16, 0, 500),
                                       -(1 - 10 ** r.uniform(-16, 0, 500)),
                                       10 ** r.uniform(-300, -1, 300)])),
])
def test_unary_accuracy(oracle, name, gen):
    mpmath.mp.dps = 40
    x = gen(np.random.default_rng(1))
    got = oracle.vec(name, x)
    ref = [getattr(mpmath, name)(mpmath.mpf(float(v))) for v in x]
    # asin just above |x| = 1/2 is pi/2 - 2 asin(sqrt((1 - |x|)/2)): the
    # doubling and the cancellation against pi/2 cost up to 1.6 ulp there
    bound = 2.0 if name == "asin" else 1.0
    assert ulp_err(got, ref).max() < bound


def test_asin_special_values(oracle):
    x = np.array([0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 5e-324])
    g = oracle.vec("asin", x)
    assert np.array_equal(np.signbit(g), np.signbit(x))
    assert g[2] == math.pi / 2 and g[3] == -math.pi / 2 and g[6] == 5e-324
    assert np.isnan(oracle.vec("asin", np.array([1.0000000000000002, math.nan]))).all()


def test_atan2_accuracy_and_special_values(oracle):
    mpmath.mp.dps = 40
    rng = np.random.default_rng(2)
    y = rng.normal(size=3000) * 10 ** rng.uniform(-6, 3, 3000)
    x = rng.normal(size=3000) * 10 ** rng.uniform(-6, 3, 3000)
    got = oracle.vec("atan2", y, x)
    ref = [mpmath.atan2(mpmath.mpf(float(a)), mpmath.mpf(float(b))) for a, b in zip(y, x)]
    assert ulp_err(got, ref).max() < 2.0
    vals = [0.0, -0.0, 1.0, -1.0, math.inf, -math.inf, 2.5e-300]
    ys = np.array([a for a in vals for _ in vals])
    xs = np.array([b for _ in vals for b in vals])
    g = oracle.vec("atan2", ys, xs)
    r = np.arctan2(ys, xs)
    assert np.allclose(g, r, rtol=0, atol=1e-15)
    assert np.array_equal(np.signbit(g), np.signbit(r))
    assert np.isnan(oracle.vec("atan2", np.array([math.nan]), np.array([1.0]))[0])
