"""The deterministic elementary functions (npmg_math.h, via the oracle) against libm in FP64."""
import numpy as np
import pytest

FNS = {  # id: (reference, domain sampler)
    0: (np.exp, lambda r: r.uniform(-87, 88, 400000)),
    1: (np.log, lambda r: np.exp(r.uniform(-87, 88, 400000))),
    2: (np.log1p, lambda r: r.uniform(-0.999, 50, 400000)),
    3: (np.expm1, lambda r: r.uniform(-20, 20, 400000)),
    4: (np.sin, lambda r: r.uniform(-7, 7, 400000)),
    5: (np.cos, lambda r: r.uniform(-7, 7, 400000)),
    6: (np.arccos, lambda r: r.uniform(-1, 1, 400000)),
    9: (lambda x: 1 / (1 + np.exp(-x)), lambda r: r.uniform(-30, 30, 400000)),
}


def ulp(x):
    return np.spacing(np.abs(x).astype(np.float32)).astype(np.float64)


@pytest.mark.parametrize("fn", sorted(FNS))
def test_accuracy(oracle, fn):
    ref, dom = FNS[fn]
    x = dom(np.random.default_rng(fn)).astype(np.float32)
    got = oracle.math(fn, x).astype(np.float64)
    want = ref(x.astype(np.float64))
    err = np.abs(got - want)
    if fn in (4, 5):  # absolute error near zeros of sin/cos
        assert np.max(err) < 2e-7
    else:
        assert np.max(err / np.maximum(ulp(want), 1e-45)) <= 4.0, np.max(err / ulp(want))


def test_atan2_and_special_values(oracle):
    r = np.random.default_rng(7)
    y, x = r.normal(size=200000).astype(np.float32), r.normal(size=200000).astype(np.float32)
    got = oracle.math(7, y, x).astype(np.float64)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(got - want) / ulp(want)) <= 4.0
    f32 = np.float32
    assert oracle.math(0, np.array([0.0], f32))[0] == 1.0
    assert oracle.math(1, np.array([1.0], f32))[0] == 0.0
    assert np.isinf(oracle.math(0, np.array([100.0], f32))[0])
    assert oracle.math(0, np.array([-200.0], f32))[0] == 0.0
    assert np.isneginf(oracle.math(1, np.array([0.0], f32))[0])
    assert oracle.math(6, np.array([1.0], f32))[0] == 0.0
    assert oracle.math(6, np.array([-1.0], f32))[0] == np.float32(np.pi)
    assert oracle.math(7, np.array([0.0], f32), np.array([-1.0], f32))[0] == np.float32(np.pi)
    assert np.isnan(oracle.math(6, np.array([1.5], f32))[0])
    # expm1 / log1p keep full relative accuracy for tiny arguments
    tiny = np.array([1e-7, -3e-6, 2e-10], f32)
    assert np.allclose(oracle.math(3, tiny), np.expm1(tiny.astype(np.float64)), rtol=1e-6, atol=0)
    assert np.allclose(oracle.math(2, tiny), np.log1p(tiny.astype(np.float64)), rtol=1e-6, atol=0)
