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


def test_round_half_matches_numpy_float16(oracle):
    """fn 12 (FP16 rounding, round to nearest even) against numpy's float16
    conversion over normals, subnormals, ties, the overflow edge and specials."""
    r = np.random.default_rng(12)
    f32 = np.float32
    mag = np.exp(r.uniform(np.log(1e-9), np.log(7e4), 200000))
    x = (mag * np.sign(r.normal(size=mag.size))).astype(f32)
    # exact ties between neighbouring FP16 values (normal and subnormal), both parities
    h = np.arange(1, 0x7bff, 37, dtype=np.uint16).view(np.float16).astype(np.float64)
    nxt = (np.arange(1, 0x7bff, 37, dtype=np.uint16) + 1).view(np.float16).astype(np.float64)
    ties = ((h + nxt) / 2).astype(f32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, 65504.0, 65519.99, 65520.0, -65520.0, 2.0 ** -24,
                        2.0 ** -25, 2.0 ** -25 * 1.0001, 2.0 ** -26, 6.1e-5, 6.103515625e-05, 3.4e38], f32)
    x = np.concatenate([x, ties, -ties, special])
    got = oracle.math(12, x)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).astype(f32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.isnan(oracle.math(12, np.array([np.nan], f32))[0])
