"""Mixture-KL loss and adjoint (train.cu:69-124) against FP64 central finite differences."""
import numpy as np
import pytest

import pyoracle


def _cfg(K=8, mapping="reference", numerics="stable", clamp=0.0, clamp_scale=128.0):
    return pyoracle.make_cfg({
        "encoding": {"otype": "Composite", "nested": [{"otype": "Identity", "n_dims_to_encode": 9}]},
        "network": {"n_neurons": 64, "n_hidden_layers": 1},
        "mixture": {"n_lobes": K, "mapping": mapping, "numerics": numerics},
        "loss": {"grad_clamp": clamp, "clamp_scale": clamp_scale}})


def _targets(n, seed):
    r = np.random.default_rng(seed)
    z = 1 - 2 * r.random(n)
    return (np.stack([np.arccos(z), 2 * np.pi * r.random(n)]).astype(np.float32),
            np.repeat(r.uniform(0.1, 2, (1, n)), 3, 0).astype(np.float32),
            r.uniform(0.05, 0.5, n).astype(np.float32))


def _raw(K, n, seed, mapping):
    r = np.random.default_rng(seed)
    raw = r.normal(0, 1, size=(((4 * K + 15) // 16) * 16, n))
    kslot = 1 if mapping == "reference" else 2
    raw[kslot:4 * K:4] = r.uniform(np.log(1e-3), np.log(1e3), size=(K, n))  # kappa log-uniform in [1e-3, 1e3]
    return raw


@pytest.mark.parametrize("mapping", ["reference", "softmax"])
def test_adjoint_matches_central_fd(oracle, mapping):
    K, n = 8, 256
    cfg = _cfg(K, mapping)
    raw = _raw(K, n, 1, mapping)
    td, tr, tp = _targets(n, 2)
    g, loss = oracle.adjoint(cfg, raw, td, tr, tp, loss_scale=float(n), dtype=np.float64)  # ls / n = 1
    h = 1e-4
    fd = np.zeros((4 * K, n))
    for j in range(4 * K):
        rp, rm = raw.copy(), raw.copy()
        rp[j] += h
        rm[j] -= h
        _, lp = oracle.adjoint(cfg, rp, td, tr, tp, float(n), dtype=np.float64)
        _, lm = oracle.adjoint(cfg, rm, td, tr, tp, float(n), dtype=np.float64)
        fd[j] = (lp - lm) / (2 * h)
    scale = np.abs(fd).max(axis=0) + 1e-300
    err = np.abs(g[:4 * K] - fd) / scale
    assert np.quantile(err, 0.999) < 1e-3 and err.max() < 1e-2, err.max()
    assert np.all(g[4 * K:] == 0)


def test_loss_closed_form_single_lobe(oracle):
    cfg = _cfg(1)
    n = 64
    r = np.random.default_rng(3)
    raw = np.zeros((16, n))
    raw[0] = r.normal(size=n)
    raw[1] = np.log(r.uniform(0.5, 50, n))
    raw[2], raw[3] = r.normal(size=n), r.normal(size=n)
    td, tr, tp = _targets(n, 4)
    _, loss = oracle.adjoint(cfg, raw, td, tr, tp, 1.0, dtype=np.float64)
    k = np.exp(raw[1])
    th, ph = np.pi / (1 + np.exp(-raw[2])), 2 * np.pi / (1 + np.exp(-raw[3]))
    mu = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    t0, t1 = td.astype(np.float64)
    w = np.stack([np.sin(t0) * np.cos(t1), np.sin(t0) * np.sin(t1), np.cos(t0)])
    c = np.minimum((mu * w).sum(0), 1)
    v = k * np.exp(k * (c - 1)) / (2 * np.pi * -np.expm1(-2 * k))
    f = tr.astype(np.float64).mean(0)
    want = -f / (tp.astype(np.float64) + 1e-5) * np.log(v + 1e-5)
    assert np.allclose(loss, want, rtol=1e-9)


@pytest.mark.parametrize("numerics", ["reference", "stable"])
def test_kappa_derivative_form(oracle, numerics):
    """K = 1: dL/dx_kappa = prefix * v * (mu.w + h(kappa)) * kappa, with h the exact 1/k - coth k
    (stable) or the reference's polynomial 0.000926 - 0.344883 k + 0.030147 k^2 below 1 (distribution.h:165-166)."""
    cfg = _cfg(1, numerics=numerics)
    kap = np.array([1e-3, 1e-2, 0.1, 0.5, 0.9, 2.0, 10.0])
    n = kap.size
    raw = np.zeros((16, n))
    raw[1] = np.log(kap)
    td = np.stack([np.full(n, 0.3), np.full(n, 1.0)]).astype(np.float32)
    tr = np.ones((3, n), np.float32)
    tp = np.full(n, 0.25, np.float32)
    g, _ = oracle.adjoint(cfg, raw, td, tr, tp, float(n), dtype=np.float64)
    k = np.exp(raw[1])
    th, ph = np.pi / 2, np.pi  # raw theta/phi slots are 0: logistic(0) = 1/2
    mu = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    t0, t1 = float(np.float32(0.3)), float(np.float32(1.0))  # targets travel as FP32
    w = np.array([np.sin(t0) * np.cos(t1), np.sin(t0) * np.sin(t1), np.cos(t0)])
    c = mu @ w
    if numerics == "reference":
        v = np.where(k < 1e-5, 1 / (4 * np.pi), np.exp(k * np.minimum(0, c - 1)) * k / (2 * np.pi * (1 - np.exp(-2 * k))))
        hk = np.where(k < 1, 0.000926 - 0.344883 * k + 0.030147 * k * k, 1 / k - 1 / np.tanh(k))
    else:
        v = k * np.exp(k * (c - 1)) / (2 * np.pi * -np.expm1(-2 * k))
        hk = 1 / k - 1 / np.tanh(k)
    prefix = -1.0 / (0.25 + 1e-5) / (v + 1e-5)
    want = prefix * v * (c + hk) * k
    assert np.allclose(g[1], want, rtol=1e-9, atol=1e-12)


def test_nonfinite_zeroed_and_upper_clamp(oracle):
    K, n = 8, 64
    raw = _raw(K, n, 5, "reference")
    td, tr, tp = _targets(n, 6)
    tr[:, :8] = np.inf  # f/q infinite -> non-finite gradients are zeroed (train.cu:106-114)
    g, _ = oracle.adjoint(_cfg(K, clamp=150.0, clamp_scale=1e9), raw, td, tr, tp, 1e9, dtype=np.float64)
    assert np.all(np.isfinite(g)) and np.all(g[:, :8] == 0)
    assert g.max() <= 150.0 and g.max() == 150.0  # min(150, g) (train.cu:116-122); no lower clamp
    assert g.min() < -150.0


def test_clamp_follows_the_loss_scale(oracle):
    """At loss scale s the clamp threshold is grad_clamp s / clamp_scale: the
    scaled gradient divided by s is the same at every scale, clamped entries
    included (the reference fixes s = 128 and clamps at 150)."""
    K, n = 8, 64
    raw = _raw(K, n, 5, "reference")
    td, tr, tp = _targets(n, 6)
    cfg = _cfg(K, clamp=150.0)
    g128, _ = oracle.adjoint(cfg, raw, td, tr, tp, 128.0 * 1e6, dtype=np.float64)
    g32k, _ = oracle.adjoint(cfg, raw, td, tr, tp, 32768.0 * 1e6, dtype=np.float64)
    hi = 150.0 * 1e6
    assert np.sum(g128 == hi) > 0, "some entries must hit the clamp"
    assert np.allclose(g32k / 256.0, g128, rtol=1e-12, atol=1e-200)  # atol: FP64 subnormals


def _h16(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


@pytest.mark.parametrize("numerics", ["reference", "stable"])
def test_fp16_intermediates(oracle, numerics):
    """The reference's __half adjoint buffers (train.cu:80-99, distribution.h:177-178):
    every dL/draw is an FP16 value; the raw outputs enter rounded (rounding them
    beforehand changes nothing); the result stays within a few FP16 steps of the
    FP32 adjoint; and a gradient beyond the FP16 range becomes infinite and is
    zeroed by the NaN test rather than clamped."""
    K, n = 8, 512
    base = dict(K=K, numerics=numerics, clamp=150.0)
    c32 = _cfg(**base)
    c16 = _cfg(**base)
    c16.fp16_intermediates = 1
    raw = _raw(K, n, 4, "reference").astype(np.float32)
    td, tr, tp = _targets(n, 5)
    g32, l32 = oracle.adjoint(c32, raw, td, tr, tp, 128.0)
    g16, l16 = oracle.adjoint(c16, raw, td, tr, tp, 128.0)
    assert np.array_equal(g16, _h16(g16))
    g16b, _ = oracle.adjoint(c16, _h16(raw), td, tr, tp, 128.0)
    assert np.array_equal(g16, g16b)
    g32h, _ = oracle.adjoint(c32, _h16(raw), td, tr, tp, 128.0)
    assert not np.array_equal(g16, g32h)  # the per-lobe dp rounding is a separate step
    scale = np.abs(g32h).max(axis=0, keepdims=True) + 1e-30
    assert np.max(np.abs(g16 - g32h) / scale) < 2.0 ** -8
    assert np.allclose(l16, oracle.adjoint(c32, _h16(raw), td, tr, tp, 128.0)[1], rtol=0, atol=0)
    # overflow: a loss scale that drives gradients past 65504 (no clamp)
    c16.grad_clamp = 0.0
    c32.grad_clamp = 0.0
    big32, _ = oracle.adjoint(c32, _h16(raw), td, tr, tp, 2.0 ** 24)
    big16, _ = oracle.adjoint(c16, raw, td, tr, tp, 2.0 ** 24)
    over = np.abs(big32) > 65520.0 * (1 + 2.0 ** -8)
    assert over.sum() > 0
    assert np.all(big16[over] == 0)
    assert np.all(np.isfinite(big16))
