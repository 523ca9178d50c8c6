"""The oracle's learned-selection loss and adjoint (compute_dL_doutput_with_
selection, train.cu:126-172; MixedSGWithSelection, distribution.h:204-241),
pinned by FP64 central differences.

With s = logistic(x[28]) and P(x) = s (p(x) + 1e-5) + (1 - s) b, the
reference's gradient is -f/P0^2 dP/dx (the sampling pdf P0 held fixed, as
the divergence gradient treats it), i.e. the derivative of
G(x) = -(f/P0) log P(x) at x0, times loss_scale / n; the selection slot is
scaled by LR_PROB = 0.01 (train.h:61) and uses p - b where dP/dx has
p + 1e-5 - b. Both are checked against central differences of G computed
from the oracle's FP64 mixture pdf, and the loss value against -f/P log P.
"""
import numpy as np
import pytest


def _cfg(mapping):
    import pyoracle

    c = {"encoding": {"otype": "Composite", "nested": [
            {"otype": "HashGrid", "n_dims_to_encode": 3, "n_levels": 2, "n_features_per_level": 2,
             "log2_hashmap_size": 10, "base_resolution": 4, "per_level_scale": 2.0},
            {"otype": "OneBlob", "n_dims_to_encode": 6, "n_bins": 4}]},
         "network": {"n_neurons": 64, "n_hidden_layers": 2},
         "mixture": {"otype": "vMF", "n_lobes": 7, "mapping": mapping, "numerics": "stable",
                     "learn_selection": True}}
    return pyoracle.make_cfg(c)


@pytest.mark.parametrize("mapping", ["reference", "softmax"])
def test_selection_adjoint_matches_central_differences(oracle, mapping):
    oc = _cfg(mapping)
    assert oc.n_output_dims == 29
    n = 24
    rng = np.random.default_rng(5)
    raw = rng.normal(0, 0.8, (32, n))
    wi32 = rng.normal(size=(3, n)).astype(np.float32)
    wi32 /= np.linalg.norm(wi32, axis=0)
    th = np.arccos(np.clip(wi32[2].astype(np.float64), -1, 1)).astype(np.float32)
    ph = np.mod(np.arctan2(wi32[1], wi32[0]).astype(np.float64), 2 * np.pi).astype(np.float32)
    tdir = np.stack([th, ph])
    wi = np.stack([np.sin(th.astype(np.float64)) * np.cos(ph), np.sin(th.astype(np.float64)) * np.sin(ph),
                   np.cos(th.astype(np.float64))]).astype(np.float32)[None]
    rgb = rng.uniform(0.1, 2, (3, n)).astype(np.float32)
    tpdf = rng.uniform(0.1, 1, n).astype(np.float32)
    bsdf = rng.uniform(0.0, 0.5, n).astype(np.float32)
    ls = 7.0
    g, loss = oracle.adjoint_selection(oc, raw, tdir, rgb, tpdf, bsdf, ls, dtype=np.float64)
    f = rgb.astype(np.float64).mean(axis=0)

    def P(r):
        p = oracle.mixture(oc, r, None, None, wi, dtype=np.float64)["eval_pdf"][0]
        s = 1 / (1 + np.exp(-r[28]))
        return s * (p + 1e-5) + (1 - s) * bsdf, p, s

    P0, p0, s0 = P(raw)
    assert np.allclose(loss, -f / P0 * np.log(P0), rtol=1e-6, atol=0)
    h = 1e-6
    for j in list(range(28)) + [28]:
        e = np.zeros_like(raw)
        e[j] = h
        fd = (-(f / P0) * (np.log(P(raw + e)[0]) - np.log(P(raw - e)[0])) / (2 * h)) * ls / n
        if j == 28:  # LR_PROB and p - b in place of p + 1e-5 - b
            fd = 0.01 * fd * (p0 - bsdf) / (p0 + 1e-5 - bsdf)
        err = np.abs(g[j] - fd) / (np.abs(fd).max() + 1e-300)
        assert err.max() < 2e-5, (j, err.max())
    assert not np.any(g[29:])


def test_selection_off_is_the_divergence_loss(oracle):
    """Without learn_selection the oracle's adjoint is the plain divergence
    adjoint (train.cu:69-124) and the outputs are 4 K wide."""
    oc = _cfg("reference")
    assert oc.learn_selection == 1
    oc.learn_selection = 0
    oc.n_output_dims = 28
    n = 8
    rng = np.random.default_rng(1)
    raw = rng.normal(0, 0.8, (32, n))
    args = (np.stack([rng.uniform(0.1, 3, n), rng.uniform(0, 6, n)]).astype(np.float32),
            rng.uniform(0.1, 2, (3, n)).astype(np.float32), rng.uniform(0.1, 1, n).astype(np.float32))
    g0, l0 = oracle.adjoint(oc, raw, *args, 3.0, dtype=np.float64)
    g1, l1 = oracle.adjoint_selection(oc, raw, *args, np.full(n, 0.3, np.float32), 3.0, dtype=np.float64)
    assert np.array_equal(g0, g1) and np.array_equal(l0, l1)
