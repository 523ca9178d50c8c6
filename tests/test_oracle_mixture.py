"""vMF mixture: closed forms, sphere quadrature, cosine product, sampling statistics, lobe-selection replay."""
import numpy as np
import pytest

import pyoracle


def _cfg(K, mapping="reference", numerics="stable"):
    return pyoracle.make_cfg({
        "encoding": {"otype": "Composite", "nested": [{"otype": "Identity", "n_dims_to_encode": 9}]},
        "network": {"n_neurons": 64, "n_hidden_layers": 1},
        "mixture": {"n_lobes": K, "mapping": mapping, "numerics": numerics}})


def _raw_single(kappa, theta, phi, n):
    """Reference-mapping raw outputs of a single lobe: lambda = 1, kappa, theta = pi s(x), phi = 2 pi s(x)."""
    logit = lambda p: np.log(p / (1 - p))
    raw = np.zeros((16, n))
    raw[1] = np.log(kappa)
    raw[2] = logit(theta / np.pi)
    raw[3] = logit(phi / (2 * np.pi))
    return raw


def _frame(mu):
    a = np.array([1.0, 0, 0]) if abs(mu[0]) < 0.9 else np.array([0, 1.0, 0])
    t = np.cross(mu, a)
    t /= np.linalg.norm(t)
    return t, np.cross(mu, t)


def _dirs_about(mu, c, psi):
    t, b = _frame(mu)
    s = np.sqrt(np.maximum(0, 1 - c * c))
    return (np.outer(mu, c) + np.outer(t, s * np.cos(psi)) + np.outer(b, s * np.sin(psi)))


@pytest.mark.parametrize("kappa", [1e-3, 0.1, 1.0, 10.0, 100.0, 1000.0])
@pytest.mark.parametrize("numerics", ["stable", "reference"])
def test_single_lobe_textbook_and_normalised(oracle, kappa, numerics):
    cfg = _cfg(1, numerics=numerics)
    theta, phi = 1.1, 4.0
    mu = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    # Gauss-Legendre in c = mu.w, refined near c = 1
    split = max(-1.0, 1 - 20 / kappa)
    xs, ws = np.polynomial.legendre.leggauss(400)
    c = np.concatenate([(split + 1) / 2 * xs + (split - 1) / 2, (1 - split) / 2 * xs + (1 + split) / 2])
    wc = np.concatenate([(split + 1) / 2 * ws, (1 - split) / 2 * ws])
    psi = np.random.default_rng(0).uniform(0, 2 * np.pi, c.size)
    d = _dirs_about(mu, c, psi)
    n = c.size
    raw = _raw_single(kappa, theta, phi, n)
    r = oracle.mixture(cfg, raw, eval_dirs=d[None].astype(np.float32), dtype=np.float64)
    p = r["eval_pdf"][0]
    cd = (d.astype(np.float32).astype(np.float64).T @ mu)
    want = kappa * np.exp(kappa * (np.minimum(cd, 1) - 1)) / (2 * np.pi * -np.expm1(-2 * kappa))
    assert np.allclose(p, want, rtol=1e-6 if numerics == "stable" else 1e-4)
    total = 2 * np.pi * np.sum(wc * p)
    assert abs(total - 1) < 1e-4


def test_small_kappa_limit(oracle):
    raw = _raw_single(np.exp(-10.0), 1.0, 2.0, 4)
    d = np.random.default_rng(1).normal(size=(1, 3, 4))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = oracle.mixture(_cfg(1), raw, eval_dirs=d.astype(np.float32), dtype=np.float64)
    assert np.allclose(r["eval_pdf"], 1 / (4 * np.pi), rtol=1e-4)


def _random_raw(K, n, seed, mapping="reference", kscale=1.0):
    rng = np.random.default_rng(seed)
    raw = rng.normal(0, 1, size=(((4 * K + 15) // 16) * 16, n))
    raw[1:4 * K:4] = rng.uniform(-1, 4, size=(K, n)) * kscale if mapping == "reference" else raw[1:4 * K:4]
    if mapping == "softmax":
        raw[2:4 * K:4] = rng.uniform(-1, 4, size=(K, n)) * kscale
    return raw


@pytest.mark.parametrize("mapping", ["reference", "softmax"])
@pytest.mark.parametrize("cosine", [False, True])
def test_mixture_normalised_fibonacci(oracle, mapping, cosine):
    K, m = 8, 400000
    cfg = _cfg(K, mapping=mapping)
    raw = _random_raw(K, 1, 3, mapping)
    i = np.arange(m) + 0.5
    z = 1 - 2 * i / m
    ph = np.pi * (1 + 5 ** 0.5) * i
    s = np.sqrt(1 - z * z)
    d = np.stack([s * np.cos(ph), s * np.sin(ph), z]).astype(np.float32)
    normal = np.array([[0.3], [-0.5], [0.81]], np.float32)
    normal /= np.linalg.norm(normal)
    # one query, m evaluation directions: evaluate in chunks of the eval axis
    tot = 0.0
    for c0 in range(0, m, 20000):
        chunk = d[:, c0:c0 + 20000]
        rr = np.repeat(raw, chunk.shape[1], axis=1)
        r = oracle.mixture(cfg, rr, normal=np.repeat(normal, chunk.shape[1], 1) if cosine else None,
                           eval_dirs=chunk[None], dtype=np.float64)
        tot += r["eval_pdf"].sum()
    assert abs(4 * np.pi * tot / m - 1) < 2e-3


@pytest.mark.parametrize("numerics", ["stable", "reference"])
def test_cosine_product_pointwise(oracle, numerics):
    """After applyCosineLobe, pdf(w) is proportional to v_cos(w) * p(w) (products of vMFs)."""
    K, n = 8, 3000
    cfg = _cfg(K, numerics=numerics)
    raw = np.repeat(_random_raw(K, 1, 4), n, axis=1)
    rng = np.random.default_rng(5)
    d = rng.normal(size=(1, 3, n))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    nrm = np.repeat(np.array([[0.6], [0.0], [0.8]], np.float32), n, 1)
    plain = oracle.mixture(cfg, raw, eval_dirs=d, dtype=np.float64)["eval_pdf"][0]
    prod = oracle.mixture(cfg, raw, normal=nrm, eval_dirs=d, dtype=np.float64)["eval_pdf"][0]
    kb = float(np.float32(2.18853))  # the configuration carries kappa_cos as a float
    nb = nrm[:, 0].astype(np.float64)
    nb /= np.linalg.norm(nb)
    vb = kb * np.exp(kb * (d[0].astype(np.float64).T @ nb - 1)) / (2 * np.pi * -np.expm1(-2 * kb))
    ratio = prod / (vb * plain)
    assert np.std(ratio) / np.mean(ratio) < (1e-9 if numerics == "stable" else 1e-5)


def test_sampling_chi_square(oracle):
    from scipy import stats

    K, n = 8, 400000
    cfg = _cfg(K)
    raw = np.repeat(_random_raw(K, 1, 6), n, axis=1)
    u = np.random.default_rng(7).random((3, n)).astype(np.float32)
    r = oracle.mixture(cfg, raw, u=u, dtype=np.float64)
    d = r["dir"]
    assert np.allclose(np.linalg.norm(d, axis=0), 1, atol=1e-12)
    nz, nph = 16, 32
    zb = np.clip(((d[2] + 1) / 2 * nz).astype(int), 0, nz - 1)
    pb = np.clip((np.mod(np.arctan2(d[1], d[0]), 2 * np.pi) / (2 * np.pi) * nph).astype(int), 0, nph - 1)
    counts = np.bincount(zb * nph + pb, minlength=nz * nph)
    # expected mass per bin: midpoint rule on a 12 x 12 sub-grid of each (z, phi) cell
    sub = 12
    zc = -1 + (np.arange(nz * sub) + 0.5) * 2 / (nz * sub)
    pc = (np.arange(nph * sub) + 0.5) * 2 * np.pi / (nph * sub)
    Z, P = np.meshgrid(zc, pc, indexing="ij")
    S = np.sqrt(1 - Z * Z)
    dirs = np.stack([S * np.cos(P), S * np.sin(P), Z]).reshape(3, -1).astype(np.float32)
    m = dirs.shape[1]
    pdf = oracle.mixture(cfg, np.repeat(raw[:, :1], m, 1), eval_dirs=dirs[None], dtype=np.float64)["eval_pdf"][0]
    cell = (2 / (nz * sub)) * (2 * np.pi / (nph * sub))
    mass = (pdf * cell).reshape(nz, sub, nph, sub).sum(axis=(1, 3)).reshape(-1)
    assert abs(mass.sum() - 1) < 1e-3
    exp = mass / mass.sum() * n
    keep = exp > 5
    chi2 = ((counts[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    p = stats.chi2.sf(chi2, keep.sum() - 1)
    assert p > 1e-4, (chi2, keep.sum())


@pytest.mark.parametrize("cosine", [False, True])
def test_lobe_selection_replay_and_pdf_consistency(oracle, cosine):
    """Replays distribution.h:132-140 (running subtraction over FP32 weights) in numpy."""
    K, n = 8, 20000
    cfg = _cfg(K)
    raw = _random_raw(K, n, 8).astype(np.float32)
    rng = np.random.default_rng(9)
    u = rng.random((3, n)).astype(np.float32)
    nrm = rng.normal(size=(3, n)).astype(np.float32)
    r = oracle.mixture(cfg, raw, normal=nrm if cosine else None, u=u)
    w = r["lobes"].reshape(K, 5, n)[:, 4].astype(np.float32)
    uu = u[0].copy()
    sel = np.full(n, -1)
    for i in range(K):
        hit = (sel < 0) & (uu < w[i])
        sel[hit] = i
        uu = np.where(sel < 0, np.float32(uu - w[i]), uu).astype(np.float32)
    sel[sel < 0] = 0
    assert np.array_equal(sel.astype(np.uint32), r["lobe"])
    again = oracle.mixture(cfg, raw, normal=nrm if cosine else None, eval_dirs=r["dir"][None])
    assert np.array_equal(again["eval_pdf"][0].view(np.uint32), r["pdf"].view(np.uint32))
