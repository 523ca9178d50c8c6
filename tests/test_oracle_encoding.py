"""OneBlob, spherical harmonics and composite padding against closed forms."""
import numpy as np
from scipy import integrate

import pyoracle


def _cfg(nested, n_in=9):
    return {"n_input_dims": n_in, "encoding": {"otype": "Composite", "nested": nested},
            "network": {"n_neurons": 64, "n_hidden_layers": 1}, "mixture": {"n_lobes": 8}}


def quartic(u):
    return np.where(np.abs(u) < 1, 15 / 16 * (1 - u * u) ** 2, 0.0)


def test_oneblob_mass_closed_form(oracle):
    nb = 16
    cfg = pyoracle.make_cfg(_cfg([{"otype": "OneBlob", "n_dims_to_encode": 9, "n_bins": nb}]))
    x = np.random.default_rng(0).random((9, 40)).astype(np.float32)
    x[0, :4] = [0.0, 0.5, 1.0, 0.03125]
    enc = oracle.encode(cfg, np.zeros(1, np.float32), x, dtype=np.float64)
    # bin k mass = integral over [k/nb, (k+1)/nb] of the wrapped kernel of radius 1/nb centred at x
    for q in range(6):
        for d in range(2):
            xv = float(x[d, q])
            for k in range(nb):
                m = sum(integrate.quad(lambda t: nb * quartic(nb * (t - xv + s)), k / nb, (k + 1) / nb)[0]
                        for s in (-1, 0, 1))
                assert abs(enc[d * nb + k, q] - m) < 1e-9
    masses = enc[: 9 * nb].reshape(9, nb, -1).sum(1)
    assert np.allclose(masses, 1.0, atol=1e-12)
    # x = 0.5 is symmetric: bins mirror
    e = enc[:nb, 1]
    assert np.allclose(e, e[::-1], atol=1e-12)


def test_sh_degree2_and_padding(oracle):
    cfg = pyoracle.make_cfg(_cfg([{"otype": "SphericalHarmonics", "n_dims_to_encode": 3, "degree": 2},
                                   {"otype": "Identity", "n_dims_to_encode": 6}]))
    ew, _ = oracle.widths(cfg)
    assert ew == 16
    x = np.random.default_rng(1).random((9, 20)).astype(np.float32)
    enc = oracle.encode(cfg, np.zeros(1, np.float32), x, dtype=np.float64)
    X, Y, Z = (2 * x[:3].astype(np.float64) - 1)
    c0, c1 = 0.5 / np.sqrt(np.pi), np.sqrt(3) / (2 * np.sqrt(np.pi))
    assert np.allclose(enc[0], c0) and np.allclose(enc[1], -c1 * Y) and np.allclose(enc[2], c1 * Z)
    assert np.allclose(enc[3], -c1 * X)
    assert np.allclose(enc[4:10], x[3:9])
    assert np.all(enc[10:16] == 1.0)  # composite padding to a multiple of 16 is 1.0 (tcnn)
    # orthonormality of the four basis functions on the sphere (directions mapped to [0,1]^3)
    n = 200000
    r = np.random.default_rng(2)
    z = 1 - 2 * r.random(n)
    ph = 2 * np.pi * r.random(n)
    s = np.sqrt(1 - z * z)
    d = np.stack([s * np.cos(ph), s * np.sin(ph), z])
    xin = np.zeros((9, n), np.float32)
    xin[:3] = (d + 1) / 2
    e = oracle.encode(cfg, np.zeros(1, np.float32), xin, dtype=np.float64)[:4]
    gram = 4 * np.pi * (e @ e.T) / n
    assert np.allclose(gram, np.eye(4), atol=0.02)


def test_reference_composite_width(oracle):
    import npmg

    cfg = pyoracle.make_cfg(npmg.reference_config())
    assert oracle.widths(cfg) == (48, 32)
    assert oracle.n_params(cfg) == (48 * 64 + 2 * 64 * 64 + 64 * 32, 4133152)
