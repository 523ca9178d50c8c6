"""MLP and grid backward (fully_fused_mlp.cu:735-836, grid.h:386-509) against FD and the adjoint identity."""
import numpy as np

import npmg
import pyoracle


def test_mlp_backward_matches_fd(oracle):
    cfg = pyoracle.make_cfg(npmg.benchmark_config("small"))
    n_mlp, n_grid = oracle.n_params(cfg)
    p = oracle.init_params(cfg, 5, passes=1)
    q = oracle.synth_queries(3, 64)
    enc = oracle.encode(cfg, p[n_mlp:], q["inputs"], dtype=np.float64)
    raw, acts = oracle.mlp_forward(cfg, p[:n_mlp], enc, acts=True)
    y = np.random.default_rng(0).normal(size=raw.shape)
    dw, denc = oracle.mlp_backward(cfg, p[:n_mlp], enc, acts, y)
    f = lambda w: (oracle.mlp_forward(cfg, w, enc) * y).sum()
    rng = np.random.default_rng(1)
    for j in rng.choice(n_mlp, 40, replace=False):
        # weights are passed as FP32: perturb by a representable step
        w0 = p[:n_mlp].astype(np.float32)
        h = np.float32(1e-4) * max(abs(w0[j]), np.float32(1e-2))
        wp, wm = w0.copy(), w0.copy()
        wp[j] += h
        wm[j] -= h
        fd = (f(wp) - f(wm)) / (float(wp[j]) - float(wm[j]))
        assert abs(fd - dw[j]) <= 1e-6 * max(1.0, abs(dw[j])) + 1e-7, (j, fd, dw[j])
    # input gradient: the network is piecewise linear, so a small step in one feature is exact
    for j in rng.choice(enc.shape[0], 10, replace=False):
        e2 = enc.copy()
        e2[j] += 1e-7
        fd = ((oracle.mlp_forward(cfg, p[:n_mlp], e2) - raw) * y).sum(0) / 1e-7
        assert np.allclose(fd, denc[j], rtol=1e-5, atol=1e-8)


def test_grid_backward_adjoint_identity(oracle):
    """The grid features are linear in the table: <J t, y> = <t, J^T y>."""
    for name in ("small", "ref"):
        conf = npmg.benchmark_config("small") if name == "small" else npmg.reference_config()
        cfg = pyoracle.make_cfg(conf)
        n_mlp, n_grid = oracle.n_params(cfg)
        ew, _ = oracle.widths(cfg)
        rng = np.random.default_rng(2)
        t = rng.normal(size=n_grid).astype(np.float32)
        inp = oracle.synth_queries(4, 500)["inputs"]
        enc = oracle.encode(cfg, t, inp, dtype=np.float64)
        LF = cfg.n_levels * cfg.n_features
        y = np.zeros((ew, 500))
        y[:LF] = rng.normal(size=(LF, 500))
        gt = oracle.grid_backward(cfg, inp, y)
        lhs = (enc[:LF] * y[:LF]).sum()
        rhs = (t.astype(np.float64) * gt).sum()
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs) + 1e-10, (name, lhs, rhs)
