"""The oracle's threaded FP32 train step (the CPU baseline of the train step)
against the composition of its single-threaded pieces: forward, adjoint,
MLP backward, grid backward and Adam (integrator.cpp:679-699). The grid
gradient and the loss follow the same operation order; the weight gradient's
per-thread partials are summed in another order, so the MLP parameters agree
to FP32 rounding of the Adam step."""
import numpy as np
import pytest

import npmg
import pyoracle


@pytest.mark.parametrize("threads", [1, 4])
def test_train_step_matches_pieces(oracle, threads):
    conf = npmg.benchmark_config("small")
    cfg = pyoracle.make_cfg(conf)
    n_mlp, n_grid = oracle.n_params(cfg)
    p0 = oracle.init_params(cfg, 7, passes=1)
    t = oracle.synth_train(3, 3000)
    # pieces
    enc = oracle.encode(cfg, p0[n_mlp:], t["inputs"])
    raw, acts = oracle.mlp_forward(cfg, p0[:n_mlp], enc, acts=True)
    dout, loss = oracle.adjoint(cfg, raw, t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], 128.0)
    dw, denc = oracle.mlp_backward(cfg, p0[:n_mlp], enc, acts, dout)
    gg = oracle.grid_backward(cfg, t["inputs"], denc)
    want = p0.copy()
    m, v, st = np.zeros_like(want), np.zeros_like(want), np.zeros(want.size, np.uint32)
    oracle.adam_step(n_mlp, 128.0, 1e-2, 0.9, 0.99, 1e-15, 1e-6, want, np.concatenate([dw, gg]).astype(np.float32),
                     m, v, st)
    # threaded step
    got = p0.copy()
    m2, v2, st2 = np.zeros_like(got), np.zeros_like(got), np.zeros(got.size, np.uint32)
    mean = oracle.train_step(cfg, got, m2, v2, st2, t, n_threads=threads)
    assert abs(mean - loss.mean()) <= 1e-6 * abs(loss.mean())
    assert np.array_equal(st2, st)
    assert np.array_equal(got[n_mlp:], want[n_mlp:]) and np.array_equal(v2[n_mlp:], v[n_mlp:])
    # matrix entries: the first Adam step moves each by about lr * sign(g), so compare
    # the moments, which carry the gradient itself
    a, b = m2[:n_mlp].astype(np.float64), m[:n_mlp].astype(np.float64)
    assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)
    assert np.max(np.abs(got[:n_mlp] - want[:n_mlp])) <= 1e-6
