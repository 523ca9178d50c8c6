"""The oracle's MLP forward (tcnn FullyFusedMLP semantics: bias-free, ReLU on
every hidden layer, no output activation, row-major [out][in] matrices in
layer order; src/fully_fused_mlp.cu:499-557, 635-705) against a textbook
chain of numpy matrix products on random weights and inputs."""
import numpy as np
import pytest

import npmg
import pyoracle


@pytest.mark.parametrize("name", ["small", "full", "ref", "wide"])
def test_mlp_forward_is_textbook_gemm_chain(oracle, name):
    conf = {"small": lambda: npmg.benchmark_config("small"), "full": lambda: npmg.benchmark_config("full"),
            "ref": npmg.reference_config,
            "wide": lambda: npmg.benchmark_config("small", n_lobes=16, width=128, hidden=3)}[name]()
    cfg = pyoracle.make_cfg(conf)
    n_mlp, _ = oracle.n_params(cfg)
    ew, po = oracle.widths(cfg)
    rng = np.random.default_rng(11)
    w = (rng.normal(size=n_mlp) * 0.15).astype(np.float32)
    enc = rng.normal(size=(ew, 200))
    raw, acts = oracle.mlp_forward(cfg, w, enc, acts=True)
    x, off, n_in = enc, 0, ew
    H = cfg.n_hidden_layers
    for layer in range(H + 1):
        n_out = cfg.width if layer < H else po
        W = w[off:off + n_out * n_in].astype(np.float64).reshape(n_out, n_in)
        y = W @ x
        off += n_out * n_in
        n_in = n_out
        if layer < H:
            x = np.maximum(y, 0.0)
            assert np.allclose(acts[layer], x, rtol=1e-12, atol=1e-12), f"hidden layer {layer}"
    assert off == n_mlp
    assert np.allclose(raw, y, rtol=1e-12, atol=1e-12)
    # the FP32 instantiation agrees to FP32 accuracy (accumulation of up to 128 terms)
    raw32 = oracle.mlp_forward(cfg, w, enc.astype(np.float32))
    assert np.max(np.abs(raw32 - y)) <= 1e-5 * np.max(np.abs(y))
