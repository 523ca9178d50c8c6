"""Grid-gradient scatter (tiny-cuda-nn kernel_grid_backward), binned by table
tile and reduced in LDS: atomic mode (FP32 sums, order varies) and
deterministic mode (exact int64 fixed-point sums). Both against the oracle's
FP64 grid backward on the library's own dL/denc. The small case has coarse
bins that span several reduction chunks; the full case is the benchmark
shape (16 levels x 2^19 entries, 2^18 samples)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("size", ["small", "reference", "full"])
def test_deterministic_mode_bitwise_repeatable_and_matches_fp64(gpu, oracle, size):
    import torch
    import npmg
    import pyoracle

    base = npmg.reference_config() if size == "reference" else npmg.benchmark_config(size)
    cfg_det = dict(base, grid_gradient="deterministic")
    net_d = npmg.GuidingNetwork(cfg_det, seed=3)
    net_a = npmg.GuidingNetwork(dict(base, grid_gradient="atomic"), seed=3)
    n = (1 << 18) if size == "full" else 20000
    q = npmg.synth_queries(12, n)
    dout = torch.from_numpy(np.random.default_rng(4).normal(size=(net_d.padded_output, n)).astype(np.float32)).cuda()
    g1, _ = net_d.backward(q["inputs"], dout)
    g2, _ = net_d.backward(q["inputs"], dout)
    ga, _ = net_a.backward(q["inputs"], dout)
    g1, g2, ga = _np(g1), _np(g2), _np(ga)
    assert np.array_equal(g1.view(np.uint32), g2.view(np.uint32)), "deterministic mode differs between runs"
    nm = net_d.n_mlp
    scale = np.abs(g1[nm:]).max()
    assert np.max(np.abs(g1[nm:] - ga[nm:])) <= 1e-5 * scale
    # against the oracle's FP64 grid backward on the library's dL/denc
    oc = pyoracle.make_cfg(cfg_det)
    _, denc = net_d.backward(q["inputs"], dout, want_denc=True)
    gg = oracle.grid_backward(oc, _np(q["inputs"]), _np(denc).astype(np.float64))
    assert np.max(np.abs(g1[nm:] - gg)) <= 1e-5 * np.abs(gg).max()
    net_d.close()
    net_a.close()
