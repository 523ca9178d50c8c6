"""MFMA (BF16/FP16 operands, FP32 accumulation) fused query path.

Checked against a numpy emulation of the same rounding points (grid table,
encoded features, weights and hidden activations rounded to the operand type,
products accumulated in FP64): the only remaining differences are FP32
accumulation order and the occasional activation whose rounding flips. The
mixture stage is FP32 and must reproduce the FP32 consumer kernel bit for bit
on the kernel's own raw outputs. The end-to-end error against the FP64 oracle
with unrounded weights is measured and bounded loosely (it is the cost of the
operand precision, reported by the bench).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _round(a, prec):
    """An operand rounded as the kernels round it: FP16, BF16, or for bf16x3
    the split pair hi = bf16(x), lo = bf16(x - hi) of the FP32 value."""
    import torch

    if prec == "fp16":
        return np.asarray(a, np.float64).astype(np.float16).astype(np.float64)
    b = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    if prec == "bf16x3":
        x = np.asarray(a, np.float32)
        hi = b(x)
        return hi + b(x - hi.astype(np.float32))
    return b(a)


def _storage(prec):
    """Rounding of the grid table and of the level-major grid features: FP16
    for bf16x3 (then split exactly), else the operand type."""
    return "fp16" if prec == "bf16x3" else prec


def _round_enc(enc, prec, oc):
    """Layer-0 inputs: the grid rows (the first n_levels x F) in the storage
    type, then every row as an operand."""
    e = np.array(enc, np.float64)
    g = oc.n_levels * oc.n_features
    e[:g] = _round(e[:g], _storage(prec))
    return _round(e, prec)


def _configs(prec):
    import npmg

    ref = npmg.reference_config()
    ref["precision"] = prec
    return {
        "small": npmg.benchmark_config("small", precision=prec),
        "full": npmg.benchmark_config("full", precision=prec),
        "reference": ref,
        "k16": npmg.benchmark_config("small", precision=prec, n_lobes=16),
        "w128": npmg.benchmark_config("small", precision=prec, width=128),
    }


@pytest.mark.parametrize("prec", ["bf16", "fp16", "bf16x3"])
@pytest.mark.parametrize("name", ["small", "full", "reference", "k16", "w128"])
def test_mfma_query(gpu, oracle, prec, name):
    import npmg
    import pyoracle

    cfg = _configs(prec)[name]
    if prec == "bf16x3" and name == "w128":
        with pytest.raises(npmg.NpmgError, match="width 64"):
            npmg.GuidingNetwork(cfg, seed=7272)
        return
    net = npmg.GuidingNetwork(cfg, seed=7272)
    oc = pyoracle.make_cfg(cfg)
    n = 4096
    q = npmg.synth_queries(41, n)
    out = net.sample_pdf(q["inputs"], q["normal"], q["u"], q["eval_dirs"], raw=True)
    h = {k: _np(v) for k, v in q.items()}
    p = _np(net.get_params(1)).astype(np.float32)
    n_mlp = net.n_mlp
    # emulation of the rounding points
    table = _round(p[n_mlp:], _storage(prec)).astype(np.float32)
    enc = oracle.encode(oc, table, h["inputs"])  # FP32 features, as the kernel computes them
    a = _round_enc(enc, prec, oc)
    layers = []
    off, n_in = 0, net.encoded_width
    for l in range(oc.n_hidden_layers + 1):
        n_out = oc.width if l < oc.n_hidden_layers else net.padded_output
        layers.append(_round(p[off:off + n_out * n_in], prec).reshape(n_out, n_in))
        off += n_out * n_in
        n_in = oc.width
    for l, W in enumerate(layers):
        a = W @ a
        if l < len(layers) - 1:
            a = _round(np.maximum(a, 0), prec)
    raw = _np(out["raw"])
    scale = np.abs(a).max(axis=0) + 1e-6
    err = np.abs(raw - a).max(axis=0) / scale
    assert np.quantile(err, 0.95) < 2e-5, np.quantile(err, 0.95)
    assert err.max() < 3e-2, err.max()
    # mixture stage: identical to the FP32 consumer kernel on the kernel's raw outputs
    c = net.mixture_sample_pdf(out["raw"], q["normal"], q["u"], q["eval_dirs"])
    for k in ("dir", "pdf", "eval_pdf", "lobe"):
        assert np.array_equal(_np(out[k]).view(np.uint32), _np(c[k]).view(np.uint32)), k
    # ... and the pdf of that mixture is within 1e-5 of its FP64 evaluation
    gold_mix = oracle.mixture(oc, raw.astype(np.float64), h["normal"], h["u"], h["eval_dirs"], dtype=np.float64)
    rel = np.abs(_np(out["eval_pdf"]) - gold_mix["eval_pdf"]) / gold_mix["eval_pdf"]
    assert rel.max() < 1e-5, rel.max()
    # end to end against the FP64 oracle with unrounded parameters (cost of the operand precision)
    gold = oracle.query(oc, p, h["inputs"], h["normal"], h["u"], h["eval_dirs"], dtype=np.float64)
    rel = np.abs(_np(out["eval_pdf"]) - gold["eval_pdf"]) / gold["eval_pdf"]
    assert np.median(rel) < {"bf16": 2e-2, "fp16": 3e-3, "bf16x3": 1e-3}[prec], np.median(rel)
    net.close()
