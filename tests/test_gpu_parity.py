"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Bitwise: elementary functions, synthetic records, parameter initialisation,
hash-grid indices and weights, encoded features, FP32 network outputs, lobe
selection, sampled directions and pdfs (the FP32 path runs the oracle's
operation sequence). Tolerance: gradients (different but fixed summation
order) against the FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _configs():
    import npmg

    ref = npmg.reference_config()
    return {
        "small": npmg.benchmark_config("small"),
        "full": npmg.benchmark_config("full"),
        "reference": ref,
        "small_refnum": dict(npmg.benchmark_config("small"),
                             mixture={"otype": "vMF", "n_lobes": 8, "mapping": "reference", "numerics": "reference"}),
        "k16": npmg.benchmark_config("small", n_lobes=16),
        "refnum_h16": dict(npmg.benchmark_config("small"),
                           mixture={"otype": "vMF", "n_lobes": 8, "mapping": "reference", "numerics": "reference"},
                           loss={"scale": 128.0, "grad_clamp": 150.0, "fp16_intermediates": True}),
        # the fast numerics (hardware transcendentals): not bit-exact, so only
        # in the tests with a tolerance
        "small_fast": dict(npmg.benchmark_config("small"),
                           mixture={"otype": "vMF", "n_lobes": 8, "mapping": "softmax", "numerics": "fast"}),
        "k16_fast": dict(npmg.benchmark_config("small", n_lobes=16),
                         mixture={"otype": "vMF", "n_lobes": 16, "mapping": "softmax", "numerics": "fast"}),
    }


CONFIGS = [k for k in _configs() if not k.endswith("_fast")]


@pytest.fixture(scope="module")
def nets(gpu):
    import npmg

    made = {}

    def get(name):
        if name not in made:
            made[name] = npmg.GuidingNetwork(_configs()[name], seed=7272)
        return made[name]

    yield get
    for n in made.values():
        n.close()


def test_debug_math_bitwise(gpu, oracle):
    import torch
    import npmg

    rng = np.random.default_rng(0)
    domains = {
        0: rng.uniform(-110, 95, 200000), 1: np.exp(rng.uniform(-100, 88, 200000)),
        2: rng.uniform(-0.999, 10, 200000), 3: rng.uniform(-30, 30, 200000),
        4: rng.uniform(-20, 20, 200000), 5: rng.uniform(-20, 20, 200000),
        6: rng.uniform(-1, 1, 200000), 7: rng.normal(size=200000), 8: rng.uniform(0, 1e6, 200000),
        9: rng.uniform(-30, 30, 200000), 10: rng.normal(size=200000),
        12: np.exp(rng.uniform(-20, 11.2, 200000)) * np.sign(rng.normal(size=200000)),
    }
    for fn, x in domains.items():
        x = x.astype(np.float32)
        y = rng.normal(size=x.size).astype(np.float32) if fn in (7, 10) else None
        g = npmg.debug_math(fn, torch.from_numpy(x).cuda(), None if y is None else torch.from_numpy(y).cuda())
        r = oracle.math(fn, x, y)
        assert np.array_equal(_np(g).view(np.uint32), r.view(np.uint32)), f"fn {fn} differs"


def test_synth_bitwise(gpu, oracle):
    import npmg

    q = npmg.synth_queries(123, 5000, first=77)
    r = oracle.synth_queries(123, 5000, first=77)
    for k in r:
        assert np.array_equal(_np(q[k]).view(np.uint32), r[k].view(np.uint32)), k
    t = npmg.synth_train(9, 5000, first=3)
    r = oracle.synth_train(9, 5000, first=3)
    for k in r:
        assert np.array_equal(_np(t[k]).view(np.uint32), r[k].view(np.uint32)), k


@pytest.mark.parametrize("name", ["small", "reference", "full"])
def test_init_and_layout(nets, oracle, name):
    import pyoracle

    net = nets(name)
    oc = pyoracle.make_cfg(net.config)
    assert (net.n_mlp, net.n_grid) == oracle.n_params(oc)
    assert (net.encoded_width, net.padded_output) == oracle.widths(oc)
    offs, scales, res = net.grid_layout()
    o_offs, o_scales, o_res = oracle.grid_layout(oc)
    assert np.array_equal(offs, o_offs) and np.array_equal(res, o_res)
    assert np.array_equal(scales.view(np.uint32), o_scales.view(np.uint32))
    p = _np(net.get_params(0))
    ref = oracle.init_params(oc, 7272, passes=2)
    assert np.array_equal(p.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("name", ["small", "reference", "full"])
def test_grid_corners_and_encoding_bitwise(nets, oracle, name):
    import npmg
    import pyoracle

    net = nets(name)
    oc = pyoracle.make_cfg(net.config)
    q = npmg.synth_queries(5, 3000)
    inp = _np(q["inputs"])
    # include the cube's faces, where the +1 corner wraps in dense levels
    inp[0:3, :8] = np.array([[0, 1, 0, 1, 0.5, 1, 0, 0.999999]] * 3, np.float32)
    q["inputs"].copy_(__import__("torch").from_numpy(inp))
    idx, w = net.grid_corners(q["inputs"])
    r_idx, r_w = oracle.grid_corners(oc, inp)
    assert np.array_equal(_np(idx).astype(np.uint32), r_idx)
    assert np.array_equal(_np(w).view(np.uint32), r_w.view(np.uint32))
    enc = net.encode(q["inputs"])
    p = _np(net.get_params(0))
    r_enc = oracle.encode(oc, p[net.n_mlp:], inp)
    assert np.array_equal(_np(enc).view(np.uint32), r_enc.view(np.uint32))


@pytest.mark.parametrize("name", CONFIGS)
def test_sample_pdf_bitwise(nets, oracle, name):
    import npmg
    import pyoracle

    net = nets(name)
    oc = pyoracle.make_cfg(net.config)
    n = 4000
    q = npmg.synth_queries(21, n)
    out = net.sample_pdf(q["inputs"], q["normal"], q["u"], q["eval_dirs"], cosine_aware=True, raw=True)
    h = {k: _np(v) for k, v in q.items()}
    p = _np(net.get_params(1))
    ref = oracle.query(oc, p, h["inputs"], h["normal"], h["u"], h["eval_dirs"])
    assert np.array_equal(_np(out["raw"]).view(np.uint32), ref["raw"].view(np.uint32)), "network outputs"
    assert np.array_equal(_np(out["lobe"]).astype(np.uint32), ref["lobe"]), "lobe selection"
    for k in ("dir", "pdf", "eval_pdf"):
        assert np.array_equal(_np(out[k]).view(np.uint32), ref[k].view(np.uint32)), k
    # against the FP64 gold: 1e-5 relative on the pdfs
    gold = oracle.query(oc, p, h["inputs"], h["normal"], h["u"], h["eval_dirs"], dtype=np.float64)
    rel = np.abs(_np(out["eval_pdf"]) - gold["eval_pdf"]) / gold["eval_pdf"]
    assert rel.max() <= 1e-5, rel.max()


@pytest.mark.parametrize("name", ["small", "small_refnum", "k16"])
@pytest.mark.parametrize("cosine", [False, True])
def test_mixture_consumer_bitwise(nets, oracle, name, cosine):
    import torch
    import npmg
    import pyoracle

    net = nets(name)
    oc = pyoracle.make_cfg(net.config)
    n = 4000
    q = npmg.synth_queries(31, n)
    rng = np.random.default_rng(1)
    raw = rng.normal(0, 2.0, size=(net.padded_output, n)).astype(np.float32)
    raw[:, :16] *= 8.0  # saturate the exp clamps and logistic tails
    raw_t = torch.from_numpy(raw).cuda()
    out = net.mixture_sample_pdf(raw_t, q["normal"] if cosine else None, q["u"], q["eval_dirs"],
                                 cosine_aware=cosine, lobes=True)
    h = {k: _np(v) for k, v in q.items()}
    ref = oracle.mixture(oc, raw, h["normal"] if cosine else None, h["u"], h["eval_dirs"])
    assert np.array_equal(_np(out["lobe"]).astype(np.uint32), ref["lobe"])
    for k in ("dir", "pdf", "eval_pdf", "lobes"):
        assert np.array_equal(_np(out[k]).view(np.uint32), ref[k].view(np.uint32)), k


@pytest.mark.parametrize("name", ["small", "small_refnum", "k16", "refnum_h16", "small_fast", "k16_fast"])
def test_adjoint_vs_fp64(nets, oracle, name):
    import torch
    import npmg
    import pyoracle

    net = nets(name)
    oc = pyoracle.make_cfg(net.config)
    n = 2048
    t = npmg.synth_train(3, n)
    rng = np.random.default_rng(2)
    raw = rng.normal(0, 1.0, size=(net.padded_output, n)).astype(np.float32)
    dout, loss = net.mixture_adjoint(torch.from_numpy(raw).cuda(), t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], 128.0)
    h = {k: _np(v) for k, v in t.items()}
    gd, gl = oracle.adjoint(oc, raw, h["tgt_dir"], h["tgt_rgb"], h["tgt_pdf"], 128.0, dtype=np.float64)
    d = _np(dout)
    scale = np.abs(gd).max(axis=0, keepdims=True) + 1e-30
    if net.config["loss"].get("fp16_intermediates"):
        # FP16 buffers on both sides: every output is an FP16 value, and a
        # rounding boundary crossed by the FP32-vs-FP64 difference costs one
        # FP16 step of a per-lobe dp or of the result (relative 2^-10 each)
        assert np.array_equal(d, d.astype(np.float16).astype(np.float32))
        assert np.max(np.abs(d - gd) / scale) < 2.0 ** -8
        assert np.mean(d == gd.astype(np.float32)) > 0.9
    else:
        assert np.max(np.abs(d - gd) / scale) < 1e-4
    # loss = -f/q log(p): an FP32 relative error e in p costs f/q e, which is
    # large against the loss itself where p is near 1 (log p near 0)
    fq = h["tgt_rgb"].astype(np.float64).mean(axis=0) / (h["tgt_pdf"].astype(np.float64) + 1e-5)
    assert np.max(np.abs(_np(loss) - gl) / (np.abs(gl) + 0.1 * fq)) < 1e-5


@pytest.mark.parametrize("name", ["small", "reference"])
def test_backward_vs_fp64(nets, oracle, name):
    import torch
    import npmg
    import pyoracle

    net = nets(name)
    oc = pyoracle.make_cfg(net.config)
    n = 1024
    q = npmg.synth_queries(8, n)
    rng = np.random.default_rng(3)
    dout = rng.normal(size=(net.padded_output, n)).astype(np.float32)
    dparams, denc = net.backward(q["inputs"], torch.from_numpy(dout).cuda(), want_denc=True)
    p = _np(net.get_params(0))
    inp = _np(q["inputs"])
    enc = oracle.encode(oc, p[net.n_mlp:], inp, dtype=np.float64)
    raw, acts = oracle.mlp_forward(oc, p[:net.n_mlp], enc, acts=True)
    dw, r_denc = oracle.mlp_backward(oc, p[:net.n_mlp], enc, acts, dout.astype(np.float64))
    gg = oracle.grid_backward(oc, inp, r_denc)
    g = _np(dparams)
    assert np.max(np.abs(g[:net.n_mlp] - dw)) / np.abs(dw).max() < 1e-5
    assert np.max(np.abs(g[net.n_mlp:] - gg)) / np.abs(gg).max() < 1e-5
    assert np.max(np.abs(_np(denc) - r_denc)) / np.abs(r_denc).max() < 1e-5


def test_train_step_matches_oracle(gpu, oracle):
    """One full step (forward, adjoint, backward, Adam, EMA) against the oracle's pieces."""
    import npmg
    import pyoracle

    cfg = npmg.reference_config()
    net = npmg.GuidingNetwork(cfg, seed=99)
    oc = pyoracle.make_cfg(cfg)
    n = 2048
    t = npmg.synth_train(4, n)
    p0 = _np(net.get_params(0)).copy()
    loss = net.train_step(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"])
    p1 = _np(net.get_params(0))
    e1 = _np(net.get_params(1))
    h = {k: _np(v) for k, v in t.items()}
    enc = oracle.encode(oc, p0[net.n_mlp:], h["inputs"], dtype=np.float64)
    raw, acts = oracle.mlp_forward(oc, p0[:net.n_mlp], enc, acts=True)
    dout, l = oracle.adjoint(oc, raw, h["tgt_dir"], h["tgt_rgb"], h["tgt_pdf"], 128.0, dtype=np.float64)
    dw, denc = oracle.mlp_backward(oc, p0[:net.n_mlp], enc, acts, dout)
    gg = oracle.grid_backward(oc, h["inputs"], denc)
    grad = np.concatenate([dw, gg]).astype(np.float32)
    master = p0.copy()
    m = np.zeros_like(master)
    v = np.zeros_like(master)
    steps = np.zeros(master.size, np.uint32)
    oracle.adam_step(net.n_mlp, 128.0, 1e-2, 0.9, 0.99, 1e-15, 1e-6, master, grad, m, v, steps)
    assert abs(loss.item() - l.mean()) / abs(l.mean()) < 1e-5
    # The first Adam step moves each parameter by about lr * sign(g): compare the step.
    d_gpu, d_ref = p1 - p0, master - p0
    big = np.abs(grad) > 1e-3 * np.abs(grad).max()
    assert np.mean(np.sign(d_gpu[big]) == np.sign(d_ref[big])) > 0.999
    assert np.max(np.abs(d_gpu - d_ref)) < 2.1e-2
    # EMA after the first step equals the weights (debiasing), tcnn ema.h:62-76
    assert np.allclose(e1, p1, rtol=1e-6, atol=1e-9)
    net.close()


def test_full_batch_properties(nets):
    """2^21 queries at the full configuration: size-independent invariants."""
    import torch
    import npmg

    net = nets("full")
    n = 1 << 21
    q = npmg.synth_queries(1234, n)
    a = net.sample_pdf(q["inputs"], q["normal"], q["u"], q["eval_dirs"], raw=True)
    b = net.sample_pdf(q["inputs"], q["normal"], q["u"], q["eval_dirs"], raw=True)
    for k in a:
        assert torch.equal(a[k], b[k]), f"non-deterministic {k}"
    assert torch.all(a["lobe"] >= 0) and torch.all(a["lobe"] < 8)
    assert torch.all(torch.isfinite(a["pdf"])) and torch.all(a["pdf"] > 0)
    assert torch.all(torch.isfinite(a["eval_pdf"])) and torch.all(a["eval_pdf"] > 0)
    norms = torch.linalg.norm(a["dir"], dim=0)
    assert torch.max(torch.abs(norms - 1)) < 1e-5
    # the consumer kernel on the same raw outputs reproduces the fused path bit for bit
    c = net.mixture_sample_pdf(a["raw"], q["normal"], q["u"], q["eval_dirs"])
    for k in ("dir", "pdf", "eval_pdf", "lobe"):
        assert torch.equal(a[k], c[k]), k
