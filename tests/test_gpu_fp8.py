"""FP8 (E4M3) inference path: "inference_precision": "fp8" with BF16 training.

The query path's MLP runs on E4M3 weights and activations with FP32
accumulation (v_mfma_f32_32x32x16_fp8_fp8), with power-of-two scales per
layer: weights from max |W_l|, inputs from the calibrated maximum activation.
Checked here:

* the device conversion is round-to-nearest-even OCP E4M3 (bitwise against
  torch's float8_e4m3fn on every representable value, every tie and random
  values, subnormals included);
* the scales are powers of two with max |W_l| sw_l in (224, 448];
* raw outputs match a numpy emulation of the rounding points (BF16 table and
  grid features, E4M3 weights and layer inputs with the library's scales,
  products summed in FP64) -- the only differences left are FP32 summation
  order and the rare activation whose rounding flips;
* the mixture stage is the FP32 consumer kernel bit for bit on those outputs;
* end to end against the FP64 oracle the pdf error is measured and bounded
  loosely (the cost of three mantissa bits, reported by the benchmark);
* the fragments follow the inference parameters through training steps, and
  the persistent queue launch gives the same bits as single calls.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _fp8(a):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a, np.float32))
    return t.to(torch.float8_e4m3fn).to(torch.float64).numpy()


def _bf16(a):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a, np.float32))
    return t.to(torch.bfloat16).to(torch.float32).numpy()


def _cfg(size="small", width=64, n_lobes=8):
    import npmg

    c = npmg.benchmark_config(size, precision="bf16", width=width, n_lobes=n_lobes)
    c["inference_precision"] = "fp8"
    return c


def _train(net, steps, seed=3):
    import torch
    import npmg

    t = npmg.synth_train(seed, 1 << 14)
    loss = torch.empty(1, device="cuda")
    for _ in range(steps):
        net.train_step(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], loss)


def _emulate(net, cfg, oracle, inputs):
    """Raw outputs of the FP8 path in FP64 from its rounding points."""
    import pyoracle

    oc = pyoracle.make_cfg(cfg)
    p = _np(net.get_params(1)).astype(np.float32)
    n_mlp = net.n_mlp
    grid = cfg["encoding"]["nested"][0]
    lf = grid["n_levels"] * grid["n_features_per_level"]
    enc = oracle.encode(oc, _bf16(p[n_mlp:]), inputs).astype(np.float32)
    enc[:lf] = _bf16(enc[:lf])  # level-major grid features are stored in BF16
    sw, sx = net.fp8_scales()
    x = _fp8(np.clip(enc.astype(np.float64) * sx[0], -448, 448))
    off, n_in = 0, net.encoded_width
    L = oc.n_hidden_layers + 1
    for l in range(L):
        n_out = oc.width if l < L - 1 else net.padded_output
        W = p[off:off + n_out * n_in].reshape(n_out, n_in).astype(np.float64)
        off += n_out * n_in
        n_in = oc.width
        acc = _fp8(np.clip(W * sw[l], -448, 448)) @ x
        if l < L - 1:
            m = sx[l + 1] / (sw[l] * sx[l])
            x = _fp8(np.clip(np.maximum(acc, 0) * m, 0, 448))
        else:
            return acc / (sw[l] * sx[l]), (sw, sx)


def test_fp8_conversion_is_ocp_e4m3_rne(gpu):
    import torch
    import npmg

    codes = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).to(torch.float32).numpy()
    vals = np.unique(codes[np.isfinite(codes)].astype(np.float64))
    pos = vals[vals >= 0]
    mids = (pos[1:] + pos[:-1]) / 2  # ties: round to the even code
    rng = np.random.default_rng(0)
    x = np.concatenate([vals, mids, -mids, np.nextafter(mids, 0), np.nextafter(mids, 1e9),
                        rng.uniform(-448, 448, 20000), rng.uniform(-2 ** -6, 2 ** -6, 20000),
                        2.0 ** rng.uniform(-12, 8.8, 20000)]).astype(np.float32)
    x = np.clip(x, -448, 448).astype(np.float32)
    got = _np(npmg.debug_math(11, torch.from_numpy(x).cuda()))
    want = _fp8(x).astype(np.float32)
    bad = got.view(np.uint32) != want.view(np.uint32)
    assert not bad.any(), (x[bad][:5], got[bad][:5], want[bad][:5])


def test_fp8_scales(gpu):
    import npmg

    cfg = _cfg()
    net = npmg.GuidingNetwork(cfg, seed=11)
    _train(net, 5)
    q = npmg.synth_queries(4, 4096)
    net.sample_pdf(q["inputs"], q["normal"], q["u"], q["eval_dirs"])  # calibrates on these queries
    sw, sx = net.fp8_scales()
    p = _np(net.get_params(1))
    off, n_in = 0, net.encoded_width
    for l, s in enumerate(sw):
        n_out = cfg["network"]["n_neurons"] if l < len(sw) - 1 else net.padded_output
        wmax = np.abs(p[off:off + n_out * n_in]).max()
        off += n_out * n_in
        n_in = cfg["network"]["n_neurons"]
        assert np.log2(s) == np.round(np.log2(s)), s
        assert 224 < wmax * s <= 448, (l, wmax, s)
    for s in sx:
        assert np.log2(s) == np.round(np.log2(s)), s
    net.close()


@pytest.mark.parametrize("size,width,n_lobes", [("small", 64, 8), ("full", 64, 8), ("small", 128, 16)])
def test_fp8_query_matches_emulation(gpu, oracle, size, width, n_lobes):
    import npmg
    import pyoracle

    cfg = _cfg(size, width, n_lobes)
    net = npmg.GuidingNetwork(cfg, seed=7272)
    _train(net, 10)
    n = 4096
    q = npmg.synth_queries(41, n)
    out = net.sample_pdf(q["inputs"], q["normal"], q["u"], q["eval_dirs"], raw=True)
    h = {k: _np(v) for k, v in q.items()}
    emu, (sw, sx) = _emulate(net, cfg, oracle, h["inputs"])
    raw = _np(out["raw"])
    scale = np.abs(emu).max(axis=0) + 1e-6
    err = np.abs(raw - emu).max(axis=0) / scale
    print(f"{size} w{width} k{n_lobes}: weight scales {sw}, input scales {sx}; "
          f"raw error vs emulation p95 {np.quantile(err, 0.95):.2e} max {err.max():.2e}")
    assert np.quantile(err, 0.95) < 2e-5, np.quantile(err, 0.95)
    assert err.max() < 0.2, err.max()
    # the mixture stage is the FP32 consumer kernel on the kernel's own raw outputs
    c = net.mixture_sample_pdf(out["raw"], q["normal"], q["u"], q["eval_dirs"])
    for k in ("dir", "pdf", "eval_pdf", "lobe"):
        assert np.array_equal(_np(out[k]).view(np.uint32), _np(c[k]).view(np.uint32)), k
    # end to end against the FP64 oracle with the unrounded parameters
    oc = pyoracle.make_cfg(cfg)
    p = _np(net.get_params(1)).astype(np.float32)
    gold = oracle.query(oc, p, h["inputs"], h["normal"], h["u"], h["eval_dirs"], dtype=np.float64)
    rel = np.abs(_np(out["eval_pdf"]) - gold["eval_pdf"]) / gold["eval_pdf"]
    print(f"  pdf vs FP64: median {np.median(rel):.2e} p90 {np.quantile(rel, 0.9):.2e}")
    assert np.median(rel) < 0.1, np.median(rel)
    net.close()


def test_fp8_follows_training(gpu, oracle):
    """After calibration, further steps repack the E4M3 fragments (new weight
    scales) while the activation scales stay calibrated."""
    import npmg

    cfg = _cfg()
    net = npmg.GuidingNetwork(cfg, seed=5)
    _train(net, 3)
    q = npmg.synth_queries(8, 2048)
    net.fp8_calibrate(q["inputs"])
    _, sx0 = net.fp8_scales()
    _train(net, 20, seed=9)
    _, sx1 = net.fp8_scales()
    assert sx0 == sx1
    out = net.forward(q["inputs"], use_inference_params=True)
    emu, _ = _emulate(net, cfg, oracle, _np(q["inputs"]))
    err = np.abs(_np(out) - emu).max(axis=0) / (np.abs(emu).max(axis=0) + 1e-6)
    assert np.quantile(err, 0.95) < 2e-5, np.quantile(err, 0.95)
    net.close()


def test_fp8_queues_match_single_calls(gpu):
    import torch
    import npmg

    net = npmg.GuidingNetwork(_cfg("full"), seed=5)
    _train(net, 5)
    counts = [20000, 13600, 9248, 0, 64, 1]
    qs = []
    for i, n in enumerate(counts):
        q = npmg.synth_queries(60 + i, n + 100)
        qs.append({"inputs": q["inputs"], "normal": q["normal"], "u": q["u"], "eval_dirs": q["eval_dirs"]})
    cnt = torch.tensor(counts, dtype=torch.int32, device="cuda")
    with pytest.raises(npmg.NpmgError):  # not calibrated yet
        net.sample_pdf_queues(qs, cnt, n_eval=4)
    net.fp8_calibrate(qs[0]["inputs"])
    outs = net.sample_pdf_queues(qs, cnt, n_eval=4)
    for q, o, n in zip(qs, outs, counts):
        if not n:
            continue
        ref = net.sample_pdf(q["inputs"][:, :n].contiguous(), q["normal"][:, :n].contiguous(),
                             q["u"][:, :n].contiguous(), q["eval_dirs"][:, :, :n].contiguous())
        for k in ("dir", "pdf", "lobe", "eval_pdf"):
            assert np.array_equal(_np(o[k])[..., :n].view(np.uint32), _np(ref[k]).view(np.uint32)), k
    net.close()


def test_fp8_config_rules(gpu):
    import npmg

    c = npmg.benchmark_config("small", precision="fp32")
    c["inference_precision"] = "fp8"
    with pytest.raises(npmg.NpmgError):
        npmg.GuidingNetwork(c)
    c = npmg.benchmark_config("small", precision="bf16")
    c["precision"] = "fp8"
    with pytest.raises(npmg.NpmgError):
        npmg.GuidingNetwork(c)
