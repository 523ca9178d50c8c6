"""Learned selection probability on the GPU (7 lobes + the selection logit in
output 29 of 32; MixedSGWithSelection, distribution.h:204-241).

* 7-lobe FP32 queries bit for bit against the oracle (raw outputs, lobes,
  samples, pdfs) and the selection probability logistic(raw[28]) bit for bit,
  from the fused path, the queue path and the consumer entry point;
* the selection adjoint against the oracle's FP64 one (pinned by central
  differences in test_oracle_selection.py) and its FP32 restatement;
* training moves s the right way: towards the guide when the BSDF pdf of the
  records is 0, towards the BSDF when it dominates (FP32 and BF16, eager and
  graph);
* the training packer carries the BSDF pdf.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cfg(prec="fp32", mapping="reference"):
    import npmg

    c = npmg.benchmark_config("small", precision=prec)
    c["mixture"] = {"otype": "vMF", "n_lobes": 7, "mapping": mapping, "numerics": "stable", "learn_selection": True}
    return c


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("mapping", ["reference", "softmax"])
def test_selection_queries_match_oracle(gpu, oracle, mapping):
    import torch
    import npmg
    import pyoracle

    cfg = _cfg("fp32", mapping)
    net = npmg.GuidingNetwork(cfg, seed=11)
    assert net.padded_output == 32
    n = 4096
    q = npmg.synth_queries(3, n)
    out = net.sample_pdf_ex(q["inputs"], q["normal"], q["u"], q["eval_dirs"])
    raw = net.forward(q["inputs"], use_inference_params=True)
    oc = pyoracle.make_cfg(cfg)
    h = {k: _np(v) for k, v in q.items()}
    ref = oracle.query(oc, _np(net.get_params(1)), h["inputs"], h["normal"], h["u"], h["eval_dirs"])
    assert np.array_equal(_np(raw).view(np.uint32), ref["raw"].view(np.uint32))
    assert np.array_equal(_np(out["lobe"]).astype(np.uint32), ref["lobe"])
    for k in ("dir", "pdf", "eval_pdf"):
        assert np.array_equal(_np(out[k]).view(np.uint32), ref[k].view(np.uint32)), k
    sel = oracle.selection(oc, ref["raw"])
    assert np.array_equal(_np(out["selection"]).view(np.uint32), sel.view(np.uint32))
    assert np.array_equal(_np(net.selection_probability(raw)).view(np.uint32), sel.view(np.uint32))
    # the queue path writes the same probabilities
    cnt = torch.tensor([n, n // 3], dtype=torch.int32, device="cuda")
    outs = net.sample_pdf_queues([{"inputs": q["inputs"], "u": q["u"]}] * 2, cnt, selection=True)
    assert np.array_equal(_np(outs[0]["selection"]).view(np.uint32), sel.view(np.uint32))
    assert np.array_equal(_np(outs[1]["selection"])[:n // 3].view(np.uint32), sel[:n // 3].view(np.uint32))
    net.close()


@pytest.mark.parametrize("mapping", ["reference", "softmax"])
def test_selection_adjoint_matches_oracle(gpu, oracle, mapping):
    import torch
    import npmg
    import pyoracle

    cfg = _cfg("fp32", mapping)
    net = npmg.GuidingNetwork(cfg, seed=2)
    n = 2048
    rng = np.random.default_rng(4)
    raw = rng.normal(0, 1.0, (32, n)).astype(np.float32)
    t = npmg.synth_train(6, n)
    bsdf = rng.uniform(0, 0.6, n).astype(np.float32)
    bsdf[::7] = 0
    dout, loss = net.mixture_adjoint_selection(torch.from_numpy(raw).cuda(), t["tgt_dir"], t["tgt_rgb"],
                                               t["tgt_pdf"], torch.from_numpy(bsdf).cuda(), 128.0)
    oc = pyoracle.make_cfg(cfg)
    h = {k: _np(v) for k, v in t.items()}
    g64, l64 = oracle.adjoint_selection(oc, raw.astype(np.float64), h["tgt_dir"], h["tgt_rgb"], h["tgt_pdf"], bsdf,
                                        128.0, dtype=np.float64)
    g = _np(dout)
    scale = np.abs(g64).max(axis=0) + 1e-30
    err = np.abs(g - g64).max(axis=0) / scale
    print(f"{mapping}: adjoint vs FP64 max {err.max():.2e}, loss {np.abs(_np(loss) - l64).max():.2e}")
    assert err.max() < 1e-4
    assert np.allclose(_np(loss), l64, rtol=1e-5, atol=1e-7)
    assert not g[29:].any()
    net.close()


@pytest.mark.parametrize("prec,graph", [("fp32", False), ("bf16", False), ("bf16", True)])
def test_selection_training_moves_s(gpu, prec, graph):
    import torch
    import npmg

    results = []
    for b in (0.0, 50.0):  # BSDF pdf of every record: none, then far above the guide's
        net = npmg.GuidingNetwork(_cfg(prec), seed=3)
        t = npmg.synth_train(8, 1 << 14)
        bsdf = torch.full((1 << 14,), b, device="cuda")
        q = npmg.synth_queries(9, 4096)
        s0 = net.sample_pdf_ex(q["inputs"])["selection"].mean().item()
        loss = torch.empty(1, device="cuda")
        for _ in range(60):
            net.train_step_selection(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], bsdf, loss, graph=graph)
        s1 = net.sample_pdf_ex(q["inputs"])["selection"].mean().item()
        assert np.isfinite(loss.item())
        results.append((s0, s1))
        net.close()
    print(prec, graph, results)
    assert results[0][1] > results[0][0] + 0.01  # no BSDF mass: guide more
    assert results[1][1] < results[1][0] - 0.01  # BSDF dominates: guide less


def test_selection_requires_bsdf_pdfs(gpu):
    import torch
    import npmg

    net = npmg.GuidingNetwork(_cfg(), seed=3)
    t = npmg.synth_train(8, 1024)
    loss = torch.empty(1, device="cuda")
    with pytest.raises(npmg.NpmgError):
        net.train_step(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], loss)
    net.close()
    plain = npmg.GuidingNetwork(npmg.benchmark_config("small"), seed=3)
    with pytest.raises(npmg.NpmgError):
        plain.train_step_selection(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"],
                                   torch.zeros(1024, device="cuda"), loss)
    plain.close()
    c = npmg.benchmark_config("small")
    c["mixture"] = dict(c["mixture"], learn_selection=True)  # 8 lobes: no room for the logit
    with pytest.raises(npmg.NpmgError):
        npmg.GuidingNetwork(c)


def test_pack_training_carries_bsdf_pdf(gpu, oracle):
    import torch
    import npmg

    rng = np.random.default_rng(2)
    D, W, H = 4, 40, 30
    P = W * H
    u = lambda lo, hi, *s: rng.uniform(lo, hi, s).astype(np.float32)
    rec = {"L": u(0, 2, D, 3, P), "thp": u(0.1, 1, D, 3, P), "pos": u(-1, 1, D, 3, P),
           "dir": np.stack([u(0, 3, D, P), u(0, 6, D, P)], 1), "wi_pdf": u(0.1, 2, D, P),
           "mis_weight": u(0, 1, D, P), "delta": (rng.random((D, P)) < 0.2).astype(np.uint8),
           "bsdf_val": u(0, 1, D, 3, P), "wo": np.stack([u(0, 3, D, P), u(0, 6, D, P)], 1),
           "normal": np.stack([u(0, 3, D, P), u(0, 6, D, P)], 1), "roughness": u(0, 1, D, P),
           "depth": rng.integers(0, D + 1, P).astype(np.uint32), "bsdf_pdf": u(0, 3, D, P)}
    tile = dict(offset=(0, 1), stride=(1, 2), res=(W, H // 2), resolution=(W, H))
    box = [-1.0, -1.0, -1.0, 1.0, 1.0, 1.0]
    want = oracle.pack_training(rec, tile, np.float32(box), D * P)
    net = npmg.GuidingNetwork(npmg.benchmark_config("small"), seed=1)
    g = {k: torch.from_numpy(v.astype(np.int32) if k == "depth" else v).cuda() for k, v in rec.items()}
    got = net.pack_training(g, tile, box, D * P)
    cnt = want[4]
    assert int(got[4].item()) == cnt
    assert np.array_equal(_np(got[5])[:cnt], want[5][:cnt])
    assert np.array_equal(_np(got[0])[:, :cnt], want[0][:, :cnt])
    net.close()
