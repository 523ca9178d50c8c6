"""The train step captured as one HIP graph (npmg_train_step_graph).

Two contexts with the same seed train on the same batches, one eagerly and one
through the graph; with the deterministic grid-gradient mode every kernel is
order-independent, so parameters, EMA, loss, loss scale and step count must
agree bit for bit after every step. The graph is replayed several times with
the program's own eager work (a query) and synchronisations in between, and
captured anew when the arguments change.
"""
import pytest

pytestmark = pytest.mark.gpu


def _cfg(prec, kind):
    import npmg

    if kind == "asg-full":  # the per-frame train step of the anisotropic frame: 16 lobes x 6, MLP 3 x 128
        cfg = npmg.benchmark_config("full", precision=prec, width=128, hidden=3)
        cfg["mixture"] = {"otype": "ASG", "n_lobes": 16, "numerics": "fast"}
    else:
        cfg = npmg.benchmark_config("small", precision=prec)
    cfg["grid_gradient"] = "deterministic"
    cfg["optimizer"] = {"otype": "Ema", "decay": 0.95, "nested": cfg["optimizer"]}
    return cfg


@pytest.mark.parametrize("prec,kind,n", [("fp32", "vmf", 4096), ("bf16", "vmf", 4096), ("bf16x3", "vmf", 4096),
                                         ("bf16", "asg-full", 1 << 18)])
def test_graph_train_step_matches_eager(gpu, prec, kind, n):
    import torch
    import npmg

    cfg = _cfg(prec, kind)
    eager = npmg.GuidingNetwork(cfg, seed=3)
    graph = npmg.GuidingNetwork(cfg, seed=3)
    q = npmg.synth_queries(1, 4096)
    cosine = kind != "asg-full"
    la = torch.empty(1, device="cuda")
    lb = torch.empty(1, device="cuda")
    m = n - 1096  # a second batch size: the graph is captured again
    for nn, seed in ((n, 9), (n, 9), (n, 9), (m, 11), (m, 11)):
        t = npmg.synth_train(seed, nn)
        for _ in range(2):
            eager.train_step(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], la)
            graph.train_step_graph(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], lb)
            graph.sample_pdf(q["inputs"], q["normal"] if cosine else None, q["u"], q["eval_dirs"],
                             cosine_aware=cosine)  # eager work between replays
            torch.cuda.synchronize()
            assert la.item() == lb.item()
            assert torch.equal(eager.get_params(0), graph.get_params(0))
            assert torch.equal(eager.get_params(1), graph.get_params(1))
    assert eager.train_state() == graph.train_state()
    eager.close()
    graph.close()


def test_graph_survives_workspace_growth(gpu):
    """A larger batch on the eager path reallocates the workspace; the next
    graph call must capture again rather than replay into freed memory."""
    import torch
    import npmg

    cfg = npmg.benchmark_config("small", precision="bf16")
    cfg["grid_gradient"] = "deterministic"
    a = npmg.GuidingNetwork(cfg, seed=5)
    b = npmg.GuidingNetwork(cfg, seed=5)
    small = npmg.synth_train(2, 2048)
    big = npmg.synth_train(3, 8192)
    la = torch.empty(1, device="cuda")
    lb = torch.empty(1, device="cuda")
    for net, loss in ((a, la), (b, lb)):
        net.train_step_graph(small["inputs"], small["tgt_dir"], small["tgt_rgb"], small["tgt_pdf"], loss)
        net.train_step(big["inputs"], big["tgt_dir"], big["tgt_rgb"], big["tgt_pdf"], loss)
    a.train_step_graph(small["inputs"], small["tgt_dir"], small["tgt_rgb"], small["tgt_pdf"], la)
    b.train_step(small["inputs"], small["tgt_dir"], small["tgt_rgb"], small["tgt_pdf"], lb)
    torch.cuda.synchronize()
    assert la.item() == lb.item()
    assert torch.equal(a.get_params(0), b.get_params(0))
    a.close()
    b.close()
