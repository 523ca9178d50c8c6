"""The wavefront packers on the GPU against the oracle, bit for bit.

npmg_pack_queries (generate_inference_data, train.cu:174-191) and
npmg_pack_training (generate_training_data, train.cu:14-67, records in
(training pixel, depth) order): random shading points and a random frame of
recorded path vertices with delta vertices, NaNs, zero pdfs, tiny
throughputs and recorded depths above the array's, compared with the
oracle's FP32 restatement; the packed records then train the network.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX = [-3.0, -1.0, 0.5, 7.0, 2.0, 4.5]


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        nan = np.isnan(a)
        return np.array_equal(nan, np.isnan(b)) and np.array_equal(a[~nan].view(np.uint32), b[~nan].view(np.uint32))
    return np.array_equal(a, b)


def _frame(D, W, H, seed):
    rng = np.random.default_rng(seed)
    P = W * H
    u = lambda lo, hi, *s: rng.uniform(lo, hi, s).astype(np.float32)
    rec = {
        "L": u(0, 4, D, 3, P), "thp": u(0, 1, D, 3, P), "pos": u(-4, 8, D, 3, P),
        "dir": np.stack([u(0, np.pi, D, P), u(0, 2 * np.pi, D, P)], 1),
        "wi_pdf": u(0, 3, D, P), "mis_weight": u(0, 1, D, P),
        "delta": (rng.random((D, P)) < 0.1).astype(np.uint8), "bsdf_val": u(0, 2, D, 3, P),
        "wo": np.stack([u(0, np.pi, D, P), u(0, 2 * np.pi, D, P)], 1),
        "normal": np.stack([u(0, np.pi, D, P), u(0, 2 * np.pi, D, P)], 1),
        "roughness": u(0, 2, D, P), "depth": rng.integers(0, D + 3, P).astype(np.uint32),
    }
    rec["thp"][rng.random((D, 3, P)) < 0.02] = 1e-6
    rec["wi_pdf"][rng.random((D, P)) < 0.02] = 0
    for k, frac in (("pos", 0.005), ("L", 0.005), ("dir", 0.005), ("wo", 0.005), ("wi_pdf", 0.005),
                    ("bsdf_val", 0.005)):
        rec[k][rng.random(rec[k].shape) < frac] = np.nan
    return rec


def _to_gpu(rec):
    import torch

    return {k: torch.from_numpy(v.astype(np.int32) if k == "depth" else v).cuda() for k, v in rec.items()}


def test_pack_queries_bitwise(gpu, oracle):
    import torch
    import npmg

    rng = np.random.default_rng(3)
    n = 1 << 16
    pos = rng.uniform(-4, 8, (3, n)).astype(np.float32)
    wo = (rng.normal(size=(3, n)) * rng.uniform(0.1, 5, n)).astype(np.float32)
    nrm = (rng.normal(size=(3, n)) * rng.uniform(0.1, 5, n)).astype(np.float32)
    rough = rng.uniform(0, 3, n).astype(np.float32)
    g = npmg.pack_queries(*[torch.from_numpy(a).cuda() for a in (pos, wo, nrm, rough)], BOX)
    want = oracle.pack_queries(np.float32(BOX), pos, wo, nrm, rough)
    assert _same(g.cpu().numpy(), want)


@pytest.mark.parametrize("W,H,D,tile", [
    (64, 36, 8, dict(offset=(1, 0), stride=(2, 2), res=(32, 18), resolution=(64, 36))),
    (640, 360, 8, dict(offset=(1, 1), stride=(2, 2), res=(320, 180), resolution=(640, 360))),
    (100, 50, 4, dict(offset=(0, 0), stride=(1, 1), res=(100, 50), resolution=(100, 50))),
    (37, 23, 3, dict(offset=(2, 1), stride=(3, 2), res=(13, 12), resolution=(37, 23))),  # partly outside
])
def test_pack_training_bitwise(gpu, oracle, W, H, D, tile):
    import npmg

    rec = _frame(D, W, H, seed=W + D)
    net = npmg.GuidingNetwork(npmg.benchmark_config("small"), seed=1)
    want = oracle.pack_training(rec, tile, np.float32(BOX), W * H * D)
    cap = want[4] + 50
    got = net.pack_training(_to_gpu(rec), tile, BOX, cap)
    assert int(got[4].item()) == want[4]
    for g, w in zip(got[:4], want[:4]):
        g = g.cpu().numpy()
        assert _same(g[..., :want[4]], w[..., :want[4]])
        assert not g[..., want[4]:].any()
    # a second call gives the same bits (fixed order), and a short capacity keeps the prefix
    again = net.pack_training(_to_gpu(rec), tile, BOX, cap)
    for a, b in zip(again[:4], got[:4]):
        assert _same(a.cpu().numpy(), b.cpu().numpy())
    short = net.pack_training(_to_gpu(rec), tile, BOX, want[4] // 3)
    assert int(short[4].item()) == want[4]
    assert _same(short[0].cpu().numpy(), got[0].cpu().numpy()[:, :want[4] // 3])
    net.close()


def test_packed_records_train(gpu):
    """Records from the packer drive the train step and the query path."""
    import torch
    import npmg

    W, H, D = 256, 128, 4
    rec = _frame(D, W, H, seed=5)
    for k in ("pos", "L", "dir", "wo", "wi_pdf", "bsdf_val"):
        rec[k] = np.nan_to_num(rec[k], nan=0.5)
    tile = dict(offset=(0, 0), stride=(1, 1), res=(W, H), resolution=(W, H))
    net = npmg.GuidingNetwork(npmg.benchmark_config("small", precision="bf16"), seed=2)
    inp, tdir, rgb, pdf, cnt = net.pack_training(_to_gpu(rec), tile, BOX, W * H * D)
    n = int(cnt.item())
    assert n > 0
    loss = torch.empty(1, device="cuda")
    first = None
    for _ in range(20):
        net.train_step(inp[:, :n].contiguous(), tdir[:, :n].contiguous(), rgb[:, :n].contiguous(),
                       pdf[:n].contiguous(), loss)
        first = loss.item() if first is None else first
    assert np.isfinite(loss.item()) and loss.item() < first
    q = npmg.pack_queries(torch.rand(3, 4096, device="cuda") * 6, torch.randn(3, 4096, device="cuda"),
                          torch.randn(3, 4096, device="cuda"), torch.rand(4096, device="cuda"), BOX)
    out = net.sample_pdf(q, None, torch.rand(3, 4096, device="cuda"), cosine_aware=False)
    assert torch.isfinite(out["pdf"]).all()
    net.close()
