"""Persistent launches over uneven bounce queues (npmg_sample_pdf_queues).

A wavefront frame holds one queue per bounce, each shorter than the last
(the guided passes of integrator.cpp:160-320 launch over each). One
persistent launch serves all of them from a device work counter. Every
query's result must be bit-identical to a plain npmg_sample_pdf call over
that queue alone: the tile a query lands in and the launch shape change
nothing in its arithmetic. Also covered: empty queues, a count above the
capacity (clamped), rows past the count left untouched, the 16-queue limit,
and the autotuner, whose chosen launch shape must not change any result.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [("fp32", "vMF"), ("bf16", "vMF"), ("fp16", "vMF"), ("bf16x3", "vMF"), ("fp32", "ASG"), ("bf16", "ASG")]
FRAME_PATHS = 2073600  # one 1920 x 1080 bounce


def _net(prec, mix, size="small", numerics=None):
    import torch
    import npmg

    if mix == "ASG":
        c = npmg.benchmark_config(size, precision=prec, width=128, hidden=3)
        c["mixture"] = {"otype": "ASG", "n_lobes": 16}
        if numerics:
            c["mixture"]["numerics"] = numerics
    else:
        c = npmg.benchmark_config(size, precision=prec)
    net = npmg.GuidingNetwork(c, seed=5)
    t = npmg.synth_train(3, 1 << 14)
    loss = torch.empty(1, device="cuda")
    for _ in range(5):
        net.train_step(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], loss)
    return net


def _queues(counts, caps, seed, cosine):
    import npmg

    qs = []
    for i, (n, cap) in enumerate(zip(counts, caps)):
        q = npmg.synth_queries(seed + i, cap)
        qs.append({"inputs": q["inputs"], "u": q["u"], "eval_dirs": q["eval_dirs"],
                   "normal": q["normal"] if cosine else None})
    return qs


def _check_queue(net, q, o, n, cap, cosine, n_eval):
    ref = net.sample_pdf(q["inputs"][:, :n].contiguous(),
                         q["normal"][:, :n].contiguous() if cosine else None,
                         q["u"][:, :n].contiguous(), q["eval_dirs"][:, :, :n].contiguous(),
                         cosine_aware=cosine) if n else None
    for key, width in (("dir", 3), ("pdf", 0), ("lobe", 0), ("eval_pdf", n_eval)):
        got = o[key].cpu().numpy()
        live = got[..., :n]
        if n:
            want = ref[key].cpu().numpy()
            assert np.array_equal(live.view(np.uint32), want.view(np.uint32)), key
        assert not got[..., n:].any(), f"{key}: rows past the count were written"


@pytest.mark.parametrize("prec,mix", CASES)
def test_queues_match_single_calls(gpu, prec, mix):
    import torch

    net = _net(prec, mix)
    cosine = mix == "vMF"
    counts = [9000, 6120, 0, 4161, 2830, 1924, 63, 1, 1000]
    caps = [9000 + 37, 6200, 64, 4161, 2900, 1924 + 64, 64, 5, 800]  # the last count exceeds its capacity
    qs = _queues(counts, caps, 100, cosine)
    cnt = torch.tensor(counts, dtype=torch.int32, device="cuda")
    outs = net.sample_pdf_queues(qs, cnt, n_eval=4, cosine_aware=cosine)
    torch.cuda.synchronize()
    for q, o, n, cap in zip(qs, outs, counts, caps):
        _check_queue(net, q, o, min(n, cap), cap, cosine, 4)
    net.close()


def test_queues_full_frame_bf16(gpu):
    """An 8-bounce frame at 1/16 of 1920x1080 paths with 0.68 survival per bounce."""
    import torch

    net = _net("bf16", "vMF", "full")
    counts = [int(129600 * 0.68 ** b) for b in range(8)]
    caps = [129600] * 8
    qs = _queues(counts, caps, 300, True)
    cnt = torch.tensor(counts, dtype=torch.int32, device="cuda")
    outs = net.sample_pdf_queues(qs, cnt, n_eval=4)
    torch.cuda.synchronize()
    for q, o, n, cap in zip(qs, outs, counts, caps):
        _check_queue(net, q, o, n, cap, True, 4)
    # a second frame through the same counters with other counts
    counts2 = [c // 3 for c in counts]
    cnt.copy_(torch.tensor(counts2, dtype=torch.int32))
    outs = net.sample_pdf_queues(qs, cnt, n_eval=4)
    torch.cuda.synchronize()
    for q, o, n, cap in zip(qs, outs, counts2, caps):
        _check_queue(net, q, o, n, cap, True, 4)
    net.close()


def test_queues_full_size_anisotropic_frame(gpu):
    """The anisotropic frame at its real size: 8 bounce queues of
    floor(2,073,600 x 0.68^b) queries (6,183,755 in all, capacity 2,073,600
    each) in one persistent launch of the 128-wide BF16 network with 16
    anisotropic lobes, bit for bit equal to one npmg_sample_pdf per queue."""
    import torch

    net = _net("bf16", "ASG", "full", numerics="fast")
    counts = [int(FRAME_PATHS * 0.68 ** b) for b in range(8)]
    assert sum(counts) == 6183755
    caps = [FRAME_PATHS] * 8
    qs = _queues(counts, caps, 900, False)
    cnt = torch.tensor(counts, dtype=torch.int32, device="cuda")
    outs = net.sample_pdf_queues(qs, cnt, n_eval=4, cosine_aware=False)
    torch.cuda.synchronize()
    for q, o, n, cap in zip(qs, outs, counts, caps):
        ref = net.sample_pdf(q["inputs"][:, :n].contiguous(), None, q["u"][:, :n].contiguous(),
                             q["eval_dirs"][:, :, :n].contiguous(), cosine_aware=False)
        for key in ("dir", "pdf", "lobe", "eval_pdf"):
            assert torch.equal(o[key][..., :n].contiguous().view(torch.int32),
                               ref[key].view(torch.int32)), (n, key)
            assert not torch.any(o[key][..., n:] != 0), f"{key}: rows past the count were written"
        del ref
    net.close()


def test_queue_limits(gpu):
    import torch
    import npmg

    net = _net("bf16", "ASG")
    qs = _queues([64] * 17, [64] * 17, 7, False)
    cnt = torch.full((17,), 64, dtype=torch.int32, device="cuda")
    with pytest.raises(npmg.NpmgError):
        net.sample_pdf_queues(qs, cnt)
    qs = _queues([64], [64], 7, True)
    with pytest.raises(npmg.NpmgError):  # the anisotropic mixture has no cosine product
        net.sample_pdf_queues(qs, cnt[:1], cosine_aware=True)
    net.close()


@pytest.mark.parametrize("prec", ["bf16", "bf16x3"])
def test_lds_levels_match_gathered(gpu, prec):
    """Batches of 2^20 queries or more gather the coarse levels whose tables
    fit in LDS (levels 0 and 1 of the full configuration) from a copy staged
    there, pipelined or in one launch; a queue launch gathers every level from
    the table in memory. The features, and so every output, are the same bits."""
    import torch

    net = _net(prec, "vMF", "full")
    n = (1 << 20) + 4099  # a ragged tail past whole 1024-query workgroups
    (q,) = _queues([n], [n], 77, True)
    cnt = torch.tensor([n], dtype=torch.int32, device="cuda")
    (ref,) = net.sample_pdf_queues([q], cnt, n_eval=4)
    torch.cuda.synchronize()
    for shape in ((1, 0), (4, 0), (3, 2)):
        net.set_tuning(*shape)
        _check_queue(net, q, ref, n, n, True, 4)
    net.close()


@pytest.mark.parametrize("prec,mix", [("bf16", "vMF"), ("bf16x3", "vMF"), ("bf16", "ASG")])
def test_autotune_keeps_results(gpu, prec, mix):
    import torch
    import npmg

    net = _net(prec, mix, "full")
    cosine = mix == "vMF"
    n = 300000
    q = npmg.synth_queries(11, n)
    nrm = q["normal"] if cosine else None
    before = net.sample_pdf(q["inputs"], nrm, q["u"], q["eval_dirs"], cosine_aware=cosine)
    chunks, bpc = net.autotune(n)
    print("tuned", prec, mix, "chunks", chunks, "blocks/CU", bpc)
    assert chunks in (1, 2, 4, 8) and bpc in (0, 1, 2, 3)
    after = net.sample_pdf(q["inputs"], nrm, q["u"], q["eval_dirs"], cosine_aware=cosine)
    torch.cuda.synchronize()
    for k in before:
        assert torch.equal(before[k], after[k]), k
    # shapes set by hand: every one gives the same bits; out-of-range ones are rejected
    for shape in ((1, 0), (8, 1), (3, 2), (2, 64)):
        net.set_tuning(*shape)
        assert net.tuning() == shape
        again = net.sample_pdf(q["inputs"], nrm, q["u"], q["eval_dirs"], cosine_aware=cosine)
        for k in before:
            assert torch.equal(before[k], again[k]), (shape, k)
    for bad in ((0, 0), (9, 0), (4, 65)):
        with pytest.raises(npmg.NpmgError):
            net.set_tuning(*bad)
    assert net.tuning() == (2, 64)
    # the queue launches' claim sizes: the tuned ones and others give the same bits
    tiles, items = net.queue_tuning()
    print("queue claims", tiles, items)
    assert 1 <= tiles <= 16 and 8 <= items <= 64
    counts = [3000, 2040, 1387, 0, 640]
    qs = _queues(counts, [3100] * 5, 40, cosine)
    cnt = torch.tensor(counts, dtype=torch.int32, device="cuda")
    ref = net.sample_pdf_queues(qs, cnt, n_eval=4, cosine_aware=cosine)
    for shape in ((1, 1), (64, 256), (3, 7)):
        net.set_queue_tuning(*shape)
        assert net.queue_tuning() == shape
        got = net.sample_pdf_queues(qs, cnt, n_eval=4, cosine_aware=cosine)
        for a, b in zip(ref, got):
            for k in a:
                assert torch.equal(a[k], b[k]), (shape, k)
    for bad in ((0, 16), (65, 16), (4, 0), (4, 257)):
        with pytest.raises(npmg.NpmgError):
            net.set_queue_tuning(*bad)
    net.close()
