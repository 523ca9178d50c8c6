"""npmg_sample_pdf_noalloc: the fused query path without any device
allocation (its grid-feature scratch is a file-scope device array), called
through the C ABI and compared bit for bit with npmg_sample_pdf on the same
inputs, on two calls in a row (the second smaller, so it reuses the first
call's scratch contents), for the split-grid MFMA precisions and FP32. A
context that never ran npmg_sample_pdf serves both calls without a change in
the device's free memory (another context ran the kernels first, since the
runtime loads a kernel's code on its first launch). Sizes beyond the scratch are rejected
with an error and leave the outputs untouched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(lib, fn, net, q, n, out, stream):
    import npmg

    P = npmg._ptr
    st = fn(net._h, P(q["inputs"]), P(q["normal"]), P(q["u"]), P(q["eval_dirs"]), 4, P(out["dir"]), P(out["pdf"]),
            P(out["eval_pdf"]), P(out["lobe"]), P(out["raw"]), n, npmg.NPMG_COSINE_AWARE, stream)
    return st


def _outputs(net, n):
    import torch

    return {"dir": torch.full((3, n), -1.0, device="cuda"), "pdf": torch.full((n,), -1.0, device="cuda"),
            "eval_pdf": torch.full((4, n), -1.0, device="cuda"),
            "lobe": torch.full((n,), -1, dtype=torch.int32, device="cuda"),
            "raw": torch.full((net.padded_output, n), -1.0, device="cuda")}


@pytest.mark.parametrize("prec", ["bf16", "bf16x3", "fp16", "fp32"])
def test_noalloc_matches_sample_pdf(gpu, prec):
    import torch
    import npmg

    lib = npmg.lib()
    stream = torch.cuda.current_stream().cuda_stream
    cfg = npmg.benchmark_config("full", precision=prec)
    cfg["mixture"] = dict(cfg["mixture"], numerics="fast")
    sizes = (1 << 18, 3000)  # the largest batch the scratch holds at 32 grid features, then a smaller one
    qs = [npmg.synth_queries(50 + k, n) for k, n in enumerate(sizes)]
    # another context runs the same kernels first: the HIP runtime loads a kernel's
    # code on its first launch in a process, which takes device memory
    warm = npmg.GuidingNetwork(cfg, seed=1)
    for q, n in zip(qs, sizes):
        assert _call(lib, lib.npmg_sample_pdf, warm, q, n, _outputs(warm, n), stream) == 0
    net = npmg.GuidingNetwork(cfg, seed=7272)  # never runs npmg_sample_pdf before the checked calls
    got = [_outputs(net, n) for n in sizes]
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for q, n, out in zip(qs, sizes, got):
        assert _call(lib, lib.npmg_sample_pdf_noalloc, net, q, n, out, stream) == 0, \
            lib.npmg_last_error(net._h).decode()
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0, "npmg_sample_pdf_noalloc changed the device's free memory"
    for q, n, out in zip(qs, sizes, got):
        want = _outputs(net, n)
        assert _call(lib, lib.npmg_sample_pdf, net, q, n, want, stream) == 0
        for k in want:
            assert torch.equal(out[k].view(torch.int32) if out[k].dtype == torch.float32 else out[k],
                               want[k].view(torch.int32) if want[k].dtype == torch.float32 else want[k]), (prec, n, k)
    net.close()
    warm.close()


def test_noalloc_rejects_sizes_beyond_scratch(gpu):
    import torch
    import npmg

    lib = npmg.lib()
    stream = torch.cuda.current_stream().cuda_stream
    net = npmg.GuidingNetwork(npmg.benchmark_config("full", precision="bf16"), seed=1)
    n = (1 << 18) + 64
    q = npmg.synth_queries(7, n)
    out = _outputs(net, n)
    st = _call(lib, lib.npmg_sample_pdf_noalloc, net, q, n, out, stream)
    assert st == 1  # NPMG_EINVAL
    assert "16 MiB" in lib.npmg_last_error(net._h).decode()
    torch.cuda.synchronize()
    assert torch.all(out["pdf"] == -1.0), "a rejected call wrote outputs"
    net.close()


def test_noalloc_fp8_needs_calibration(gpu):
    import torch
    import npmg

    lib = npmg.lib()
    stream = torch.cuda.current_stream().cuda_stream
    cfg = npmg.benchmark_config("full", precision="bf16")
    cfg["inference_precision"] = "fp8"
    net = npmg.GuidingNetwork(cfg, seed=2)
    n = 4096
    q = npmg.synth_queries(9, n)
    out = _outputs(net, n)
    assert _call(lib, lib.npmg_sample_pdf_noalloc, net, q, n, out, stream) != 0
    assert "calibrat" in lib.npmg_last_error(net._h).decode()
    net.fp8_calibrate(q["inputs"])
    assert _call(lib, lib.npmg_sample_pdf_noalloc, net, q, n, out, stream) == 0
    want = _outputs(net, n)
    assert _call(lib, lib.npmg_sample_pdf, net, q, n, want, stream) == 0
    assert torch.equal(out["eval_pdf"], want["eval_pdf"]) and torch.equal(out["lobe"], want["lobe"])
    net.close()
