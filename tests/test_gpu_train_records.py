"""npmg_train_records: trainStep's batch loop over packed records
(integrator.cpp:660-702) gives the same parameters, bit for bit, as training
the same batches one by one with npmg_train_step."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cfg(selection):
    import npmg

    c = npmg.benchmark_config("small", precision="bf16")
    c["grid_gradient"] = "deterministic"
    if selection:
        c["mixture"] = {"otype": "vMF", "n_lobes": 7, "mapping": "softmax", "learn_selection": True}
    return c


@pytest.mark.parametrize("selection", [False, True])
def test_train_records_equals_manual_batches(gpu, selection):
    import torch
    import npmg

    n, cap = 300000, 320000
    t = npmg.synth_train(12, n)
    # packer-style buffers: row stride cap, n records kept
    pad = lambda x: torch.cat([x, torch.zeros(*x.shape[:-1], cap - n, device="cuda")], dim=-1).contiguous()
    rec = {k: pad(v) for k, v in t.items()}
    bsdf = pad(torch.rand(n, device="cuda")) if selection else None
    a = npmg.GuidingNetwork(_cfg(selection), seed=4)
    b = npmg.GuidingNetwork(_cfg(selection), seed=4)
    policy = dict(batch_size=131072, batches_per_frame=8, granularity=128, min_batch=65536)
    batches = npmg.train_batches(n, **policy)
    assert batches == [(0, 131072), (131072, 131072)]  # the third (37 856 -> 37 760) is below the minimum
    losses = torch.zeros(8, device="cuda")
    nb = a.train_records(rec["inputs"], rec["tgt_dir"], rec["tgt_rgb"], rec["tgt_pdf"], n, tgt_bsdf_pdf=bsdf,
                         losses=losses, **policy)
    assert nb == len(batches)
    loss = torch.empty(1, device="cuda")
    manual = []
    for o, m in batches:
        args = [t[k][..., o:o + m].contiguous() for k in ("inputs", "tgt_dir", "tgt_rgb", "tgt_pdf")]
        if selection:
            b.train_step_selection(*args, bsdf[o:o + m].contiguous(), loss)
        else:
            b.train_step(*args, loss)
        manual.append(loss.item())
    assert np.array_equal(losses[:nb].cpu().numpy(), np.float32(manual))
    for which in (0, 1):
        assert torch.equal(a.get_params(which), b.get_params(which))
    a.close()
    b.close()
