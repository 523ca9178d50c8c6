"""What the learned guiding buys: the trained network against the method
without it (cosine-weighted sampling), on the analytic radiance field the
network is trained on.

The integrand at a shading point is g(w) = f(x, w) max(0, n . w), with f the
field of three vMF lobes (kappa in [1, 200]) whose axes rotate with position.
Each method draws directions from its pdf p and forms g / p; both estimators
are unbiased, so their means must agree, and the guided one (the trained
mixture times the cosine lobe, sampled and evaluated by the library) must have
a much smaller relative variance. The untrained network is no better than
cosine sampling, which shows the gain comes from training.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELD_SEED = 31


def _frame_cosine_samples(nf, u1, u2):
    r = np.sqrt(u1)
    ph = 2 * np.pi * u2
    loc = np.stack([r * np.cos(ph), r * np.sin(ph), np.sqrt(np.maximum(0.0, 1 - u1))])
    # orthonormal frame about nf
    a = np.where(np.abs(nf[0]) > 0.9, 0.0, 1.0)
    t = np.stack([a, 1 - a, np.zeros_like(a)])
    b1 = np.cross(nf, t, axis=0)
    b1 /= np.linalg.norm(b1, axis=0)
    b2 = np.cross(nf, b1, axis=0)
    return b1 * loc[0] + b2 * loc[1] + nf * loc[2]


def _estimate(net, oracle, rec, nf, spp, rng, guided):
    import torch

    pos = rec["inputs"][:3]
    m = pos.shape[1]
    vals = np.zeros((spp, m))
    for s in range(spp):
        u = rng.random((3, m)).astype(np.float32)
        if guided:
            out = net.sample_pdf(torch.from_numpy(rec["inputs"]).cuda(), torch.from_numpy(nf.astype(np.float32)).cuda(),
                                 torch.from_numpy(u).cuda())
            w = out["dir"].cpu().numpy().astype(np.float64)
            p = out["pdf"].cpu().numpy().astype(np.float64)
        else:
            w = _frame_cosine_samples(nf, u[1].astype(np.float64), u[2].astype(np.float64))
            p = np.maximum((nf * w).sum(0), 0) / np.pi
        f = oracle.field_eval(FIELD_SEED, pos, w.astype(np.float32))
        g = f * np.maximum((nf * w).sum(0), 0.0)
        vals[s] = np.where(p > 0, g / np.maximum(p, 1e-30), 0.0)
    mean = vals.mean(0)
    var = vals.var(0)
    return mean, var


def test_trained_guiding_beats_cosine_sampling(gpu, oracle):
    import torch
    import npmg

    cfg = npmg.benchmark_config("small", precision="bf16")
    net = npmg.GuidingNetwork(cfg, seed=17)
    fresh = npmg.GuidingNetwork(cfg, seed=17)
    n_train = 1 << 16
    batch = npmg.synth_train(FIELD_SEED, n_train)
    loss = torch.empty(1, device="cuda")
    lib = npmg.lib()
    for k in range(400):
        lib.npmg_synth_train(FIELD_SEED, k * n_train, n_train, npmg._ptr(batch["inputs"]), npmg._ptr(batch["tgt_dir"]),
                             npmg._ptr(batch["tgt_rgb"]), npmg._ptr(batch["tgt_pdf"]), npmg._stream())
        net.train_step_graph(batch["inputs"], batch["tgt_dir"], batch["tgt_rgb"], batch["tgt_pdf"], loss)
    torch.cuda.synchronize()
    # shading points of the same field, not seen in training
    m = 1024
    rec = oracle.synth_train(FIELD_SEED, m, first=1 << 30)
    inp = rec["inputs"].astype(np.float64)
    th, ph = inp[6] * np.pi, inp[7] * 2 * np.pi
    n = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    wo = 2 * inp[3:6] - 1
    nf = np.where((n * wo).sum(0) > 0, n, -n)
    spp = 48
    rng = np.random.default_rng(4)
    mc, vc = _estimate(net, oracle, rec, nf, spp, rng, guided=False)
    mg, vg = _estimate(net, oracle, rec, nf, spp, rng, guided=True)
    m0, v0 = _estimate(fresh, oracle, rec, nf, spp, rng, guided=True)
    ok = mc > 1e-3 * mc.max()
    rel_c = float(np.mean(vc[ok] / mc[ok] ** 2))
    rel_g = float(np.mean(vg[ok] / mc[ok] ** 2))
    rel_0 = float(np.mean(v0[ok] / mc[ok] ** 2))
    print(f"relative variance: cosine {rel_c:.3f}  untrained {rel_0:.3f}  trained {rel_g:.3f}  "
          f"gain {rel_c / rel_g:.1f}x")
    # unbiased: the overall means agree within their pooled Monte Carlo error
    # (per-point z-scores are skewed by the heavy tail of the cosine estimator)
    z = (mg.mean() - mc.mean()) / np.sqrt((vc.mean() + vg.mean()) / (spp * mc.size))
    print(f"pooled z of the mean difference: {z:.2f}")
    assert abs(z) < 4.0
    assert rel_g < 0.5 * rel_c
    assert rel_0 > 0.5 * rel_c
    net.close()
    fresh.close()
