"""The anisotropic lobe mixture (no counterpart in the reference, so it is
pinned by closed forms): each lobe is a vMF pushed through
M = diag(a, 1/a, 1) in a frame rotated by psi about its axis.

* normalisation: FP64 sphere quadrature of the library's FP32 pdf is 1;
* a = 1 reduces every lobe to the vMF lobe with the same axis and kappa
  (checked against a numpy FP64 vMF mixture);
* sampling: chi-square of drawn directions against the pdf;
* adjoint: FP64 central differences of the FP64 restatement of the loss;
* the MFMA path reproduces the FP32 consumer kernel on its own raw outputs,
  and training lowers the loss in FP32 and BF16.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NZ, NP, SUB = 32, 64, 8


def _cfg(prec="fp32", size="small", width=128):
    import npmg

    c = npmg.benchmark_config(size, precision=prec, width=width, hidden=3)
    c["mixture"] = {"otype": "ASG", "n_lobes": 16}
    c["loss"] = dict(c["loss"], grad_clamp=0.0)
    return c


def _raw(n, seed, kmax=300.0, log_a=2.5, a_one=False):
    r = np.random.default_rng(seed)
    raw = r.normal(0, 1, size=(96, n))
    raw[3:96:6] = r.uniform(np.log(1e-2), np.log(kmax), size=(16, n))
    raw[4:96:6] = 0.0 if a_one else r.uniform(-log_a, log_a, size=(16, n))
    return raw.astype(np.float32)


def test_asg_pdf_integrates_to_one(gpu):
    import torch
    import npmg

    net = npmg.GuidingNetwork(_cfg(), seed=1)
    n = 32
    raw = torch.from_numpy(_raw(n, 3)).cuda()
    integ = net.pdf_integral(raw, None, n_z=2048, n_phi=4096).cpu().numpy()
    print("max |I - 1|:", np.abs(integ - 1).max())
    assert np.abs(integ - 1).max() < 1e-4
    net.close()


def test_asg_with_unit_anisotropy_is_the_vmf_mixture(gpu):
    import torch
    import npmg

    net = npmg.GuidingNetwork(_cfg(), seed=1)
    n = 4096
    raw = _raw(n, 5, kmax=1e3, a_one=True)
    q = npmg.synth_queries(9, n)
    out = net.mixture_sample_pdf(torch.from_numpy(raw).cuda(), None, None, q["eval_dirs"], cosine_aware=False)
    got = out["eval_pdf"].cpu().numpy().astype(np.float64)
    x = raw.astype(np.float64)
    sig = lambda v: 1 / (1 + np.exp(-v))
    th, ph = np.pi * sig(x[0::6][:16]), 2 * np.pi * sig(x[1::6][:16])
    kap = np.clip(np.exp(x[3::6][:16]), 1e-2, 1e3)
    logit = x[5::6][:16]
    w = np.exp(logit - logit.max(0))
    w /= w.sum(0)
    mu = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])  # [3][16][n]
    ed = q["eval_dirs"].cpu().numpy().astype(np.float64)  # [4][3][n]
    for e in range(4):
        c = (mu * ed[e][:, None, :]).sum(0)
        v = kap / (2 * np.pi * -np.expm1(-2 * kap)) * np.exp(kap * np.minimum(0, c - 1))
        ref = (w * v).sum(0)
        rel = np.abs(got[e] - ref) / ref
        print("eval", e, "max rel", rel.max())
        assert np.quantile(rel, 0.999) < 1e-4 and rel.max() < 1e-3
    net.close()


def _bin_grid():
    z = -1 + (np.arange(NZ * SUB) + 0.5) * (2.0 / (NZ * SUB))
    ph = (np.arange(NP * SUB) + 0.5) * (2 * np.pi / (NP * SUB))
    Z, PH = np.meshgrid(z, ph, indexing="ij")
    r = np.sqrt(1 - Z ** 2)
    d = np.stack([r * np.cos(PH), r * np.sin(PH), Z]).reshape(3, NZ, SUB, NP, SUB).transpose(0, 1, 3, 2, 4)
    return d.reshape(3, -1).astype(np.float32)


def test_asg_samples_follow_pdf(gpu):
    import torch
    import npmg
    from scipy import stats

    net = npmg.GuidingNetwork(_cfg(), seed=2)
    grid = torch.from_numpy(_bin_grid()).cuda()
    n_grid = grid.shape[1]
    n = 1 << 20
    rng = np.random.default_rng(3)
    for i in range(3):
        raw1 = _raw(1, 10 + i, kmax=100.0, log_a=1.5)
        raw = torch.from_numpy(np.repeat(raw1, n, axis=1)).cuda()
        u = torch.from_numpy(rng.random((3, n), dtype=np.float32)).cuda()
        out = net.mixture_sample_pdf(raw, None, u, None, cosine_aware=False)
        w = out["dir"].cpu().numpy().astype(np.float64)
        z = np.clip(w[2], -1, 1)
        phi = np.mod(np.arctan2(w[1], w[0]), 2 * np.pi)
        idx = np.minimum((z + 1) / 2 * NZ, NZ - 1).astype(np.int64) * NP + np.minimum(phi / (2 * np.pi) * NP, NP - 1).astype(np.int64)
        counts = np.bincount(idx, minlength=NZ * NP)
        rawg = torch.from_numpy(np.repeat(raw1, n_grid, axis=1)).cuda()
        pdf = net.mixture_sample_pdf(rawg, None, None, grid.reshape(1, 3, n_grid).contiguous(),
                                     cosine_aware=False)["eval_pdf"][0].cpu().numpy().astype(np.float64)
        prob = pdf.reshape(NZ * NP, SUB * SUB).sum(1) * (4 * np.pi / n_grid)
        assert abs(prob.sum() - 1) < 5e-3, prob.sum()
        expected = prob / prob.sum() * n
        keep = expected >= 5
        chi2 = float((((counts - expected) ** 2 / expected)[keep]).sum()
                     + (counts[~keep].sum() - expected[~keep].sum()) ** 2 / max(expected[~keep].sum(), 1.0))
        p = float(stats.chi2.sf(chi2, int(keep.sum())))
        print(f"query {i}: chi2 {chi2:.0f} dof {int(keep.sum())} p {p:.3f}")
        assert p > 1e-4
        # the sampled pdf is the pdf at the sampled direction
        ev = net.mixture_sample_pdf(raw[:, :4096].contiguous(), None, None,
                                    out["dir"][:, :4096].reshape(1, 3, 4096).contiguous(), cosine_aware=False)
        assert torch.allclose(ev["eval_pdf"][0], out["pdf"][:4096], rtol=1e-6, atol=0)
    net.close()


def test_asg_adjoint_matches_fp64_finite_differences(gpu):
    import torch
    import npmg

    net = npmg.GuidingNetwork(_cfg(), seed=1)
    n = 256
    raw = torch.from_numpy(_raw(n, 7, kmax=1e3)).cuda()
    t = npmg.synth_train(7, n)
    dout, loss = net.mixture_adjoint(raw, t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], loss_scale=float(n))
    fd = net.loss_fd(raw, t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], h=1e-4).cpu().numpy()
    g = dout.cpu().numpy().astype(np.float64)
    scale = np.abs(fd).max(axis=0, keepdims=True) + 1e-30
    err = np.abs(g - fd) / scale
    print("max relative error vs FD:", float(err.max()), "median:", float(np.median(err)))
    assert float(err.max()) < 1e-2
    net.close()


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_asg_mfma_query_matches_consumer(gpu, prec):
    import npmg

    net = npmg.GuidingNetwork(_cfg(prec), seed=4)
    ref = npmg.GuidingNetwork(_cfg("fp32"), seed=4)
    n = 8192
    q = npmg.synth_queries(3, n)
    a = net.sample_pdf(q["inputs"], None, q["u"], q["eval_dirs"], cosine_aware=False, raw=True)
    c = net.mixture_sample_pdf(a["raw"], None, q["u"], q["eval_dirs"], cosine_aware=False)
    for k in ("dir", "pdf", "eval_pdf", "lobe"):
        assert (a[k] == c[k]).all(), k
    b = ref.sample_pdf(q["inputs"], None, q["u"], q["eval_dirs"], cosine_aware=False, raw=True)
    rel = ((a["raw"] - b["raw"]).abs().max(0).values / (b["raw"].abs().max(0).values + 1e-6)).cpu().numpy()
    print(prec, "raw output relative error: median", np.median(rel), "max", rel.max())
    assert np.median(rel) < 2e-2
    net.close()
    ref.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_asg_training_lowers_the_loss(gpu, prec):
    import torch
    import npmg

    net = npmg.GuidingNetwork(_cfg(prec), seed=5)
    t = npmg.synth_train(21, 1 << 14)
    loss = torch.empty(1, device="cuda")
    ls = []
    for _ in range(60):
        ls.append(net.train_step(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], loss).item())
    print(prec, ls[0], np.mean(ls[-5:]))
    assert np.isfinite(ls).all()
    assert np.mean(ls[-5:]) < 0.8 * ls[0]
    net.close()
