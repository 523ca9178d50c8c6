"""Sampling matches the pdf (distribution.h:132-140, vmf.h:48-71): a chi-square
test of the directions the library draws against the pdf it evaluates.

For a few queries, 2^20 directions are drawn with independent uniforms and
binned on the sphere (32 equal-width bins in z = cos(theta) times 64 in phi,
equal areas). The expected count of each bin is the library's own pdf
integrated over the bin (8 x 8 midpoint rule), so the test checks that the
sampler and the pdf describe the same distribution, through the fused path and
the consumer kernel, with and without the cosine product.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NZ, NP, SUB = 32, 64, 8


def _bin_grid():
    """Midpoints of SUB x SUB cells inside every bin: directions [3][NZ NP SUB^2]."""
    z = -1 + (np.arange(NZ * SUB) + 0.5) * (2.0 / (NZ * SUB))
    ph = (np.arange(NP * SUB) + 0.5) * (2 * np.pi / (NP * SUB))
    Z, PH = np.meshgrid(z, ph, indexing="ij")  # [NZ SUB][NP SUB]
    r = np.sqrt(1 - Z ** 2)
    d = np.stack([r * np.cos(PH), r * np.sin(PH), Z])  # [3][NZ SUB][NP SUB]
    # group as bins: [3][NZ][NP][SUB][SUB]
    d = d.reshape(3, NZ, SUB, NP, SUB).transpose(0, 1, 3, 2, 4)
    return d.reshape(3, -1).astype(np.float32)


def _bin_index(w):
    z = np.clip(w[2], -1, 1)
    ph = np.mod(np.arctan2(w[1], w[0]), 2 * np.pi)
    iz = np.minimum((z + 1) / 2 * NZ, NZ - 1).astype(np.int64)
    ip = np.minimum(ph / (2 * np.pi) * NP, NP - 1).astype(np.int64)
    return iz * NP + ip


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("cosine", [False, True])
def test_samples_follow_pdf(gpu, prec, cosine):
    import torch
    import npmg
    from scipy import stats

    net = npmg.GuidingNetwork(npmg.benchmark_config("small", precision=prec), seed=23)
    t = npmg.synth_train(5, 1 << 14)
    loss = torch.empty(1, device="cuda")
    for _ in range(30):  # a few steps so the lobes are not the initial near-uniform ones
        net.train_step(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], loss)
    q = npmg.synth_queries(77, 3)
    grid = torch.from_numpy(_bin_grid()).cuda()
    n_grid = grid.shape[1]
    area = 4 * np.pi / n_grid
    n = 1 << 20
    rng = np.random.default_rng(1)
    for i in range(3):
        inp = q["inputs"][:, i:i + 1].repeat(1, n).contiguous()
        nrm = q["normal"][:, i:i + 1].repeat(1, n).contiguous() if cosine else None
        u = torch.from_numpy(rng.random((3, n), dtype=np.float32)).cuda()
        out = net.sample_pdf(inp, nrm, u, raw=True, cosine_aware=cosine)
        counts = np.bincount(_bin_index(out["dir"].cpu().numpy().astype(np.float64)), minlength=NZ * NP)
        # the pdf over the bin grid, from the same raw outputs through the consumer kernel
        raw = out["raw"][:, :1].repeat(1, n_grid).contiguous()
        nrm_g = q["normal"][:, i:i + 1].repeat(1, n_grid).contiguous() if cosine else None
        pdf = net.mixture_sample_pdf(raw, nrm_g, None, grid.reshape(1, 3, n_grid).contiguous(),
                                     cosine_aware=cosine)["eval_pdf"][0].cpu().numpy().astype(np.float64)
        prob = pdf.reshape(NZ * NP, SUB * SUB).sum(1) * area
        assert abs(prob.sum() - 1) < 2e-3, prob.sum()
        expected = prob / prob.sum() * n
        keep = expected >= 5
        chi2 = float((((counts - expected) ** 2 / np.maximum(expected, 1e-30))[keep]).sum()
                     + (counts[~keep].sum() - expected[~keep].sum()) ** 2 / max(expected[~keep].sum(), 1.0))
        dof = int(keep.sum())
        p = float(stats.chi2.sf(chi2, dof))
        print(f"{prec} cosine={cosine} query {i}: chi2 {chi2:.0f} dof {dof} p {p:.3f}")
        assert p > 1e-4
    net.close()
