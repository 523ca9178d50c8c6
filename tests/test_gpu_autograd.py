"""torch.autograd wrapping of the network and the mixture loss (npmg.autograd;
tiny-cuda-nn's torch binding, src/ext/tcnn/bindings/torch/tinycudann/
modules.py:82-151, and the loss of compute_dL_doutput_divergence,
train.cu:69-124).

* loss.backward() through network_apply + mixture_loss gives dL/dparams
  within 1e-3 of FP64 central differences (step 1e-4, 256 queries) of the
  oracle's FP64 forward and loss: per coordinate on the largest gradients of
  every layer and of the grid, and along random directions over all
  parameters;
* the autograd gradient is bit for bit the split passes' (adjoint, then
  backward), in FP32 and BF16;
* the loss value matches the FP64 loss to FP32 accuracy.
The upper gradient clamp is switched off ("grad_clamp": 0) so that the loss is
smooth where the differences are taken.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 256
H = 1e-4


def _np(t):
    return t.detach().cpu().numpy()


def _cfg(prec):
    import npmg

    c = npmg.benchmark_config("small", precision=prec)
    c["loss"] = {"scale": 128.0, "grad_clamp": 0.0}
    c["grid_gradient"] = "deterministic"  # fixed summation order: repeatable bits
    return c


def _loss64(oracle, oc, n_mlp, p32, h, terms=False):
    """Mean loss in FP64 of FP32 parameters p32 (the oracle's FP64 encoding,
    MLP and loss); terms: also the per-query losses."""
    enc = oracle.encode(oc, p32[n_mlp:], h["inputs"], dtype=np.float64)
    raw = oracle.mlp_forward(oc, p32[:n_mlp], enc)
    _, loss = oracle.adjoint(oc, raw, h["tgt_dir"], h["tgt_rgb"], h["tgt_pdf"], 1.0, dtype=np.float64)
    return (float(loss.mean()), loss) if terms else float(loss.mean())


def _central(oracle, oc, n_mlp, p, h, g, direction):
    """Relative error of g against the central difference along direction.
    Both ends are FP32-representable points (the oracle reads FP32
    parameters), so the difference is compared with g times the displacement
    actually taken rather than 2 H direction."""
    hi = (p.astype(np.float64) + H * direction).astype(np.float32)
    lo = (p.astype(np.float64) - H * direction).astype(np.float32)
    disp = hi.astype(np.float64) - lo.astype(np.float64)
    dl = _loss64(oracle, oc, n_mlp, hi, h) - _loss64(oracle, oc, n_mlp, lo, h)
    return abs(g @ disp - dl) / abs(dl)


def test_autograd_gradient_matches_fp64_differences(gpu, oracle):
    import torch
    import npmg
    import pyoracle
    from npmg.autograd import mixture_loss, network_apply

    cfg = _cfg("fp32")
    net = npmg.GuidingNetwork(cfg, seed=31)
    oc = pyoracle.make_cfg(cfg)
    t = npmg.synth_train(13, N)
    params = net.get_params(0).clone().requires_grad_(True)
    raw = network_apply(net, t["inputs"], params)
    loss = mixture_loss(net, raw, t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"])
    loss.backward()
    g = _np(params.grad).astype(np.float64)
    p = _np(params)
    h = {k: _np(v) for k, v in t.items()}
    l64, terms = _loss64(oracle, oc, net.n_mlp, p, h, terms=True)
    assert abs(loss.item() - l64) <= 1e-5 * np.abs(terms).mean()  # terms of both signs: scale by their size

    # per coordinate: the 4 largest gradients of every weight matrix and of the grid
    bounds, off, n_in = [], 0, net.encoded_width
    for l in range(oc.n_hidden_layers + 1):
        n_out = oc.width if l < oc.n_hidden_layers else net.padded_output
        bounds.append((off, off + n_out * n_in))
        off += n_out * n_in
        n_in = oc.width
    bounds.append((net.n_mlp, net.n_params))
    worst = 0.0
    for a, b in bounds:
        for i in a + np.argsort(-np.abs(g[a:b]))[:4]:
            e = np.zeros(net.n_params)
            e[i] = 1.0
            worst = max(worst, _central(oracle, oc, net.n_mlp, p, h, g, e))
    # along random directions over all parameters (MLP and grid scaled alike)
    rng = np.random.default_rng(5)
    dworst = 0.0
    for _ in range(3):
        v = rng.normal(size=net.n_params)
        v /= np.linalg.norm(v)
        dworst = max(dworst, _central(oracle, oc, net.n_mlp, p, h, g, v))
    print(f"loss {loss.item():.6e} (FP64 {l64:.6e}); coordinate worst {worst:.2e}, direction worst {dworst:.2e}")
    assert worst < 1e-3 and dworst < 1e-3
    net.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_autograd_equals_split_passes(gpu, prec):
    import torch
    import npmg
    from npmg.autograd import mixture_loss, network_apply

    net = npmg.GuidingNetwork(_cfg(prec), seed=8)
    t = npmg.synth_train(4, 3000)
    params = net.get_params(0).clone().requires_grad_(True)
    loss = mixture_loss(net, network_apply(net, t["inputs"], params), t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"])
    loss.backward()
    raw = net.forward(t["inputs"])
    dout, per = net.mixture_adjoint(raw, t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], 1.0)
    dparams, _ = net.backward(t["inputs"], dout)
    assert np.array_equal(_np(params.grad).view(np.uint32), _np(dparams).view(np.uint32))
    assert loss.item() == per.mean().item()
    # a scaled upstream gradient scales dL/draw: 2 x loss gives 2 x the gradient
    params2 = net.get_params(0).clone().requires_grad_(True)
    (2.0 * mixture_loss(net, network_apply(net, t["inputs"], params2), t["tgt_dir"], t["tgt_rgb"],
                        t["tgt_pdf"])).backward()
    assert np.array_equal(_np(params2.grad), 2.0 * _np(params.grad))
    net.close()
