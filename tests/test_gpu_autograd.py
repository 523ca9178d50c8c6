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


def _cfg(prec, size="small"):
    import npmg

    c = npmg.benchmark_config(size, precision=prec)
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


@pytest.mark.parametrize("size", ["small", "full"])
def test_autograd_gradient_matches_fp64_differences(gpu, oracle, size):
    """size "full": the network of the full-frame configuration (16 levels of
    2^19 entries, MLP 3 x 64), the differentiation check's own network."""
    import torch
    import npmg
    import pyoracle
    from npmg.autograd import mixture_loss, network_apply

    cfg = _cfg("fp32", size)
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
    print(f"{size}: loss {loss.item():.6e} (FP64 {l64:.6e}); coordinate worst {worst:.2e}, "
          f"direction worst {dworst:.2e}")
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


def test_autograd_backward_uses_forward_time_params(gpu):
    """The backward differentiates at the parameters its forward saw, even when
    another forward with other parameters (or an optimiser step) ran in
    between, and network_apply leaves the inference (EMA) copy alone."""
    import torch
    import npmg
    from npmg.autograd import mixture_loss, network_apply

    cfg = _cfg("fp32")
    cfg["optimizer"] = {"otype": "Ema", "decay": 0.99, "nested": {"otype": "Adam", "learning_rate": 1e-2,
                                                                  "beta1": 0.9, "beta2": 0.99, "epsilon": 1e-15,
                                                                  "l2_reg": 1e-6}}
    net = npmg.GuidingNetwork(cfg, seed=21)
    t = npmg.synth_train(6, 2048)
    p_a = net.get_params(0).clone()
    ema0 = net.get_params(1).clone()
    gen = torch.Generator(device="cuda").manual_seed(3)
    p_b = p_a + 1e-2 * torch.randn(p_a.shape, device="cuda", generator=gen)
    pa = p_a.clone().requires_grad_(True)
    pb = p_b.clone().requires_grad_(True)
    loss_a = mixture_loss(net, network_apply(net, t["inputs"], pa), t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"])
    loss_b = mixture_loss(net, network_apply(net, t["inputs"], pb), t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"])
    assert torch.equal(net.get_params(1), ema0), "network_apply changed the inference copy"
    net.train_step(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"])  # moves the master again
    loss_a.backward()
    loss_b.backward()
    # split passes at each parameter vector
    for p, got, l in ((p_a, pa.grad, loss_a), (p_b, pb.grad, loss_b)):
        net.set_params(p, which=0)
        raw = net.forward(t["inputs"])
        dout, per = net.mixture_adjoint(raw, t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], 1.0)
        want, _ = net.backward(t["inputs"], dout)
        assert l.item() == per.mean().item()
        assert np.array_equal(_np(got).view(np.uint32), _np(want).view(np.uint32))
    assert not torch.equal(pa.grad, pb.grad)
    net.close()
