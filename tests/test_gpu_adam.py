"""The GPU optimiser (k_adam: Adam + EMA in one pass) against the oracle's
restatement of tiny-cuda-nn's AdamOptimizer and EmaOptimizer
(optimizers/adam.h:48-119, optimizers/ema.h:45-76, 102-136), step by step.

npmg_optimizer_step is driven for three steps with crafted gradient vectors:
exact zeros in grid entries (skipped: no moment update, no step count) and in
matrix entries (not skipped: the L2 term still moves them), values over
eleven decades including ones whose square is subnormal, a loss scale of 128.
After every step the master weights, both moments, the per-parameter step
counters and the EMA (read back through the snapshot document) equal the
oracle's bit for bit: both run the same IEEE operation sequence. The oracle's
recurrences are pinned by hand-computed steps on the CPU (test_oracle_adam.py).

A dynamic-loss-scale step that overflows leaves every part of the optimiser
state untouched and halves the scale.
"""
import msgpack
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EMA_DECAY = 0.95


def _cfg(prec, full=None):
    import npmg

    c = npmg.benchmark_config("small", precision=prec)
    c["optimizer"] = {"otype": "Ema", "decay": EMA_DECAY,
                      "nested": {"otype": "Adam", "learning_rate": 1e-2, "beta1": 0.9, "beta2": 0.99,
                                 "epsilon": 1e-15, "l2_reg": 1e-6}}
    if full is not None:
        c["optimizer"]["full_precision"] = full
    return c


def _state(net):
    """(master, m, v, steps, ema, current_step) from the snapshot document."""
    doc = msgpack.unpackb(net.snapshot(True), raw=False)
    opt = doc["optimizer"]
    adam = opt["nested"]
    f = lambda b: np.frombuffer(b, np.float32).copy()  # noqa: E731
    return (f(doc["npmg"]["master_binary"]), f(adam["first_moments_binary"]), f(adam["second_moments_binary"]),
            np.frombuffer(adam["param_steps_binary"], np.uint32).copy(), f(opt["weights_ema_binary"]),
            adam["current_step"])


def _gradient(rng, n_mlp, n, step):
    """Crafted gradients (already multiplied by the loss scale 128)."""
    g = np.zeros(n, np.float32)
    mag = 10.0 ** rng.uniform(-9.0, 2.0, size=n)
    g[:] = (rng.normal(size=n) * mag * 128.0).astype(np.float32)
    # exact zeros: 5 % of the matrix entries, a different half of the grid each step
    g[:n_mlp][rng.random(n_mlp) < 0.05] = 0.0
    grid = g[n_mlp:]
    grid[rng.random(grid.size) < 0.5] = 0.0
    # entries whose unscaled square is subnormal, and the largest magnitudes
    idx = rng.choice(n, size=64, replace=False)
    g[idx[:32]] = np.float32(3e-21 * 128.0) * np.sign(rng.normal(size=32)).astype(np.float32)
    g[idx[32:]] = np.float32(5e3 * 128.0) * np.sign(rng.normal(size=32)).astype(np.float32)
    if step == 2:  # grid entries skipped at steps 1 and 3 but updated at 2 and the reverse
        grid[: grid.size // 8] = np.where(grid[: grid.size // 8] == 0, np.float32(1e-3 * 128.0), 0.0)
    return g


@pytest.mark.parametrize("prec,full", [("fp32", None), ("bf16", None), ("fp16", None), ("fp32", False),
                                       ("fp16", True)])
def test_adam_ema_three_steps_match_oracle(gpu, oracle, prec, full):
    """EMA in FP32, or in FP16 like tcnn's ema_step_half_precision: explicitly
    ("full_precision": false) or by default where the parameters are FP16."""
    import torch
    import npmg

    half = full is False or (full is None and prec == "fp16")
    net = npmg.GuidingNetwork(_cfg(prec, full), seed=17)
    n, n_mlp = net.n_params, net.n_mlp
    master = net.get_params(0).cpu().numpy().copy()
    ema = master.copy()
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    steps = np.zeros(n, np.uint32)
    rng = np.random.default_rng(2024)
    for t in (1, 2, 3):
        g = _gradient(rng, n_mlp, n, t)
        net.optimizer_step(torch.from_numpy(g).cuda(), 128.0)
        oracle.adam_step(n_mlp, 128.0, 1e-2, 0.9, 0.99, 1e-15, 1e-6, master, g, m, v, steps)
        oracle.ema_step(EMA_DECAY, t, master, ema, half=half)
        gm, gmm, gv, gs, ge, cur = _state(net)
        assert cur == t
        zero_grid = np.flatnonzero(g[n_mlp:] == 0) + n_mlp
        assert zero_grid.size > 0 and np.all(gs[zero_grid] == steps[zero_grid])
        for name, got, want in (("master", gm, master), ("first moment", gmm, m), ("second moment", gv, v),
                                ("EMA", ge, ema)):
            bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
            assert bad.size == 0, (f"step {t}: {name} differs at {bad.size} entries, e.g. {bad[:4]}: "
                                   f"{got[bad[:4]]} vs {want[bad[:4]]}")
        assert np.array_equal(gs, steps), f"step {t}: per-parameter step counts differ"
    # the zero-gradient rule was exercised both ways and the L2 term moved zero-gradient matrix entries
    assert steps[n_mlp:].min() < 3 and steps[n_mlp:].max() == 3 and steps[:n_mlp].min() == 3
    assert np.all(net.get_params(1).cpu().numpy().view(np.uint32) == ema.view(np.uint32))
    net.close()


def test_overflowing_step_leaves_optimizer_state(gpu):
    """A step whose gradients overflow (dynamic scaling) is skipped: master,
    moments, step counters and EMA keep their bits; the scale halves."""
    import npmg

    cfg = _cfg("fp16")
    cfg["loss"] = {"scale": 4096.0, "dynamic": True, "scale_window": 1000, "grad_clamp": 0.0}
    net = npmg.GuidingNetwork(cfg, seed=5)
    t = npmg.synth_train(5, 4096)
    net.train_step(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"])  # one finite step: non-zero state
    before = _state(net)
    scale0, step0 = net.train_state()
    big = t["tgt_rgb"] * 1e6
    net.train_step(t["inputs"], t["tgt_dir"], big, t["tgt_pdf"])
    after = _state(net)
    scale1, step1 = net.train_state()
    assert step1 == step0 == 1 and scale1 == scale0 / 2, (scale0, step0, scale1, step1)
    for name, a, b in zip(("master", "first moment", "second moment", "steps", "EMA"), before[:5], after[:5]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{name} changed by a skipped step"
    net.close()
