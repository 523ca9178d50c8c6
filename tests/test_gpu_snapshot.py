"""Snapshots in tcnn's format (Trainer::serialize, trainer.h:267-307; Adam
adam.h:282-298; Ema ema.h:194-206), read back with an independent
MessagePack implementation (the msgpack package), resumed exactly, and
loaded from a document shaped like the reference's own (FP16 parameters).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cfg(prec, ema=True):
    import npmg

    c = npmg.benchmark_config("small", precision=prec)
    c["grid_gradient"] = "deterministic"
    if ema:  # as the reference's base.json: Ema wrapping Adam
        c["optimizer"] = {"otype": "Ema", "decay": 0.95, "nested": c["optimizer"]}
    return c


def _steps(net, k, seed=4):
    import torch
    import npmg

    t = npmg.synth_train(seed, 1 << 13)
    loss = torch.empty(1, device="cuda")
    for _ in range(k):
        net.train_step(t["inputs"], t["tgt_dir"], t["tgt_rgb"], t["tgt_pdf"], loss)
    return loss.item()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_snapshot_format(gpu, prec):
    import msgpack
    import npmg

    net = npmg.GuidingNetwork(_cfg(prec), seed=3)
    _steps(net, 7)
    doc = msgpack.unpackb(net.snapshot(True))
    assert list(doc) == sorted(doc)  # nlohmann's key order
    assert set(doc) == {"n_params", "npmg", "optimizer", "params_binary", "params_type"}
    assert doc["n_params"] == net.n_params and doc["params_type"] == "float"
    assert doc["params_binary"] == net.get_params(1).cpu().numpy().tobytes()
    assert doc["npmg"]["master_binary"] == net.get_params(0).cpu().numpy().tobytes()
    opt = doc["optimizer"]
    assert set(opt) == {"nested", "weights_ema_binary"}
    adam = opt["nested"]
    assert set(adam) == {"base_learning_rate", "current_step", "first_moments_binary", "param_steps_binary",
                         "second_moments_binary"}
    scale, step = net.train_state()
    assert adam["current_step"] == step == 7
    assert abs(adam["base_learning_rate"] - 1e-2) < 1e-9
    for k in ("first_moments_binary", "second_moments_binary", "param_steps_binary"):
        assert len(adam[k]) == 4 * net.n_params
    assert np.frombuffer(adam["param_steps_binary"], np.uint32).max() == 7
    assert doc["npmg"]["loss_scale"] == scale
    assert "optimizer" not in msgpack.unpackb(net.snapshot(False))
    net.close()
    # plain Adam: the optimizer map is Adam's own
    net = npmg.GuidingNetwork(_cfg(prec, ema=False), seed=3)
    _steps(net, 2)
    doc = msgpack.unpackb(net.snapshot(True))
    assert doc["optimizer"]["current_step"] == 2 and "nested" not in doc["optimizer"]
    assert doc["params_binary"] == doc["npmg"]["master_binary"]
    net.close()


@pytest.mark.parametrize("prec,ema", [("fp32", True), ("bf16", True), ("bf16", False)])
def test_snapshot_resume_is_exact(gpu, prec, ema):
    import npmg

    a = npmg.GuidingNetwork(_cfg(prec, ema), seed=3)
    _steps(a, 10)
    snap = a.snapshot(True)
    b = npmg.GuidingNetwork(_cfg(prec, ema), seed=99)
    b.load_snapshot(snap)
    for which in (0, 1):
        assert np.array_equal(a.get_params(which).cpu().numpy(), b.get_params(which).cpu().numpy())
    assert a.train_state() == b.train_state()
    la, lb = _steps(a, 5, seed=8), _steps(b, 5, seed=8)
    assert la == lb
    for which in (0, 1):
        assert np.array_equal(a.get_params(which).cpu().numpy(), b.get_params(which).cpu().numpy())
    a.close()
    b.close()


def test_snapshot_from_reference_shaped_document(gpu):
    """A document as the reference's Trainer<float, __half> writes it: FP16
    inference parameters, FP16 EMA weights, FP32 moments, no npmg map."""
    import msgpack
    import npmg

    net = npmg.GuidingNetwork(_cfg("bf16"), seed=3)
    n = net.n_params
    rng = np.random.default_rng(0)
    p16 = (rng.normal(0, 0.05, n)).astype(np.float16)
    e16 = (rng.normal(0, 0.05, n)).astype(np.float16)
    m = rng.normal(0, 1e-3, n).astype(np.float32)
    v = rng.uniform(0, 1e-4, n).astype(np.float32)
    doc = {"n_params": n, "params_type": "__half", "params_binary": p16.tobytes(),
           "optimizer": {"nested": {"current_step": 42, "base_learning_rate": 0.01,
                                    "first_moments_binary": m.tobytes(), "second_moments_binary": v.tobytes()},
                         "weights_ema_binary": e16.tobytes()}}
    net.load_snapshot(msgpack.packb(doc))
    assert np.array_equal(net.get_params(0).cpu().numpy(), p16.astype(np.float32))  # master from params_binary
    assert np.array_equal(net.get_params(1).cpu().numpy(), e16.astype(np.float32))  # EMA
    assert net.train_state()[1] == 42
    back = msgpack.unpackb(net.snapshot(True))
    assert back["optimizer"]["nested"]["first_moments_binary"] == m.tobytes()
    assert np.frombuffer(back["optimizer"]["nested"]["param_steps_binary"], np.uint32).max() == 0
    # only parameters: the optimiser state stays
    net.load_snapshot(msgpack.packb({"n_params": n, "params_binary": np.zeros(n, np.float32).tobytes()}))
    assert not net.get_params(0).cpu().numpy().any()
    assert net.train_state()[1] == 42
    net.close()


def test_snapshot_rejects_bad_documents(gpu):
    import msgpack
    import npmg

    net = npmg.GuidingNetwork(_cfg("fp32"), seed=3)
    good = net.snapshot(True)
    n = net.n_params
    for bad in (good[:len(good) // 2], good + b"\x00",
                msgpack.packb({"n_params": n + 1, "params_binary": b"\x00" * 4 * (n + 1)}),
                msgpack.packb({"n_params": n, "params_binary": b"\x00" * 3}),
                msgpack.packb({"n_params": n, "params_type": "double", "params_binary": b"\x00" * 8 * n}),
                msgpack.packb([1, 2, 3])):
        with pytest.raises(npmg.NpmgError):
            net.load_snapshot(bad)
    net.load_snapshot(good)  # still usable
    net.close()
