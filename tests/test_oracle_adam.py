"""Adam and EMA as tiny-cuda-nn implements them (optimizers/adam.h:47-119, optimizers/ema.h:62-136)."""
import numpy as np


def test_adam_three_steps_hand_computed(oracle):
    n, n_matrix = 6, 3
    lr, b1, b2, eps, l2, ls = 1e-2, 0.9, 0.99, 1e-15, 1e-6, 128.0
    master = np.array([0.5, -0.25, 0.125, 1e-4, -2e-4, 3e-4], np.float32)
    grads = [np.array([1.0, -2.0, 0.5, 0.0, 4.0, 0.25], np.float32) * 128,
             np.array([0.5, 0.0, -1.0, 1.0, 0.0, -0.25], np.float32) * 128,
             np.array([-1.0, 1.0, 2.0, 0.0, 3.0, 0.5], np.float32) * 128]
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    steps = np.zeros(n, np.uint32)
    w = master.astype(np.float64).copy()
    M, V, T = np.zeros(n), np.zeros(n), np.zeros(n, int)
    for g in grads:
        oracle.adam_step(n_matrix, ls, lr, b1, b2, eps, l2, master, g, m, v, steps)
        for i in range(n):
            gi = g[i] / ls
            if i >= n_matrix and gi == 0:
                continue  # zero gradient: grid parameter untouched, its step count stays
            if i < n_matrix:
                gi += l2 * w[i]  # L2 only on matrix (MLP) parameters
            M[i] = b1 * M[i] + (1 - b1) * gi
            V[i] = b2 * V[i] + (1 - b2) * gi * gi
            T[i] += 1
            lr_t = lr * np.sqrt(1 - b2 ** T[i]) / (1 - b1 ** T[i])
            w[i] -= lr_t / (np.sqrt(V[i]) + eps) * M[i]
    assert np.array_equal(steps, T.astype(np.uint32))
    assert np.allclose(master, w, rtol=2e-6, atol=1e-9)
    assert np.allclose(m, M, rtol=1e-6) and np.allclose(v, V, rtol=1e-5)


def test_ema_debias(oracle):
    d = 0.95
    w1 = np.array([1.0, -2.0, 3.0], np.float32)
    ema = np.zeros(3, np.float32)
    oracle.ema_step(d, 1, w1, ema)
    assert np.allclose(ema, w1)  # step 1: debias_old = 0, result = w
    w2 = np.array([2.0, 0.0, -1.0], np.float32)
    oracle.ema_step(d, 2, w2, ema)
    want = (w1 * d * (1 - d) + w2 * (1 - d)) / (1 - d * d)
    assert np.allclose(ema, want, rtol=1e-6)


def test_half_ema_recurrence(oracle):
    """tcnn's FP16 EMA (ema_step_half_precision, optimizers/ema.h:62-76): every
    stored average is an FP16 value; the first step returns the FP16 weight
    (debiasing); a weight within half an FP16 step of the average leaves it
    unchanged (the stall the FP32 recurrence does not have); with weights held
    constant at an FP16 value the average stays on it exactly."""
    r = np.random.default_rng(5)
    n = 50000
    w = r.normal(0, 0.3, n).astype(np.float32)
    ema = np.zeros(n, np.float32)
    oracle.ema_step(0.95, 1, w, ema, half=True)
    h = w.astype(np.float16).astype(np.float32)
    assert np.array_equal(ema, h)
    for t in range(2, 6):
        w2 = (w + r.normal(0, 1e-2, n)).astype(np.float32)
        oracle.ema_step(0.95, t, w2, ema, half=True)
        assert np.array_equal(ema, ema.astype(np.float16).astype(np.float32))
    # stall: 1 + 2^-11 is within half an FP16 step of 1 after the (1 - decay) weighting
    e = np.ones(4, np.float32)
    oracle.ema_step(0.95, 200, np.full(4, 1.0 + 2.0 ** -10, np.float32), e, half=True)
    assert np.all(e == 1.0)
    f = np.ones(4, np.float32)
    oracle.ema_step(0.95, 200, np.full(4, 1.0 + 2.0 ** -10, np.float32), f, half=False)
    assert np.all(f > 1.0)
    # constant FP16 weights are a fixed point
    c = h.copy()
    for t in range(7, 12):
        oracle.ema_step(0.95, t, h, c, half=True)
    assert np.max(np.abs(c - h) / np.maximum(np.abs(h), 1e-3)) <= 2.0 ** -10
