"""pcg32 known answers and tcnn's parameter initialisation, restated in numpy."""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MASK = (1 << 64) - 1


class NpPcg:
    """pcg32 in Python integers (O'Neill 2014), for tiny cases."""

    def __init__(self, s, q=1):
        self.state, self.inc = 0, ((q << 1) | 1) & MASK
        self.next()
        self.state = (self.state + s) & MASK
        self.next()

    def next(self):
        old = self.state
        self.state = (old * 0x5851F42D4C957F2D + self.inc) & MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def next_float(self):
        return np.float32(np.uint32((self.next() >> 9) | 0x3F800000).view(np.float32) - np.float32(1.0))


def test_pcg32_known_answers(oracle):
    g = json.load(open(os.path.join(HERE, "golden", "pcg32_known_answers.json")))
    out = np.zeros(6, np.uint32)
    oracle.lib.orc_pcg32(ctypes.c_uint64(g["initstate"]), ctypes.c_uint64(g["initseq"]), 6,
                         ctypes.c_void_p(out.ctypes.data))
    assert [hex(v) for v in out] == g["outputs"]
    r = NpPcg(g["initstate"], g["initseq"])
    assert [hex(r.next()) for _ in range(6)] == g["outputs"]


def _tiny_config():
    return {
        "encoding": {"otype": "Composite", "nested": [
            {"otype": "HashGrid", "n_dims_to_encode": 3, "n_levels": 2, "n_features_per_level": 2,
             "log2_hashmap_size": 5, "base_resolution": 2, "per_level_scale": 1.5},
            {"otype": "OneBlob", "n_dims_to_encode": 6, "n_bins": 2}]},
        "network": {"n_neurons": 64, "n_hidden_layers": 1},
        "mixture": {"n_lobes": 8},
    }


def test_init_matches_numpy_restatement(oracle):
    import random

    import pyoracle

    cfg = pyoracle.make_cfg(_tiny_config())
    n_mlp, n_grid = oracle.n_params(cfg)
    p = oracle.init_params(cfg, 1234, passes=2)
    # std::seed_seq{1234}.generate(2) -> first value (C++ standard algorithm [rand.util.seedseq])
    seeds = _seed_seq([1234], 2)
    rng = NpPcg(seeds[0])
    ew, po = oracle.widths(cfg)
    layers = [(64, ew), (po, 64)]
    for _ in range(2):
        want = []
        for out_w, in_w in layers:
            s = np.float32(np.sqrt(np.float32(6.0) / np.float32(in_w + out_w)))
            for _ in range(out_w * in_w):
                want.append(np.float32(np.float32(rng.next_float() * np.float32(2.0)) * s) - s)
        nt = (n_grid + 3) // 4
        r2 = NpPcg(0)
        r2.state, r2.inc = rng.state, rng.inc
        grid = np.zeros(n_grid, np.float32)
        for d in range(4 * nt):
            v = r2.next_float()
            idx = d // 4 + nt * (d % 4)
            if idx < n_grid:
                grid[idx] = np.float32(np.float64(v) * np.float64(np.float32(2e-4)) + np.float64(np.float32(-1e-4)))
        for _ in range(n_grid):
            rng.next()
    assert np.array_equal(p[:n_mlp], np.array(want, np.float32))
    assert np.array_equal(p[n_mlp:], grid)
    assert np.all(np.abs(p[n_mlp:]) <= 1e-4)


def _seed_seq(v, n):
    """std::seed_seq::generate ([rand.util.seedseq] of the C++ standard)."""
    s, m = len(v), n
    b = [0x8B8B8B8B] * n
    t = 11 if n >= 623 else (7 if n >= 68 else (5 if n >= 39 else (3 if n >= 7 else (n - 1) // 2)))
    p = (n - t) // 2
    q = p + t
    mm = max(s + 1, n)
    T = lambda x: x ^ (x >> 27)
    for k in range(mm):
        r1 = (1664525 * T(b[k % n] ^ b[(k + p) % n] ^ b[(k - 1) % n])) & 0xFFFFFFFF
        r2 = (r1 + (s if k == 0 else ((k % n) + v[k - 1] if k <= s else k % n))) & 0xFFFFFFFF
        b[(k + p) % n] = (b[(k + p) % n] + r1) & 0xFFFFFFFF
        b[(k + q) % n] = (b[(k + q) % n] + r2) & 0xFFFFFFFF
        b[k % n] = r2
    for k in range(mm, mm + n):
        r3 = (1566083941 * T((b[k % n] + b[(k + p) % n] + b[(k - 1) % n]) & 0xFFFFFFFF)) & 0xFFFFFFFF
        r4 = (r3 - (k % n)) & 0xFFFFFFFF
        b[(k + p) % n] ^= r3
        b[(k + q) % n] ^= r4
        b[k % n] = r4
    return b
