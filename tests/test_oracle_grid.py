"""Hash/dense grid layout, corner indices and interpolation against closed forms."""
import numpy as np
import pytest

import pyoracle


def _cfg(grid, F=2):
    g = dict(otype="HashGrid", n_dims_to_encode=3, n_features_per_level=F)
    g.update(grid)
    return {"encoding": {"otype": "Composite", "nested": [g, {"otype": "OneBlob", "n_dims_to_encode": 6, "n_bins": 4}]},
            "network": {"n_neurons": 64, "n_hidden_layers": 1}, "mixture": {"n_lobes": 8}}


GRIDS = {
    "small": dict(n_levels=8, log2_hashmap_size=15, base_resolution=16, per_level_scale=1.5),
    "full": dict(n_levels=16, log2_hashmap_size=19, base_resolution=16, per_level_scale=1.5),
    "dense_ref": dict(otype="DenseGrid", n_levels=8, base_resolution=8, per_level_scale=1.405,
                      interpolation="Smoothstep"),
}


@pytest.mark.parametrize("name", GRIDS)
def test_layout_closed_form(oracle, name):
    g = GRIDS[name]
    cfg = pyoracle.make_cfg(_cfg(g))
    offs, scales, res = oracle.grid_layout(cfg)
    L = g["n_levels"]
    T = 1 << g.get("log2_hashmap_size", 19)
    off = 0
    for l in range(L):
        s = np.float32(np.exp2(np.float32(l) * np.float32(np.log2(np.float32(g["per_level_scale"]))))) \
            * np.float32(g["base_resolution"]) - np.float32(1)
        assert abs(scales[l] - s) <= 2 * np.spacing(s)
        r = int(res[l])
        assert r == int(np.ceil(scales[l])) + 1
        size = -(-r ** 3 // 8) * 8
        if g.get("otype", "HashGrid") == "HashGrid":
            size = min(size, T)
        assert offs[l] == off
        off += size
    assert offs[L] == off


def _closed_form_index(x, scale, res, size, hashed):
    """(cell + corner) -> index, with uint32 wrap-around, for one level."""
    p = (x.astype(np.float64) * np.float64(scale) + 0.5).astype(np.float32)  # exact fma(x, scale, 0.5)
    cell = np.floor(p).astype(np.int64).astype(np.uint64)
    frac = p - np.floor(p)
    idx, w = [], []
    for k in range(8):
        c = [cell[d] + ((k >> d) & 1) for d in range(3)]
        c = [ci & 0xFFFFFFFF for ci in c]
        if hashed:
            h = (c[0] * 1) ^ ((c[1] * 2654435761) & 0xFFFFFFFF) ^ ((c[2] * 805459861) & 0xFFFFFFFF)
        else:
            h = (c[0] + c[1] * res + c[2] * res * res) & 0xFFFFFFFF
        idx.append((h & 0xFFFFFFFF) % size)
        ww = np.ones_like(frac[0])
        for d in range(3):
            ww = ww * (frac[d] if (k >> d) & 1 else np.float32(1) - frac[d])
        w.append(ww)
    return np.array(idx, np.uint64), np.array(w, np.float32)


@pytest.mark.parametrize("name", ["small", "full"])
def test_corner_indices_brute_force(oracle, name):
    cfg = pyoracle.make_cfg(_cfg(GRIDS[name]))
    rng = np.random.default_rng(0)
    n = 3000
    inp = rng.random((9, n)).astype(np.float32)
    inp[:3, :6] = np.array([[0, 1, 0.5, 1 - 2 ** -24, 0.25, 1]] * 3, np.float32)
    idx, w = oracle.grid_corners(cfg, inp)
    offs, scales, res = oracle.grid_layout(cfg)
    for l in range(cfg.n_levels):
        size = int(offs[l + 1] - offs[l])
        hashed = int(res[l]) ** 3 > size
        ci, cw = _closed_form_index(inp[:3], scales[l], int(res[l]), size, hashed)
        assert np.array_equal(idx[l].astype(np.uint64), ci), f"level {l}"
        assert np.array_equal(w[l], cw)
        assert np.allclose(w[l].sum(0), 1.0, atol=1e-6)
    # x = 0 puts p = 0.5 on every level: cell 0, weights 1/8
    assert np.all(idx[:, 0, 0] == 0) and np.all(w[:, :, 0] == 0.125)


def test_grid_forward_is_trilinear_gather(oracle):
    cfg = pyoracle.make_cfg(_cfg(GRIDS["small"]))
    _, n_grid = oracle.n_params(cfg)
    rng = np.random.default_rng(1)
    table = rng.normal(size=n_grid).astype(np.float32)
    inp = rng.random((9, 500)).astype(np.float32)
    enc = oracle.encode(cfg, table, inp, dtype=np.float64)
    idx, w = oracle.grid_corners(cfg, inp)
    offs, _, _ = oracle.grid_layout(cfg)
    for l in range(cfg.n_levels):
        for f in range(2):
            vals = table[(offs[l] + idx[l].astype(np.int64)) * 2 + f].astype(np.float64)
            want = (w[l].astype(np.float64) * vals).sum(0)
            # the FP64 oracle forms the weights in FP64 from the FP32 fractions
            assert np.allclose(enc[l * 2 + f], want, rtol=0, atol=1e-6 * np.abs(table).max())


def test_smoothstep_dense_reference_grid(oracle):
    cfg = pyoracle.make_cfg(_cfg(GRIDS["dense_ref"], F=4))
    inp = np.random.default_rng(2).random((9, 400)).astype(np.float32)
    idx, w = oracle.grid_corners(cfg, inp)
    offs, scales, res = oracle.grid_layout(cfg)
    for l in range(cfg.n_levels):
        p = (inp[:3].astype(np.float64) * np.float64(scales[l]) + 0.5).astype(np.float32)
        f = p - np.floor(p)
        s = f * f * (np.float32(3) - np.float32(2) * f)
        assert np.allclose(w[l][0], (1 - s[0]) * (1 - s[1]) * (1 - s[2]), rtol=1e-6)
        cell = np.floor(p).astype(np.int64)
        r = int(res[l])
        assert np.array_equal(idx[l][0], ((cell[0] + cell[1] * r + cell[2] * r * r) % (offs[l + 1] - offs[l])))
