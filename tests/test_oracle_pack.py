"""The oracle's wavefront packers against closed forms and constructed cases.

generate_inference_data (train.cu:174-191) and generate_training_data
(train.cu:14-67): position normalised to the inflated scene box
(train.h:173-177), direction warp (w + 1) / 2, normal as (theta/pi, phi/2pi),
roughness 1 - e^-r; training records filtered and weighted as the reference
does. Values are chosen so that the expected results are exact or follow
from textbook functions in FP64.
"""
import numpy as np
import pytest

BOX = np.array([-2.0, 0.0, 1.0, 6.0, 4.0, 3.0], np.float32)


def _box64(x):
    lo, hi = BOX[:3].astype(np.float64), BOX[3:].astype(np.float64)
    inf = np.linalg.norm(hi - lo) * 0.005
    lo, hi = lo - inf, hi + inf
    return 0.5 + (x - ((lo + hi) / 2)[:, None]) / (hi - lo)[:, None]


def test_pack_queries_closed_forms(oracle):
    rng = np.random.default_rng(0)
    n = 4096
    pos = rng.uniform(-3, 7, (3, n)).astype(np.float32)
    wo = rng.normal(size=(3, n)).astype(np.float32) * rng.uniform(0.1, 10, n).astype(np.float32)
    nrm = rng.normal(size=(3, n)).astype(np.float32)
    rough = rng.uniform(0, 3, n).astype(np.float32)
    # special cases: centre, +z, unnormalised axes
    pos[:, 0] = [2, 2, 2]
    wo[:, 0], nrm[:, 0] = [0, 0, 1], [0, 0, 1]
    wo[:, 1], nrm[:, 1] = [0, 0, 5], [1, 0, 0]
    wo[:, 2], nrm[:, 2] = [-3, 0, 0], [0, -2, 0]
    rough[0] = 0
    got = oracle.pack_queries(BOX, pos, wo, nrm, rough).astype(np.float64)
    # a few FP32 roundings of values up to about 2.5 (points outside the box)
    assert np.allclose(got[:3], _box64(pos.astype(np.float64)), rtol=0, atol=6e-7)
    assert np.allclose(got[:3, 0], 0.5, atol=1e-7)
    w = wo.astype(np.float64) / np.linalg.norm(wo.astype(np.float64), axis=0)
    assert np.allclose(got[3:6], (w + 1) / 2, atol=2e-7)
    assert np.array_equal(got[3:6, 0], [0.5, 0.5, 1.0]) and np.array_equal(got[3:6, 1], [0.5, 0.5, 1.0])
    assert np.array_equal(got[3:6, 2], [0.0, 0.5, 0.5])
    v = nrm.astype(np.float64) / np.linalg.norm(nrm.astype(np.float64), axis=0)
    th = np.arccos(np.clip(v[2], -1, 1)) / np.pi
    ph = np.mod(np.arctan2(v[1], v[0]), 2 * np.pi) / (2 * np.pi)
    assert np.allclose(got[6], th, atol=1e-6)
    dph = np.abs(got[7] - ph)
    assert np.all(np.minimum(dph, 1 - dph) < 1e-6)  # phi wraps at 2 pi
    assert np.array_equal(got[6:8, 0], [0.0, 0.0])
    assert np.allclose(got[6:8, 1], [0.5, 0.0], atol=1e-7)
    assert np.allclose(got[6:8, 2], [0.5, 0.75], atol=1e-7)
    assert np.allclose(got[8], 1 - np.exp(-rough.astype(np.float64)), atol=2e-7)
    assert got[8, 0] == 0


def _frame(D=3, W=4, H=3, seed=1):
    rng = np.random.default_rng(seed)
    P = W * H
    f = lambda c: rng.uniform(0.5, 2.0, (D, c, P)).astype(np.float32)
    rec = {
        "L": f(3), "thp": f(3), "pos": rng.uniform(-1, 5, (D, 3, P)).astype(np.float32),
        "dir": np.stack([rng.uniform(0, np.pi, (D, P)), rng.uniform(0, 2 * np.pi, (D, P))], 1).astype(np.float32),
        "wi_pdf": rng.uniform(0.1, 2, (D, P)).astype(np.float32),
        "mis_weight": rng.uniform(0, 1, (D, P)).astype(np.float32),
        "delta": np.zeros((D, P), np.uint8), "bsdf_val": f(3),
        "wo": np.stack([rng.uniform(0, np.pi, (D, P)), rng.uniform(0, 2 * np.pi, (D, P))], 1).astype(np.float32),
        "normal": np.stack([rng.uniform(0, np.pi, (D, P)), rng.uniform(0, 2 * np.pi, (D, P))], 1).astype(np.float32),
        "roughness": rng.uniform(0, 1, (D, P)).astype(np.float32),
        "depth": np.full(P, D, np.uint32),
    }
    return rec


def test_pack_training_rules(oracle):
    D, W, H = 3, 4, 3
    rec = _frame(D, W, H)
    # pixel 1: exact target arithmetic on vertex 0, delta vertex 1, thp below 1e-5 in channel 2 of vertex 2
    rec["L"][0, :, 1], rec["thp"][0, :, 1] = [2.0, 3.0, 1.0], [0.5, 0.25, 2.0]
    rec["mis_weight"][0, 1], rec["bsdf_val"][0, :, 1] = 0.25, [3.0, 1.0, 0.5]
    rec["delta"][1, 1] = 1
    rec["thp"][2, 2, 1] = 1e-6
    # pixel 3: NaN position (dropped), zero pdf (dropped), NaN bsdf (kept: the reference tests L before it)
    rec["pos"][0, 1, 3] = np.nan
    rec["wi_pdf"][1, 3] = 0
    rec["bsdf_val"][2, 0, 3] = np.nan
    # pixel 9: one vertex only; pixel 11: recorded depth above the array's (clamped), NaN radiance on vertex 1
    rec["depth"][9] = 1
    rec["depth"][11] = 7
    rec["L"][1, 0, 11] = np.nan
    # pixel 0 (not a training pixel) holds NaNs everywhere: never read
    for k in ("L", "pos"):
        rec[k][:, :, 0] = np.nan
    tile = dict(offset=(1, 0), stride=(2, 2), res=(2, 2), resolution=(W, H))  # pixels 1, 3, 9, 11
    inp, tdir, rgb, pdf, cnt = oracle.pack_training(rec, tile, BOX, 64)
    kept = [(1, 0), (1, 2), (3, 2), (9, 0), (11, 0), (11, 2)]
    assert cnt == len(kept)
    for j, (p, d) in enumerate(kept):
        assert np.array_equal(tdir[:, j], rec["dir"][d, :, p])
        assert pdf[j] == rec["wi_pdf"][d, p]
        x = _box64(rec["pos"][d, :, p].astype(np.float64)[:, None])[:, 0]
        assert np.allclose(inp[:3, j], x, atol=3e-7)
        th, ph = rec["wo"][d, :, p].astype(np.float64)
        w = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        assert np.allclose(inp[3:6, j], (w + 1) / 2, atol=3e-7)
        assert np.allclose(inp[6:8, j], rec["normal"][d, :, p] / np.array([np.pi, 2 * np.pi]), rtol=1e-6)
        assert np.allclose(inp[8, j], 1 - np.exp(-rec["roughness"][d, p].astype(np.float64)), atol=2e-7)
    # exact target: (L / thp) * mis * bsdf = (4, 12, 0.5) * 0.25 * (3, 1, 0.5)
    assert np.array_equal(rgb[:, 0], np.float32([3.0, 3.0, 0.0625]))
    assert rgb[2, 1] == 0  # thp <= 1e-5
    assert np.isnan(rgb[0, 2]) and np.isfinite(rgb[1:, 2]).all()
    assert not np.any(inp[:, cnt:]) and not np.any(pdf[cnt:])
    # capacity below the count: the first records, the full count
    inp2, _, rgb2, _, cnt2 = oracle.pack_training(rec, tile, BOX, 4)
    assert cnt2 == cnt
    assert np.array_equal(inp2, inp[:, :4]) and np.array_equal(rgb2[:, :4][:, :2], rgb[:, :2])


def test_pack_training_tile_outside_image(oracle):
    rec = _frame(2, 4, 3)
    tile = dict(offset=(3, 2), stride=(2, 2), res=(2, 2), resolution=(4, 3))  # only pixel (3, 2) is inside
    _, _, _, _, cnt = oracle.pack_training(rec, tile, BOX, 16)
    assert cnt == 2
