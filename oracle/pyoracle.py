"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg. The product never uses it.
Arrays are numpy, structure-of-arrays [d][n] like the library's device
buffers.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MAX_ENC = 4
ENC = {"grid": 0, "oneblob": 1, "sh": 2, "identity": 3}


class OrcCfg(ctypes.Structure):
    _fields_ = [
        ("n_input_dims", ctypes.c_int), ("n_enc", ctypes.c_int),
        ("enc_type", ctypes.c_int * MAX_ENC), ("enc_begin", ctypes.c_int * MAX_ENC),
        ("enc_ndims", ctypes.c_int * MAX_ENC), ("enc_param", ctypes.c_int * MAX_ENC),
        ("grid_type", ctypes.c_int), ("hash_type", ctypes.c_int),
        ("n_levels", ctypes.c_int), ("n_features", ctypes.c_int), ("log2_hashmap_size", ctypes.c_int),
        ("base_resolution", ctypes.c_int), ("per_level_scale", ctypes.c_float), ("interpolation", ctypes.c_int),
        ("width", ctypes.c_int), ("n_hidden_layers", ctypes.c_int), ("n_output_dims", ctypes.c_int),
        ("mixture", ctypes.c_int), ("n_lobes", ctypes.c_int), ("mapping", ctypes.c_int), ("numerics", ctypes.c_int),
        ("cosine_kappa", ctypes.c_float), ("grad_clamp", ctypes.c_float), ("clamp_scale", ctypes.c_float),
        ("learn_selection", ctypes.c_int), ("fp16_intermediates", ctypes.c_int),
    ]


def make_cfg(config: dict) -> OrcCfg:
    """Maps the library's JSON configuration onto the oracle's struct (same defaults)."""
    c = OrcCfg()
    c.n_input_dims = int(config.get("n_input_dims", 9))
    enc = config["encoding"]
    nested = enc["nested"] if enc.get("otype", "Composite").lower() == "composite" else [enc]
    c.n_enc = len(nested)
    specified = sum(int(e["n_dims_to_encode"]) for e in nested if "n_dims_to_encode" in e)
    begin = 0
    for i, e in enumerate(nested):
        ot = e.get("otype", "").lower()
        begin = int(e.get("dims_to_encode_begin", begin))
        nd = int(e["n_dims_to_encode"]) if "n_dims_to_encode" in e else c.n_input_dims - specified
        c.enc_begin[i], c.enc_ndims[i] = begin, nd
        if ot in ("hashgrid", "densegrid", "grid"):
            c.enc_type[i] = ENC["grid"]
            gt = e.get("type", "Dense" if ot == "densegrid" else "Hash").lower()
            c.grid_type = 0 if gt == "hash" else 1
            c.hash_type = {"coherentprime": 0, "prime": 1, "reversedprime": 2}[e.get("hash", "CoherentPrime").lower()]
            c.n_features = int(e.get("n_features_per_level", 2))
            c.n_levels = int(e.get("n_levels", 16))
            c.log2_hashmap_size = int(e.get("log2_hashmap_size", 19))
            c.base_resolution = int(e.get("base_resolution", 16))
            c.per_level_scale = float(e.get("per_level_scale", 2.0))
            c.interpolation = {"linear": 0, "smoothstep": 1, "nearest": 2}[e.get("interpolation", "Linear").lower()]
        elif ot == "oneblob":
            c.enc_type[i] = ENC["oneblob"]
            c.enc_param[i] = int(e.get("n_bins", 16))
        elif ot == "sphericalharmonics":
            c.enc_type[i] = ENC["sh"]
            c.enc_param[i] = int(e.get("degree", 2))
        else:
            c.enc_type[i] = ENC["identity"]
        begin += nd
    net = config["network"]
    c.width = int(net.get("n_neurons", 64))
    c.n_hidden_layers = int(net.get("n_hidden_layers", 2))
    mix = config.get("mixture", {})
    c.n_lobes = int(mix.get("n_lobes", 8))
    c.learn_selection = 1 if mix.get("learn_selection", False) else 0
    c.n_output_dims = 4 * c.n_lobes + c.learn_selection
    c.mixture = 0
    c.mapping = 0 if mix.get("mapping", "reference").lower() == "reference" else 1
    c.numerics = 0 if mix.get("numerics", "stable").lower() == "reference" else 1
    c.cosine_kappa = float(mix.get("cosine_kappa", 2.18853))
    c.grad_clamp = float(config.get("loss", {}).get("grad_clamp", 150.0))
    c.clamp_scale = float(config.get("loss", {}).get("clamp_scale", 128.0))
    c.fp16_intermediates = 1 if config.get("loss", {}).get("fp16_intermediates", False) else 0
    return c


_f = ctypes.c_float
_vp = ctypes.c_void_p


def _p(a):
    """Pointer to an array that the caller keeps alive for the duration of the call."""
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class _Keep:
    """Holds converted temporaries so that their memory outlives the C call."""

    def __init__(self):
        self.refs = []

    def __call__(self, a, dtype=np.float32):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype)
        self.refs.append(a)
        return ctypes.c_void_p(a.ctypes.data)


class Oracle:
    def __init__(self, path: str = LIB_PATH):
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built (make -C oracle)")
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.orc_encoded_width.restype = ctypes.c_int
        L.orc_padded_output.restype = ctypes.c_int
        L.orc_grid_layout.restype = ctypes.c_size_t
        L.orc_adam_step.argtypes = [ctypes.c_size_t, ctypes.c_size_t, _f, _f, _f, _f, _f, _f, _vp, _vp, _vp, _vp, _vp]
        L.orc_ema_step.argtypes = [ctypes.c_size_t, _f, ctypes.c_uint32, _vp, _vp]
        L.orc_ema_step_half.argtypes = [ctypes.c_size_t, _f, ctypes.c_uint32, _vp, _vp]
        L.orc_train_step_f.argtypes = [_vp] * 9 + [ctypes.c_int, _f, _f, _f, _f, _f, _f, ctypes.c_int]
        L.orc_train_step_f.restype = _f
        L.orc_adjoint_f.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _f, _vp, _vp]
        L.orc_adjoint_d.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_double, _vp, _vp]
        L.orc_synth_queries.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _vp, _vp, _vp, _vp]
        L.orc_synth_train.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _vp, _vp, _vp, _vp]
        L.orc_field_eval.argtypes = [ctypes.c_uint64, _vp, _vp, ctypes.c_int, _vp]
        L.orc_init_params.argtypes = [_vp, ctypes.c_uint32, ctypes.c_int, _vp]
        L.orc_pack_queries_f.argtypes = [_vp] * 5 + [ctypes.c_int, _vp]
        L.orc_pack_training_f.argtypes = [_vp] * 13 + [ctypes.c_uint32, ctypes.c_uint32, _vp, _vp,
                                                        ctypes.c_uint32] + [_vp] * 5
        L.orc_adjoint_sel_f.argtypes = [_vp] * 6 + [ctypes.c_int, _f, _vp, _vp]
        L.orc_adjoint_sel_d.argtypes = [_vp] * 6 + [ctypes.c_int, ctypes.c_double, _vp, _vp]
        L.orc_selection_f.argtypes = [_vp, _vp, ctypes.c_int, _vp]
        L.orc_pack_training_f.restype = ctypes.c_uint32

    # -- layout -----------------------------------------------------------
    def widths(self, cfg):
        return self.lib.orc_encoded_width(ctypes.byref(cfg)), self.lib.orc_padded_output(ctypes.byref(cfg))

    def n_params(self, cfg):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        self.lib.orc_n_params(ctypes.byref(cfg), ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def grid_layout(self, cfg):
        offs = np.zeros(33, np.uint32)
        scales = np.zeros(32, np.float32)
        res = np.zeros(32, np.uint32)
        self.lib.orc_grid_layout(ctypes.byref(cfg), _p(offs), _p(scales), _p(res))
        L = cfg.n_levels
        return offs[:L + 1], scales[:L], res[:L]

    def grid_corners(self, cfg, inputs):
        k = _Keep()
        n = inputs.shape[1]
        idx = np.zeros((cfg.n_levels, 8, n), np.uint32)
        w = np.zeros((cfg.n_levels, 8, n), np.float32)
        self.lib.orc_grid_corners(ctypes.byref(cfg), k(inputs, np.float32), n, _p(idx), _p(w))
        return idx, w

    # -- forward ----------------------------------------------------------
    def encode(self, cfg, grid_params, inputs, dtype=np.float32):
        k = _Keep()
        n = inputs.shape[1]
        ew, _ = self.widths(cfg)
        out = np.zeros((ew, n), dtype)
        fn = self.lib.orc_encode_f if dtype == np.float32 else self.lib.orc_encode_d
        fn(ctypes.byref(cfg), k(grid_params, np.float32),
           k(inputs, np.float32), n, _p(out))
        return out

    def mlp_forward(self, cfg, mlp_params, enc, acts=False):
        k = _Keep()
        dtype = enc.dtype
        n = enc.shape[1]
        _, po = self.widths(cfg)
        raw = np.zeros((po, n), dtype)
        a = np.zeros((cfg.n_hidden_layers, cfg.width, n), dtype) if acts else None
        fn = self.lib.orc_mlp_forward_f if dtype == np.float32 else self.lib.orc_mlp_forward_d
        fn(ctypes.byref(cfg), k(mlp_params, np.float32), k(enc, enc.dtype), n,
           _p(raw), _p(a))
        return (raw, a) if acts else raw

    def mixture(self, cfg, raw, normal=None, u=None, eval_dirs=None, dtype=np.float32):
        n = raw.shape[1]
        raw = np.ascontiguousarray(raw, dtype)
        n_eval = 0 if eval_dirs is None else eval_dirs.shape[0]
        out_dir = np.zeros((3, n), dtype)
        out_pdf = np.zeros(n, dtype)
        out_eval = np.zeros((max(n_eval, 1), n), dtype)
        lobe = np.zeros(n, np.uint32)
        lobes = np.zeros((cfg.n_lobes * 5, n), dtype)
        fn = self.lib.orc_mixture_f if dtype == np.float32 else self.lib.orc_mixture_d
        c32 = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        normal, u, eval_dirs = c32(normal), c32(u), c32(eval_dirs)
        fn(ctypes.byref(cfg), _p(raw), _p(normal), _p(u), _p(eval_dirs), n_eval, n, _p(out_dir), _p(out_pdf),
           _p(out_eval), _p(lobe), _p(lobes))
        res = {"lobes": lobes}
        if u is not None:
            res.update(dir=out_dir, pdf=out_pdf, lobe=lobe)
        if n_eval:
            res["eval_pdf"] = out_eval[:n_eval]
        return res

    def query(self, cfg, params, inputs, normal=None, u=None, eval_dirs=None, dtype=np.float32, n_threads=8):
        k = _Keep()
        n = inputs.shape[1]
        n_eval = 0 if eval_dirs is None else eval_dirs.shape[0]
        _, po = self.widths(cfg)
        out_dir = np.zeros((3, n), dtype)
        out_pdf = np.zeros(n, dtype)
        out_eval = np.zeros((max(n_eval, 1), n), dtype)
        lobe = np.zeros(n, np.uint32)
        raw = np.zeros((po, n), dtype)
        fn = self.lib.orc_query_f if dtype == np.float32 else self.lib.orc_query_d
        c32 = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        fn(ctypes.byref(cfg), k(c32(params)), k(c32(inputs)), k(c32(normal)), k(c32(u)), k(c32(eval_dirs)),
           n_eval, n, _p(out_dir), _p(out_pdf), _p(out_eval), _p(lobe), _p(raw), n_threads)
        res = {"raw": raw}
        if u is not None:
            res.update(dir=out_dir, pdf=out_pdf, lobe=lobe)
        if n_eval:
            res["eval_pdf"] = out_eval[:n_eval]
        return res

    # -- training ---------------------------------------------------------
    def adjoint(self, cfg, raw, tgt_dir, tgt_rgb, tgt_pdf, loss_scale=1.0, dtype=np.float32):
        k = _Keep()
        n = raw.shape[1]
        _, po = self.widths(cfg)
        dout = np.zeros((po, n), dtype)
        loss = np.zeros(n, dtype)
        fn = self.lib.orc_adjoint_f if dtype == np.float32 else self.lib.orc_adjoint_d
        c32 = lambda a: np.ascontiguousarray(a, np.float32)
        fn(ctypes.byref(cfg), k(raw, dtype), k(c32(tgt_dir)), k(c32(tgt_rgb)),
           k(c32(tgt_pdf)), n, loss_scale, _p(dout), _p(loss))
        return dout, loss

    def adjoint_selection(self, cfg, raw, tgt_dir, tgt_rgb, tgt_pdf, tgt_bsdf, loss_scale=1.0, dtype=np.float32):
        k = _Keep()
        n = raw.shape[1]
        _, po = self.widths(cfg)
        dout = np.zeros((po, n), dtype)
        loss = np.zeros(n, dtype)
        fn = self.lib.orc_adjoint_sel_f if dtype == np.float32 else self.lib.orc_adjoint_sel_d
        c32 = lambda a: np.ascontiguousarray(a, np.float32)
        fn(ctypes.byref(cfg), k(raw, dtype), k(c32(tgt_dir)), k(c32(tgt_rgb)), k(c32(tgt_pdf)), k(c32(tgt_bsdf)),
           n, loss_scale, _p(dout), _p(loss))
        return dout, loss

    def selection(self, cfg, raw):
        k = _Keep()
        out = np.zeros(raw.shape[1], np.float32)
        self.lib.orc_selection_f(ctypes.byref(cfg), k(raw, np.float32), raw.shape[1], _p(out))
        return out

    def mlp_backward(self, cfg, mlp_params, enc, acts, dout):
        k = _Keep()
        dtype = enc.dtype
        n = enc.shape[1]
        n_mlp, _ = self.n_params(cfg)
        dw = np.zeros(n_mlp, dtype)
        denc = np.zeros_like(enc)
        fn = self.lib.orc_mlp_backward_f if dtype == np.float32 else self.lib.orc_mlp_backward_d
        fn(ctypes.byref(cfg), k(mlp_params, np.float32), k(enc, enc.dtype),
           k(acts, acts.dtype), k(dout, dtype), n, _p(dw), _p(denc))
        return dw, denc

    def grid_backward(self, cfg, inputs, denc):
        k = _Keep()
        dtype = denc.dtype
        n = inputs.shape[1]
        _, n_grid = self.n_params(cfg)
        gg = np.zeros(n_grid, dtype)
        fn = self.lib.orc_grid_backward_f if dtype == np.float32 else self.lib.orc_grid_backward_d
        fn(ctypes.byref(cfg), k(inputs, np.float32), k(denc, denc.dtype), n, _p(gg))
        return gg

    def adam_step(self, n_matrix, loss_scale, lr, beta1, beta2, eps, l2, master, grad, m, v, steps):
        self.lib.orc_adam_step(master.size, n_matrix, loss_scale, lr, beta1, beta2, eps, l2, _p(master), _p(grad),
                               _p(m), _p(v), _p(steps))

    def train_step(self, cfg, params, m, v, steps, t, loss_scale=128.0, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-15,
                   l2=1e-6, n_threads=8):
        """One FP32 train step (the CPU baseline); params, m, v, steps updated in place."""
        k = _Keep()
        n = t["inputs"].shape[1]
        return self.lib.orc_train_step_f(ctypes.byref(cfg), _p(params), _p(m), _p(v), _p(steps),
                                         k(t["inputs"]), k(t["tgt_dir"]), k(t["tgt_rgb"]), k(t["tgt_pdf"]), n,
                                         loss_scale, lr, beta1, beta2, eps, l2, n_threads)

    def ema_step(self, decay, step, weights, ema, half=False):
        """half: tcnn's FP16 recurrence (ema_step_half_precision)."""
        (self.lib.orc_ema_step_half if half else self.lib.orc_ema_step)(weights.size, decay, step, _p(weights),
                                                                       _p(ema))

    def init_params(self, cfg, seed, passes=2):
        n_mlp, n_grid = self.n_params(cfg)
        p = np.zeros(n_mlp + n_grid, np.float32)
        self.lib.orc_init_params(ctypes.byref(cfg), seed, passes, _p(p))
        return p

    # -- synthetic records --------------------------------------------------
    def synth_queries(self, seed, n, first=0):
        a = dict(inputs=np.zeros((9, n), np.float32), normal=np.zeros((3, n), np.float32),
                 u=np.zeros((3, n), np.float32), eval_dirs=np.zeros((4, 3, n), np.float32))
        self.lib.orc_synth_queries(seed, first, n, _p(a["inputs"]), _p(a["normal"]), _p(a["u"]), _p(a["eval_dirs"]))
        return a

    def synth_train(self, seed, n, first=0):
        a = dict(inputs=np.zeros((9, n), np.float32), tgt_dir=np.zeros((2, n), np.float32),
                 tgt_rgb=np.zeros((3, n), np.float32), tgt_pdf=np.zeros(n, np.float32))
        self.lib.orc_synth_train(seed, first, n, _p(a["inputs"]), _p(a["tgt_dir"]), _p(a["tgt_rgb"]),
                                 _p(a["tgt_pdf"]))
        return a

    # -- wavefront packers -----------------------------------------------------
    def pack_queries(self, box, pos, wo, normal, roughness):
        k = _Keep()
        n = pos.shape[1]
        out = np.zeros((9, n), np.float32)
        self.lib.orc_pack_queries_f(k(box, np.float32), k(pos, np.float32), k(wo, np.float32),
                                    k(normal, np.float32), k(roughness, np.float32), n, _p(out))
        return out

    PACK_FIELDS = ("L", "thp", "pos", "dir", "wi_pdf", "mis_weight", "delta", "bsdf_val", "wo", "normal",
                   "roughness", "depth", "bsdf_pdf")

    def pack_training(self, rec, tile, box, capacity):
        """rec: dict of numpy arrays [D][C][P] (delta uint8, depth uint32 [P],
        optional bsdf_pdf [D][P]); tile: dict(offset, stride, res, resolution).
        Returns (in, dir, rgb, pdf, count) plus bsdf_pdf when rec has it."""
        k = _Keep()
        ptrs = [k(rec.get(f), np.uint8 if f == "delta" else (np.uint32 if f == "depth" else np.float32))
                for f in self.PACK_FIELDS]
        t = np.array(list(tile["offset"]) + list(tile["stride"]) + list(tile["res"]) + list(tile["resolution"]),
                     np.uint32)
        out = (np.zeros((9, capacity), np.float32), np.zeros((2, capacity), np.float32),
               np.zeros((3, capacity), np.float32), np.zeros(capacity, np.float32))
        bsdf = np.zeros(capacity, np.float32) if rec.get("bsdf_pdf") is not None else None
        cnt = self.lib.orc_pack_training_f(*ptrs, rec["L"].shape[0], rec["depth"].shape[0], k(t, np.uint32),
                                           k(box, np.float32), capacity, *[_p(o) for o in out], _p(bsdf))
        return out + (int(cnt),) + ((bsdf,) if bsdf is not None else ())

    def field_eval(self, seed, pos, dirs):
        k = _Keep()
        n = pos.shape[1]
        f = np.zeros(n, np.float64)
        self.lib.orc_field_eval(seed, k(pos, np.float32),
                                k(dirs, np.float32), n, _p(f))
        return f

    def math(self, fn, x, y=None):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros_like(x)
        y = None if y is None else np.ascontiguousarray(y, np.float32)
        self.lib.orc_math(fn, _p(x), _p(y), x.size, _p(out))
        return out
