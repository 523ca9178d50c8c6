"""Python harness of the MI355X neural parametric mixture guiding library.

The product is the C ABI in ``include/npmg.h`` implemented by ``libnpmg.so``
(hand-written HIP for gfx950). This module binds it with ctypes for tests,
benchmarks and autograd wrapping; PyTorch is used only for device memory and
streams. The classes mirror the interface the reference's guided integrator
uses (``src/render/guided/integrator.cpp:555-702``): a network with
``inference``/``forward``/``backward``, a trainer ``optimizer_step`` and the
mixture consumers' ``sample``/``pdf``.

There is no CPU fallback: every call goes to the HIP library, and importing a
missing or unbuilt library raises ``ImportError``.
"""
from __future__ import annotations

import ctypes
import json
import os
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnpmg.so")
HEADER_PATH = os.path.normpath(os.path.join(_HERE, "..", "..", "include", "npmg.h"))

NPMG_OK, NPMG_EINVAL, NPMG_ENOMEM, NPMG_EHIP, NPMG_ESTATE = range(5)
NPMG_COSINE_AWARE = 1
NPMG_TIME_GRID_ENCODE, NPMG_TIME_QUERY, NPMG_TIME_TRAIN_STEP = 0, 1, 2

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_f32 = ctypes.c_float
_int = ctypes.c_int
_sz = ctypes.c_size_t

class NpmgQueue(ctypes.Structure):
    """npmg_queue of include/npmg.h: one queue's SoA buffers, row stride `capacity`."""
    _fields_ = [("in_", _vp), ("normal", _vp), ("u", _vp), ("eval_dirs", _vp), ("out_dir", _vp),
                ("out_pdf", _vp), ("out_eval_pdf", _vp), ("out_lobe", _vp), ("capacity", _u32),
                ("out_selection", _vp)]


class NpmgPathRecords(ctypes.Structure):
    """npmg_path_records of include/npmg.h (device pointers, SoA per depth)."""
    FIELDS = ("L", "thp", "pos", "dir", "wi_pdf", "mis_weight", "delta", "bsdf_val", "wo", "normal", "roughness",
              "depth", "bsdf_pdf")
    _fields_ = [(k, _vp) for k in FIELDS] + [("max_depth", _u32), ("n_pixels", _u32)]


class NpmgTrainTile(ctypes.Structure):
    _fields_ = [("offset", _u32 * 2), ("stride", _u32 * 2), ("res", _u32 * 2), ("resolution", _u32 * 2)]


# name -> (restype, argtypes); every function declared in include/npmg.h.
SIGNATURES = {
    "npmg_create": (_int, [ctypes.c_char_p, _u64, ctypes.POINTER(_vp)]),
    "npmg_destroy": (None, [_vp]),
    "npmg_last_error": (ctypes.c_char_p, [_vp]),
    "npmg_n_params": (_int, [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
    "npmg_widths": (_int, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "npmg_grid_layout": (_int, [_vp, ctypes.POINTER(_u32), _vp, _vp, _vp]),
    "npmg_set_params": (_int, [_vp, _vp, _sz, _int]),
    "npmg_get_params": (_int, [_vp, _vp, _sz, _int, _int]),
    "npmg_set_params_which": (_int, [_vp, _vp, _sz, _int, _int]),
    "npmg_infer": (_int, [_vp, _vp, _vp, _u32, _vp]),
    "npmg_sample_pdf": (_int, [_vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp]),
    "npmg_mixture_sample_pdf": (_int, [_vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp]),
    "npmg_train_step": (_int, [_vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp]),
    "npmg_encode": (_int, [_vp, _vp, _vp, _u32, _int, _vp]),
    "npmg_grid_corners": (_int, [_vp, _vp, _vp, _vp, _u32, _vp]),
    "npmg_forward": (_int, [_vp, _vp, _vp, _u32, _int, _vp]),
    "npmg_mixture_adjoint": (_int, [_vp, _vp, _vp, _vp, _vp, _u32, _f32, _vp, _vp, _vp]),
    "npmg_backward": (_int, [_vp, _vp, _vp, _u32, _vp, _vp, _vp]),
    "npmg_optimizer_step": (_int, [_vp, _vp, _f32, _vp]),
    "npmg_train_state": (_int, [_vp, ctypes.POINTER(_f32), ctypes.POINTER(_u32)]),
    "npmg_synth_queries": (_int, [_u64, _u64, _u32, _vp, _vp, _vp, _vp, _vp]),
    "npmg_synth_train": (_int, [_u64, _u64, _u32, _vp, _vp, _vp, _vp, _vp]),
    "npmg_debug_math": (_int, [_int, _vp, _vp, _u32, _vp, _vp]),
    "npmg_train_step_graph": (_int, [_vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp]),
    "npmg_timing_enable": (_int, [_vp, _int]),
    "npmg_timing_read": (_int, [_vp, _u32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_u32)]),
    "npmg_validate_pdf_integral": (_int, [_vp, _vp, _vp, _u32, _u32, _u32, _vp, _vp]),
    "npmg_validate_loss_fd": (_int, [_vp, _vp, _vp, _vp, _vp, _u32, ctypes.c_double, _vp, _vp]),
    "npmg_sample_pdf_queues": (_int, [_vp, ctypes.POINTER(NpmgQueue), _vp, _u32, _u32, _u32, _vp]),
    "npmg_autotune": (_int, [_vp, _u32, _vp]),
    "npmg_fp8_calibrate": (_int, [_vp, _vp, _u32, _vp]),
    "npmg_snapshot_save": (_int, [_vp, _int, _vp, _sz, ctypes.POINTER(_sz)]),
    "npmg_snapshot_load": (_int, [_vp, _vp, _sz]),
    "npmg_pack_queries": (_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(_f32), _u32, _vp, _vp]),
    "npmg_pack_training": (_int, [_vp, ctypes.POINTER(NpmgPathRecords), ctypes.POINTER(NpmgTrainTile),
                                  ctypes.POINTER(_f32), _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "npmg_sample_pdf_ex": (_int, [_vp, ctypes.POINTER(NpmgQueue), _u32, _u32, _u32, _vp]),
    "npmg_sample_pdf_noalloc": (_int, [_vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp]),
    "npmg_train_step_selection": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _int, _vp, _vp]),
    "npmg_mixture_adjoint_selection": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _f32, _vp, _vp, _vp]),
    "npmg_selection_probability": (_int, [_vp, _vp, _vp, _u32, _vp]),
    "npmg_training_stride": (_int, [ctypes.POINTER(_u32), ctypes.POINTER(_u32), _u32, _u32, _u32,
                                    ctypes.POINTER(_u32)]),
    "npmg_train_batches": (_int, [_u32, _u32, _u32, _u32, _u32, ctypes.POINTER(_u32), ctypes.POINTER(_u32), _u32,
                                  ctypes.POINTER(_u32)]),
    "npmg_train_records": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _vp,
                                  ctypes.POINTER(_u32), _vp]),
    "npmg_fp8_scales": (_int, [_vp, _vp, _vp, _u32]),
    "npmg_tuning": (_int, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "npmg_set_tuning": (_int, [_vp, _u32, _u32]),
    "npmg_queue_tuning": (_int, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "npmg_set_queue_tuning": (_int, [_vp, _u32, _u32]),
}


class NpmgError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"npmg status {status}: {message}")
        self.status = status


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run __graft_entry__.build() (no CPU fallback exists)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


def _ptr(t) -> Optional[int]:
    if t is None:
        return None
    import torch

    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor")
    if not t.is_cuda:
        raise ValueError("device tensor required")
    if not t.is_contiguous():
        raise ValueError("contiguous tensor required")
    return t.data_ptr()


def _stream():
    import torch

    return torch.cuda.current_stream().cuda_stream


# ---------------------------------------------------------------------------
# Configurations

def reference_config() -> dict:
    """The reference's hyper-parameters (common/configs/nn/base.json "nn")."""
    return {
        "loss": {"otype": "L2"},
        "optimizer": {"otype": "Ema", "decay": 0.95,
                      "nested": {"otype": "Adam", "learning_rate": 1e-2, "beta1": 0.9, "beta2": 0.99,
                                 "epsilon": 1e-15, "l2_reg": 1e-6, "adabound": False}},
        "encoding": {"otype": "Composite", "nested": [
            {"otype": "DenseGrid", "n_levels": 8, "n_features_per_level": 4, "base_resolution": 8,
             "per_level_scale": 1.405, "interpolation": "Smoothstep", "n_dims_to_encode": 3},
            {"n_dims_to_encode": 3, "otype": "SphericalHarmonics", "degree": 2},
            {"n_dims_to_encode": 3, "otype": "OneBlob", "n_bins": 4}]},
        "network": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "None",
                    "n_neurons": 64, "n_hidden_layers": 3},
        "mixture": {"otype": "vMF", "n_lobes": 8, "mapping": "reference", "numerics": "reference"},
    }


def benchmark_config(size: str = "full", precision: str = "fp32", n_lobes: int = 8, width: int = 64,
                     hidden: Optional[int] = None) -> dict:
    """Hash-grid benchmark configurations: "small" (8 levels x 2^15) or "full" (16 levels x 2^19)."""
    small = size == "small"
    grid = {"otype": "HashGrid", "n_dims_to_encode": 3, "n_levels": 8 if small else 16,
            "n_features_per_level": 2, "log2_hashmap_size": 15 if small else 19,
            "base_resolution": 16, "per_level_scale": 1.5, "interpolation": "Linear"}
    return {
        "encoding": {"otype": "Composite", "nested": [
            grid, {"otype": "OneBlob", "n_dims_to_encode": 6, "n_bins": 16}]},
        "network": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "None",
                    "n_neurons": width, "n_hidden_layers": hidden if hidden is not None else (2 if small else 3)},
        "optimizer": {"otype": "Adam", "learning_rate": 1e-2, "beta1": 0.9, "beta2": 0.99, "epsilon": 1e-15,
                      "l2_reg": 1e-6},
        "mixture": {"otype": "vMF", "n_lobes": n_lobes, "mapping": "softmax", "numerics": "stable"},
        # FP32: the reference's static scale; BF16/FP16: dynamic from 2^15 (the
        # gradient clamp follows the scale, so both train the same function)
        "loss": ({"scale": 128.0, "grad_clamp": 150.0} if precision == "fp32" else
                 {"scale": 32768.0, "dynamic": True, "scale_window": 1000, "grad_clamp": 150.0}),
        "precision": precision,
    }


# ---------------------------------------------------------------------------

class GuidingNetwork:
    """Owns one npmg context: the guiding network, its optimizer and workspaces."""

    def __init__(self, config: dict, seed: int = 7272):
        self._lib = lib()
        self.config = dict(config)
        h = _vp()
        st = self._lib.npmg_create(json.dumps(self.config).encode(), seed, ctypes.byref(h))
        if st != NPMG_OK:
            raise NpmgError(st, self._lib.npmg_last_error(None).decode())
        self._h = h
        a, b = _sz(), _sz()
        self._check(self._lib.npmg_n_params(self._h, ctypes.byref(a), ctypes.byref(b)))
        self.n_mlp, self.n_grid = a.value, b.value
        self.n_params = self.n_mlp + self.n_grid
        ew, po = _u32(), _u32()
        self._check(self._lib.npmg_widths(self._h, ctypes.byref(ew), ctypes.byref(po)))
        self.encoded_width, self.padded_output = ew.value, po.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.npmg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int):
        if st != NPMG_OK:
            raise NpmgError(st, self._lib.npmg_last_error(self._h).decode())

    # -- layout ---------------------------------------------------------------
    def grid_layout(self):
        import numpy as np

        n = _u32()
        offs = np.zeros(33, np.uint32)
        scales = np.zeros(32, np.float32)
        res = np.zeros(32, np.uint32)
        self._check(self._lib.npmg_grid_layout(self._h, ctypes.byref(n), offs.ctypes.data, scales.ctypes.data,
                                               res.ctypes.data))
        L = n.value
        return offs[:L + 1].copy(), scales[:L].copy(), res[:L].copy()

    # -- parameters -----------------------------------------------------------
    def get_params(self, which: int = 0):
        import torch

        out = torch.empty(self.n_params, dtype=torch.float32, device="cuda")
        self._check(self._lib.npmg_get_params(self._h, _ptr(out), self.n_params, 1, which))
        return out

    def set_params(self, params, which: int = 2):
        """which: 0 the training master only, 1 the inference (EMA) copy only, 2 both."""
        import torch

        p = params.to(device="cuda", dtype=torch.float32).contiguous()
        if which == 2:
            self._check(self._lib.npmg_set_params(self._h, _ptr(p), self.n_params, 1))
        else:
            self._check(self._lib.npmg_set_params_which(self._h, _ptr(p), self.n_params, 1, which))

    # -- inference ------------------------------------------------------------
    def inference(self, inputs):
        """Network::inference: in [9][n] -> raw outputs [n_out][n] (inference/EMA weights)."""
        return self.forward(inputs, use_inference_params=True)

    def forward(self, inputs, use_inference_params: bool = False):
        import torch

        n = inputs.shape[1]
        out = torch.empty(self.padded_output, n, dtype=torch.float32, device="cuda")
        self._check(self._lib.npmg_forward(self._h, _ptr(inputs), _ptr(out), n, int(use_inference_params),
                                           _stream()))
        return out

    def encode(self, inputs, use_inference_params: bool = False):
        import torch

        n = inputs.shape[1]
        out = torch.empty(self.encoded_width, n, dtype=torch.float32, device="cuda")
        self._check(self._lib.npmg_encode(self._h, _ptr(inputs), _ptr(out), n, int(use_inference_params), _stream()))
        return out

    def grid_corners(self, inputs):
        import torch

        n = inputs.shape[1]
        L = len(self.grid_layout()[1])
        idx = torch.empty(L, 8, n, dtype=torch.int32, device="cuda")
        w = torch.empty(L, 8, n, dtype=torch.float32, device="cuda")
        self._check(self._lib.npmg_grid_corners(self._h, _ptr(inputs), _ptr(idx), _ptr(w), n, _stream()))
        return idx, w

    def sample_pdf(self, inputs, normal=None, u=None, eval_dirs=None, cosine_aware: bool = True,
                   raw: bool = False):
        """Fused encode -> MLP -> mixture -> sample + pdfs (infer + sample + pdf)."""
        import torch

        n = inputs.shape[1]
        n_eval = 0 if eval_dirs is None else eval_dirs.shape[0]
        dev = "cuda"
        out = {}
        if u is not None:
            out["dir"] = torch.empty(3, n, dtype=torch.float32, device=dev)
            out["pdf"] = torch.empty(n, dtype=torch.float32, device=dev)
            out["lobe"] = torch.empty(n, dtype=torch.int32, device=dev)
        if n_eval:
            out["eval_pdf"] = torch.empty(n_eval, n, dtype=torch.float32, device=dev)
        if raw:
            out["raw"] = torch.empty(self.padded_output, n, dtype=torch.float32, device=dev)
        flags = NPMG_COSINE_AWARE if (cosine_aware and normal is not None) else 0
        self._check(self._lib.npmg_sample_pdf(
            self._h, _ptr(inputs), _ptr(normal) if flags else None, _ptr(u), _ptr(eval_dirs), n_eval,
            _ptr(out.get("dir")), _ptr(out.get("pdf")), _ptr(out.get("eval_pdf")), _ptr(out.get("lobe")),
            _ptr(out.get("raw")), n, flags, _stream()))
        return out

    def mixture_sample_pdf(self, raw_out, normal=None, u=None, eval_dirs=None, cosine_aware: bool = True,
                           lobes: bool = False):
        """Consumer side: raw outputs -> mixture -> sample + pdfs (+ lobe parameters)."""
        import torch

        n = raw_out.shape[1]
        n_eval = 0 if eval_dirs is None else eval_dirs.shape[0]
        K = self.config.get("mixture", {}).get("n_lobes", 8)
        dev = "cuda"
        out = {}
        if u is not None:
            out["dir"] = torch.empty(3, n, dtype=torch.float32, device=dev)
            out["pdf"] = torch.empty(n, dtype=torch.float32, device=dev)
            out["lobe"] = torch.empty(n, dtype=torch.int32, device=dev)
        if n_eval:
            out["eval_pdf"] = torch.empty(n_eval, n, dtype=torch.float32, device=dev)
        if lobes:
            out["lobes"] = torch.empty(K * 5, n, dtype=torch.float32, device=dev)
        flags = NPMG_COSINE_AWARE if (cosine_aware and normal is not None) else 0
        self._check(self._lib.npmg_mixture_sample_pdf(
            self._h, _ptr(raw_out), _ptr(normal) if flags else None, _ptr(u), _ptr(eval_dirs), n_eval,
            _ptr(out.get("dir")), _ptr(out.get("pdf")), _ptr(out.get("eval_pdf")), _ptr(out.get("lobe")),
            _ptr(out.get("lobes")), n, flags, _stream()))
        return out

    # -- training -------------------------------------------------------------
    def mixture_adjoint(self, raw_out, tgt_dir, tgt_rgb, tgt_pdf, loss_scale: float = 1.0):
        import torch

        n = raw_out.shape[1]
        dout = torch.empty(self.padded_output, n, dtype=torch.float32, device="cuda")
        loss = torch.empty(n, dtype=torch.float32, device="cuda")
        self._check(self._lib.npmg_mixture_adjoint(self._h, _ptr(raw_out), _ptr(tgt_dir), _ptr(tgt_rgb),
                                                   _ptr(tgt_pdf), n, float(loss_scale), _ptr(dout), _ptr(loss),
                                                   _stream()))
        return dout, loss

    def backward(self, inputs, dout, want_denc: bool = False):
        import torch

        n = inputs.shape[1]
        dparams = torch.empty(self.n_params, dtype=torch.float32, device="cuda")
        denc = torch.empty(self.encoded_width, n, dtype=torch.float32, device="cuda") if want_denc else None
        self._check(self._lib.npmg_backward(self._h, _ptr(inputs), _ptr(dout.contiguous()), n, _ptr(dparams),
                                            _ptr(denc), _stream()))
        return dparams, denc

    def optimizer_step(self, dparams, loss_scale: float):
        """Trainer::optimizer_step (trainer.h:153-155)."""
        self._check(self._lib.npmg_optimizer_step(self._h, _ptr(dparams), float(loss_scale), _stream()))

    def train_step(self, inputs, tgt_dir, tgt_rgb, tgt_pdf, loss_out=None):
        """One full training step; returns the device scalar with the mean loss."""
        import torch

        n = inputs.shape[1]
        if loss_out is None:
            loss_out = torch.empty(1, dtype=torch.float32, device="cuda")
        self._check(self._lib.npmg_train_step(self._h, _ptr(inputs), _ptr(tgt_dir), _ptr(tgt_rgb), _ptr(tgt_pdf),
                                               n, _ptr(loss_out), _stream()))
        return loss_out

    # -- FP64 validation ---------------------------------------------------------
    def pdf_integral(self, raw_out, normal=None, n_z: int = 1024, n_phi: int = 2048):
        """Sphere integral (FP64 quadrature) of each query's mixture pdf."""
        import torch

        n = raw_out.shape[1]
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        self._check(self._lib.npmg_validate_pdf_integral(self._h, _ptr(raw_out), _ptr(normal), n, n_z, n_phi,
                                                         _ptr(out), _stream()))
        return out

    def loss_fd(self, raw_out, tgt_dir, tgt_rgb, tgt_pdf, h: float = 1e-4):
        """FP64 central differences of the per-sample loss with respect to the raw outputs."""
        import torch

        n = raw_out.shape[1]
        out = torch.empty(self.padded_output, n, dtype=torch.float64, device="cuda")
        self._check(self._lib.npmg_validate_loss_fd(self._h, _ptr(raw_out), _ptr(tgt_dir), _ptr(tgt_rgb),
                                                    _ptr(tgt_pdf), n, float(h), _ptr(out), _stream()))
        return out

    def sample_pdf_queues(self, queues, counts, n_eval: int = 0, cosine_aware: bool = True, selection: bool = False):
        """sample_pdf over several queues in one persistent launch.

        queues: list of dicts with "inputs" [9][cap] and optionally "normal",
        "u" [3][cap], "eval_dirs" [n_eval][3][cap]; counts: device uint32/int32
        tensor [len(queues)] of live queries per queue. Returns one output dict
        per queue (rows of stride cap; entries past the count are untouched).
        """
        import torch

        nq = len(queues)
        flags = NPMG_COSINE_AWARE if cosine_aware and nq and queues[0].get("normal") is not None else 0
        arr = (NpmgQueue * max(nq, 1))()
        outs = []
        for i, q in enumerate(queues):
            cap = q["inputs"].shape[1]
            o = {}
            if q.get("u") is not None:
                o["dir"] = torch.zeros(3, cap, dtype=torch.float32, device="cuda")
                o["pdf"] = torch.zeros(cap, dtype=torch.float32, device="cuda")
                o["lobe"] = torch.zeros(cap, dtype=torch.int32, device="cuda")
            if n_eval:
                o["eval_pdf"] = torch.zeros(n_eval, cap, dtype=torch.float32, device="cuda")
            if selection:
                o["selection"] = torch.zeros(cap, dtype=torch.float32, device="cuda")
            arr[i] = NpmgQueue(_ptr(q["inputs"]), _ptr(q.get("normal")) if flags else None, _ptr(q.get("u")),
                               _ptr(q.get("eval_dirs")), _ptr(o.get("dir")), _ptr(o.get("pdf")),
                               _ptr(o.get("eval_pdf")), _ptr(o.get("lobe")), cap, _ptr(o.get("selection")))
            outs.append(o)
        self._check(self._lib.npmg_sample_pdf_queues(self._h, arr, _ptr(counts), nq, n_eval, flags, _stream()))
        return outs

    def sample_pdf_ex(self, inputs, normal=None, u=None, eval_dirs=None, cosine_aware: bool = True,
                      selection: bool = True):
        """sample_pdf through npmg_sample_pdf_ex, with the learned selection probability."""
        import torch

        n = inputs.shape[1]
        n_eval = 0 if eval_dirs is None else eval_dirs.shape[0]
        o = {}
        if u is not None:
            o["dir"] = torch.empty(3, n, device="cuda")
            o["pdf"] = torch.empty(n, device="cuda")
            o["lobe"] = torch.empty(n, dtype=torch.int32, device="cuda")
        if n_eval:
            o["eval_pdf"] = torch.empty(n_eval, n, device="cuda")
        if selection:
            o["selection"] = torch.empty(n, device="cuda")
        flags = NPMG_COSINE_AWARE if (cosine_aware and normal is not None) else 0
        q = NpmgQueue(_ptr(inputs), _ptr(normal) if flags else None, _ptr(u), _ptr(eval_dirs), _ptr(o.get("dir")),
                      _ptr(o.get("pdf")), _ptr(o.get("eval_pdf")), _ptr(o.get("lobe")), n, _ptr(o.get("selection")))
        self._check(self._lib.npmg_sample_pdf_ex(self._h, ctypes.byref(q), n, n_eval, flags, _stream()))
        return o

    def selection_probability(self, raw_out):
        import torch

        out = torch.empty(raw_out.shape[1], device="cuda")
        self._check(self._lib.npmg_selection_probability(self._h, _ptr(raw_out), _ptr(out), raw_out.shape[1],
                                                         _stream()))
        return out

    def train_step_selection(self, inputs, tgt_dir, tgt_rgb, tgt_pdf, tgt_bsdf_pdf, loss_out=None, graph=False):
        n = inputs.shape[1]
        self._check(self._lib.npmg_train_step_selection(self._h, _ptr(inputs), _ptr(tgt_dir), _ptr(tgt_rgb),
                                                        _ptr(tgt_pdf), _ptr(tgt_bsdf_pdf), n, int(graph),
                                                        _ptr(loss_out), _stream()))
        return loss_out

    def mixture_adjoint_selection(self, raw_out, tgt_dir, tgt_rgb, tgt_pdf, tgt_bsdf_pdf, loss_scale=1.0):
        import torch

        n = raw_out.shape[1]
        dout = torch.empty(self.padded_output, n, device="cuda")
        loss = torch.empty(n, device="cuda")
        self._check(self._lib.npmg_mixture_adjoint_selection(self._h, _ptr(raw_out), _ptr(tgt_dir), _ptr(tgt_rgb),
                                                             _ptr(tgt_pdf), _ptr(tgt_bsdf_pdf), n, loss_scale,
                                                             _ptr(dout), _ptr(loss), _stream()))
        return dout, loss

    def train_records(self, inputs, tgt_dir, tgt_rgb, tgt_pdf, n_records, batch_size=524288, batches_per_frame=8,
                      granularity=128, min_batch=65536, tgt_bsdf_pdf=None, losses=None):
        """trainStep's batch loop over packed records (row stride inputs.shape[1]); returns the batch count."""
        nb = _u32()
        self._check(self._lib.npmg_train_records(self._h, _ptr(inputs), _ptr(tgt_dir), _ptr(tgt_rgb), _ptr(tgt_pdf),
                                                 _ptr(tgt_bsdf_pdf), inputs.shape[1], n_records, batch_size,
                                                 batches_per_frame, granularity, min_batch, _ptr(losses),
                                                 ctypes.byref(nb), _stream()))
        return nb.value

    def autotune(self, n: int):
        """Times the query path's launch shapes for batches of about n; returns (chunks, blocks_per_cu)."""
        self._check(self._lib.npmg_autotune(self._h, n, _stream()))
        return self.tuning()

    def fp8_calibrate(self, inputs):
        self._check(self._lib.npmg_fp8_calibrate(self._h, _ptr(inputs), inputs.shape[1], _stream()))

    def fp8_scales(self):
        """(weight scales, input scales) per layer of the FP8 inference path."""
        L = self.config["network"]["n_hidden_layers"] + 1
        w = (_f32 * L)()
        a = (_f32 * L)()
        self._check(self._lib.npmg_fp8_scales(self._h, w, a, L))
        return list(w), list(a)

    def pack_training(self, records, tile, box, capacity):
        """generate_training_data: records = dict of device tensors named as the
        npmg_path_records fields ([D][C][P] float, delta uint8, depth int32 [P],
        optional bsdf_pdf [D][P]); tile = dict(offset, stride, res, resolution)
        of pairs; box = 6 floats. Returns (in [9][cap], tgt_dir [2][cap],
        tgt_rgb [3][cap], tgt_pdf [cap], count) plus tgt_bsdf_pdf [cap] when
        the records hold bsdf_pdf."""
        import torch

        d = "cuda"
        out = (torch.zeros(9, capacity, device=d), torch.zeros(2, capacity, device=d),
               torch.zeros(3, capacity, device=d), torch.zeros(capacity, device=d))
        bsdf = torch.zeros(capacity, device=d) if records.get("bsdf_pdf") is not None else None
        count = torch.zeros(1, dtype=torch.int32, device=d)
        r = NpmgPathRecords(*[_ptr(records.get(k)) for k in NpmgPathRecords.FIELDS],
                            records["L"].shape[0], records["depth"].shape[0])
        t = NpmgTrainTile((_u32 * 2)(*tile["offset"]), (_u32 * 2)(*tile["stride"]), (_u32 * 2)(*tile["res"]),
                          (_u32 * 2)(*tile["resolution"]))
        b = (_f32 * 6)(*box)
        self._check(self._lib.npmg_pack_training(self._h, ctypes.byref(r), ctypes.byref(t), b, capacity,
                                                 *[_ptr(o) for o in out], _ptr(bsdf), _ptr(count), _stream()))
        return out + (count,) + ((bsdf,) if bsdf is not None else ())

    def snapshot(self, with_optimizer: bool = True) -> bytes:
        """tcnn-format MessagePack snapshot (Trainer::serialize) as bytes."""
        size = _sz()
        self._check(self._lib.npmg_snapshot_save(self._h, int(with_optimizer), None, 0, ctypes.byref(size)))
        buf = ctypes.create_string_buffer(size.value)
        self._check(self._lib.npmg_snapshot_save(self._h, int(with_optimizer), buf, size.value, ctypes.byref(size)))
        return buf.raw[:size.value]

    def load_snapshot(self, data: bytes):
        buf = ctypes.create_string_buffer(data, len(data))
        self._check(self._lib.npmg_snapshot_load(self._h, buf, len(data)))

    def tuning(self):
        c, b = _u32(), _u32()
        self._check(self._lib.npmg_tuning(self._h, ctypes.byref(c), ctypes.byref(b)))
        return c.value, b.value

    def queue_tuning(self):
        """(64-query tiles, grid-encode items) claimed per work-counter atomic by queue launches."""
        t, i = _u32(), _u32()
        self._check(self._lib.npmg_queue_tuning(self._h, ctypes.byref(t), ctypes.byref(i)))
        return t.value, i.value

    def set_queue_tuning(self, tiles_per_claim: int, items_per_claim: int):
        self._check(self._lib.npmg_set_queue_tuning(self._h, int(tiles_per_claim), int(items_per_claim)))

    def set_tuning(self, chunks: int, blocks_per_cu: int):
        """Sets the query path's launch shape (e.g. a saved autotune result)."""
        self._check(self._lib.npmg_set_tuning(self._h, int(chunks), int(blocks_per_cu)))

    def train_step_graph(self, inputs, tgt_dir, tgt_rgb, tgt_pdf, loss_out):
        """train_step as one captured HIP graph (replayed while the buffers are the same)."""
        n = inputs.shape[1]
        self._check(self._lib.npmg_train_step_graph(self._h, _ptr(inputs), _ptr(tgt_dir), _ptr(tgt_rgb),
                                                     _ptr(tgt_pdf), n, _ptr(loss_out), _stream()))
        return loss_out

    def timing(self, enable: bool):
        self._check(self._lib.npmg_timing_enable(self._h, int(enable)))

    def timing_read(self, tag: int):
        """(total ms, launches) of a timing tag since the last read (synchronises)."""
        ms, cnt = ctypes.c_double(), _u32()
        self._check(self._lib.npmg_timing_read(self._h, tag, ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def train_state(self):
        s, t = _f32(), _u32()
        self._check(self._lib.npmg_train_state(self._h, ctypes.byref(s), ctypes.byref(t)))
        return s.value, t.value


# ---------------------------------------------------------------------------
# Synthetic wavefront records

def synth_queries(seed: int, n: int, first: int = 0):
    import torch

    d = "cuda"
    t = dict(inputs=torch.empty(9, n, device=d), normal=torch.empty(3, n, device=d),
             u=torch.empty(3, n, device=d), eval_dirs=torch.empty(4, 3, n, device=d))
    st = lib().npmg_synth_queries(seed, first, n, _ptr(t["inputs"]), _ptr(t["normal"]), _ptr(t["u"]),
                                  _ptr(t["eval_dirs"]), _stream())
    if st != NPMG_OK:
        raise NpmgError(st, "npmg_synth_queries failed")
    return t


def synth_train(seed: int, n: int, first: int = 0):
    import torch

    d = "cuda"
    t = dict(inputs=torch.empty(9, n, device=d), tgt_dir=torch.empty(2, n, device=d),
             tgt_rgb=torch.empty(3, n, device=d), tgt_pdf=torch.empty(n, device=d))
    st = lib().npmg_synth_train(seed, first, n, _ptr(t["inputs"]), _ptr(t["tgt_dir"]), _ptr(t["tgt_rgb"]),
                                _ptr(t["tgt_pdf"]), _stream())
    if st != NPMG_OK:
        raise NpmgError(st, "npmg_synth_train failed")
    return t


def pack_queries(pos, wo, normal, roughness, box):
    """generate_inference_data: shading points [3][n] x 3 + roughness [n] -> in [9][n]."""
    import torch

    n = pos.shape[1]
    out = torch.empty(9, n, device="cuda")
    b = (_f32 * 6)(*box)
    st = lib().npmg_pack_queries(_ptr(pos), _ptr(wo), _ptr(normal), _ptr(roughness), b, n, _ptr(out), _stream())
    if st != NPMG_OK:
        raise NpmgError(st, "npmg_pack_queries failed")
    return out


def training_stride(cur, min_stride, batch_size, batches_per_frame, n_training_samples):
    """updateTrainingStride: (accepted, new stride)."""
    out = (_u32 * 2)()
    acc = lib().npmg_training_stride((_u32 * 2)(*cur), (_u32 * 2)(*min_stride), batch_size, batches_per_frame,
                                     n_training_samples, out)
    return bool(acc), (out[0], out[1])


def train_batches(n_samples, batch_size=524288, batches_per_frame=8, granularity=128, min_batch=65536):
    """The trainStep batch policy: [(offset, size), ...]."""
    m = max(batches_per_frame, 1)
    off, size, nb = (_u32 * m)(), (_u32 * m)(), _u32()
    st = lib().npmg_train_batches(n_samples, batch_size, batches_per_frame, granularity, min_batch, off, size, m,
                                  ctypes.byref(nb))
    if st != NPMG_OK:
        raise NpmgError(st, "npmg_train_batches failed")
    return [(off[i], size[i]) for i in range(nb.value)]


def debug_math(fn: int, x, y=None):
    import torch

    out = torch.empty_like(x)
    st = lib().npmg_debug_math(fn, _ptr(x), _ptr(y), x.numel(), _ptr(out), _stream())
    if st != NPMG_OK:
        raise NpmgError(st, "npmg_debug_math failed")
    return out


from .autograd import NetworkFunction, network_apply  # noqa: E402,F401
