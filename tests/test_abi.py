"""The C ABI library loads on a CPU-only host and exports every symbol include/npmg.h declares."""
import ctypes
import json
import re

import pytest

import npmg


def _header_functions():
    text = open(npmg.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(npmg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header():
    lib = npmg.load_library()
    names = _header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(npmg.SIGNATURES), set(names) ^ set(npmg.SIGNATURES)


@pytest.mark.parametrize("bad, msg", [
    ("{not json", "JSON"),
    (json.dumps({"network": {}}), "encoding"),
    (json.dumps({"encoding": {"otype": "Composite", "nested": [{"otype": "Frequency"}]}, "network": {}}),
     "unsupported encoding"),
    (json.dumps({"encoding": {"otype": "OneBlob", "n_bins": 3}, "network": {}}), "power of two"),
    (json.dumps({"encoding": {"otype": "OneBlob"}, "network": {"n_neurons": 48}}), "n_neurons"),
    (json.dumps({"encoding": {"otype": "OneBlob"}, "network": {}, "precision": "int4"}), "precision"),
    (json.dumps({"encoding": {"otype": "OneBlob"}, "network": {}, "optimizer": {"otype": "Shampoo"}}), "Adam"),
])
def test_invalid_config_rejected(bad, msg):
    lib = npmg.load_library()
    h = ctypes.c_void_p()
    st = lib.npmg_create(bad.encode(), 1, ctypes.byref(h))
    assert st == npmg.NPMG_EINVAL
    assert msg.lower() in lib.npmg_last_error(None).decode().lower()
    assert not h.value


def test_null_arguments():
    lib = npmg.load_library()
    assert lib.npmg_create(None, 1, None) == npmg.NPMG_EINVAL
    assert lib.npmg_n_params(None, None, None) == npmg.NPMG_EINVAL
    assert lib.npmg_debug_math(99, None, None, 0, None, None) == npmg.NPMG_EINVAL
