"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

from rtlflow_amd.build import build_runtime

ROOT = Path(__file__).resolve().parent.parent


def declared(header: str):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(rf_[a-z0-9_]+)\s*\(", text, re.M)
    return sorted(set(names))


def test_runtime_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(build_runtime()))
    names = declared("rf_verilated.h")
    assert "rf_create" in names and "rf_submit_queue" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_runtime_exports_cost_model_symbols():
    lib = ctypes.CDLL(str(build_runtime()))
    names = declared("rf_mlp.h")
    assert "rf_mlp_train_step" in names and "rf_mlp_forward" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    lib.rf_mlp_param_count.restype = ctypes.c_uint32
    assert lib.rf_mlp_param_count(6) == 64 * 6 + 64 + 64 * 64 + 64 + 64 + 1


def test_error_path_without_design():
    lib = ctypes.CDLL(str(build_runtime()))
    lib.rf_create.restype = ctypes.c_int
    lib.rf_last_error.restype = ctypes.c_char_p
    out = ctypes.c_void_p()
    rc = lib.rf_create(None, 16, None, ctypes.byref(out))
    assert rc == -1
    assert b"null design" in lib.rf_last_error()


def test_runtime_exports_gatenet_symbols():
    lib = ctypes.CDLL(str(build_runtime()))
    names = declared("rf_gatenet.h")
    assert "rf_gn_create" in names and "rf_gn_run" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_gatenet_rejects_a_bad_program_without_a_gpu():
    import numpy as np
    lib = ctypes.CDLL(str(build_runtime()))
    lib.rf_gn_create.restype = ctypes.c_int
    lib.rf_gn_create.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32,
                                 ctypes.POINTER(ctypes.c_void_p)]
    lib.rf_last_error.restype = ctypes.c_char_p
    out = ctypes.c_void_p()
    bad = np.zeros(64, np.uint32)
    rc = lib.rf_gn_create(bad.ctypes.data, 64, 16, ctypes.byref(out))
    assert rc == -1 and b"gate-net program" in lib.rf_last_error()
