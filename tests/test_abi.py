"""The C-ABI library loads and exports every entry point include/qgpu.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from qsource_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "qgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qgpu_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(engine.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(engine.SYMBOLS) == names


def test_error_path_without_device_or_input():
    lib = engine.load_library()
    h = ctypes.c_void_p()
    arr = (ctypes.c_char_p * 1)(b"/nonexistent/input.inp")
    rc = lib.qgpu_create(arr, 1, 0, ctypes.byref(h))
    assert rc != 0
    assert lib.qgpu_last_error().decode() != ""
