"""The C-ABI library loads and exports every entry point include/g4acc.h declares (no compute
calls without a GPU)."""
import ctypes
import os
import re

import pytest

from dca_amd import g4acc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "g4acc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(g4acc_[a-zA-Z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(g4acc.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = g4acc.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    # Cross-check against the dynamic symbol table itself.
    out = os.popen(f"nm -D --defined-only {g4acc.LIB_PATH}").read()
    for name in declared_functions():
        assert re.search(rf"\bT {name}\b", out), name


def test_library_is_built_for_gfx950():
    data = open(g4acc.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_null_handle_calls_fail_cleanly():
    lib = g4acc.load_library()
    assert lib.g4acc_reset(None) == g4acc.INVALID_ARG
    assert lib.g4acc_sync(None) == g4acc.INVALID_ARG
    assert lib.g4acc_g4_size(None) == 0
    assert lib.g4acc_destroy(None) == g4acc.OK
    assert lib.g4acc_create(None, None, None) == g4acc.INVALID_ARG


def test_create_rejects_bad_geometry_before_touching_a_device():
    from dca_amd import synth
    wl = synth.tiny_workload()
    lib = g4acc.load_library()
    gc = g4acc.Geometry_c()
    gc.n_bands = 0
    h = ctypes.c_void_p()
    opt = g4acc.Options_c()
    assert lib.g4acc_create(ctypes.byref(gc), ctypes.byref(opt), ctypes.byref(h)) == g4acc.INVALID_ARG
    assert b"n_bands" in lib.g4acc_last_error(None)
    # An unsupported channel is reported as such (tp_accumulator_cpu.hpp:663-664).
    with pytest.raises(g4acc.LogicError, match="four point type"):
        wl.geometry.channels = [7]
        g4acc.TpAccumulator(wl.geometry)
