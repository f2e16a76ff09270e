"""The C-ABI library loads on a host without a GPU and exports every function include/omni_gpu.h declares."""
import ctypes
import os
import re

from omnioperator_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "omni_gpu.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(omni_gpu_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_bound_entry_points():
    names = declared_functions()
    assert len(names) >= 20
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_load():
    L = capi.lib()
    assert L.omni_gpu_last_error() is not None
