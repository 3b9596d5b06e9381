"""The C-ABI library loads and exports every entry point include/pfh.h declares
(no device calls: runs without a GPU)."""
import ctypes
import os
import re

import pfh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "pfh.h")).read()
    return sorted(set(re.findall(r"\b(pfh_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert set(declared_functions()) == set(pfh.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pfh.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_create_rejects_bad_device_without_crashing():
    lib = pfh.load()
    h = ctypes.c_void_p()
    assert lib.pfh_create(-1, ctypes.byref(h)) != 0
    assert not h.value
    assert lib.pfh_rasterize(None, None, 0, 0) == -1
