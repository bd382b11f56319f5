"""The C-ABI library loads and exports every entry point include/wfx.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(wfx_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "wfx_create" in names and "wfx_step_api" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    from weldformfem_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from weldformfem_amd import _lib
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        return
    from weldformfem_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.wfx_create(ctypes.byref(h), 0)
    assert rc == -3
    assert b"device" in lib.wfx_last_error()
