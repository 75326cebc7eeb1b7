"""The C ABI library builds for gfx950, loads without a GPU and exports
every function include/heddle.h declares (no compute calls here)."""
import ctypes
import os
import re

from heddlemd_amd import _lib


def _declared():
    with open(_lib.HEADER_PATH) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hd_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_bound():
    declared = _declared()
    assert len(declared) > 30
    missing = [n for n in declared if n not in _lib.SIGNATURES]
    assert missing == [], f"declared in heddle.h but not bound: {missing}"


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert missing == []


def test_version_and_gfx950_code_object():
    lib = _lib.load()
    assert b"gfx950" in lib.hd_version()
    with open(_lib.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"gfx950" in blob


def test_context_creation_fails_cleanly_without_gpu():
    # on a machine without a GPU the library reports an error, it does not
    # fall back to the CPU
    import torch

    if torch.cuda.is_available():
        return
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    rc = lib.hd_ctx_create(0, ctypes.byref(ctx))
    assert rc != 0 and not ctx.value
