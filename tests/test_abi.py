"""C-ABI checks that need no GPU: the library loads and exports every entry
point that include/sphadj.h declares, with the signatures the binding uses."""
import ctypes
import os
import re

from sphadjoint_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_entry_points():
    src = open(os.path.join(ROOT, "include", "sphadj.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(sph_\w+)\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = header_entry_points()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_lib.SIGNATURES) == header_entry_points()


def test_last_error_callable_without_gpu():
    lib = _lib.load()
    assert isinstance(lib.sph_last_error(), bytes)


def test_ctx_create_fails_cleanly_without_gpu():
    import torch
    if torch.cuda.is_available():
        return
    h = ctypes.c_void_p()
    rc = _lib.load().sph_ctx_create(0, ctypes.byref(h))
    assert rc != _lib.SPH_OK
    assert _lib.load().sph_last_error()
