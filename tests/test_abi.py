"""CPU-side checks of the C ABI: the library loads and exports exactly what
include/dwg_hip.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

from dwg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "dwg_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dwg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_list():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(L, s)]
    assert not missing, missing


def test_binding_prototypes_load():
    L = _lib.load()
    assert L.dwg_ctx_error.restype is ctypes.c_char_p


def test_null_arguments_are_rejected_without_a_device():
    L = _lib.load()
    assert L.dwg_ctx_create(0, None, None) == _lib.DWG_ERR_ARG
    assert L.dwg_octree_info(None, None, None, None) == _lib.DWG_ERR_ARG
    assert L.dwg_ctx_error(None) == b"null context"
