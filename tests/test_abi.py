"""CPU checks of the boundary: the C-ABI library loads and exports every
symbol include/disorder.h declares, and the ctypes binding covers them all
(no compute calls: there is no GPU here)."""
import ctypes
import os
import re

from disorder_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "disorder.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dis_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 15
    for name in syms:
        assert hasattr(lib, name), name


def test_binding_covers_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_status_strings_and_version():
    lib = _lib.load()
    assert lib.dis_version() >= 100
    assert lib.dis_status_string(0) == b"ok"
    assert lib.dis_status_string(_lib.DIS_E_UNSUPPORTED) == b"unsupported"


def test_null_arguments_are_rejected_without_a_device():
    lib = _lib.load()
    assert lib.dis_ctx_create(0, None) == _lib.DIS_E_INVALID
    assert lib.dis_ctx_destroy(None) == _lib.DIS_E_INVALID
    assert lib.dis_encode(None, None, None, None) == _lib.DIS_E_INVALID
    assert lib.dis_plan_destroy(None) == _lib.DIS_E_INVALID
    assert lib.dis_last_error(None) == b"null context"
