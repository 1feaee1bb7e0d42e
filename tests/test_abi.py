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


def test_new_entry_points_reject_bad_arguments_without_a_device():
    import ctypes
    lib = _lib.load()
    assert lib.dis_mlp_param_count() == 256 * 13 + 256 + 256 * 256 + 256 + 6 * 256 + 6 + 6 * 6 * 5 + 6
    assert lib.dis_set_coil_schedule(None, 0) == _lib.DIS_E_INVALID
    assert lib.dis_set_cache_budget(None, 0) == _lib.DIS_E_INVALID
    assert lib.dis_plan_set_lm_filter(None, None, 0) == _lib.DIS_E_INVALID
    assert lib.dis_autotune(None, None, None, 1, None, None) == _lib.DIS_E_INVALID
    assert lib.dis_autotune_clear(None) == _lib.DIS_E_INVALID
    assert lib.dis_set_pass_shape(None, None, 0, 0, 0, 16) == _lib.DIS_E_INVALID
    assert lib.dis_cg_batched(None, 0, None, None, None, None, None, 1, 1, None, None) == _lib.DIS_E_INVALID
    assert lib.dis_coil_gram(None, None, 0, 0, None) == _lib.DIS_E_INVALID
    assert lib.dis_coil_basis(None, 4, None, 0, None, None, None) == _lib.DIS_E_INVALID
    assert lib.dis_resample(None, None, None, 1, None, None) == _lib.DIS_E_INVALID
    assert lib.dis_mlp_forward(None, None, None, 1, None, 1, None) == _lib.DIS_E_INVALID
    # the host eigensolver needs no device
    P = (ctypes.c_double * 8)(2.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0)
    perc = (ctypes.c_double * 1)(0.5)
    cc = (ctypes.c_int32 * 1)()
    D = (ctypes.c_double * 2)()
    assert lib.dis_coil_basis(P, 2, perc, 1, cc, D, None) == _lib.DIS_OK
    assert cc[0] == 1 and D[0] == 2.0 and D[1] == 1.0
