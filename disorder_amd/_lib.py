"""ctypes binding of libdisorder_hip.so (the C-ABI declared in include/disorder.h).

The library is built in-tree (``make -C disorder_amd/csrc``, hipcc for
gfx950).  Loading fails loudly when it is missing: there is no
CPU fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdisorder_hip.so")

c_int = ctypes.c_int
c_int32 = ctypes.c_int32
c_void_p = ctypes.c_void_p
c_size_t = ctypes.c_size_t
c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_int32_p = ctypes.POINTER(ctypes.c_int32)

DIS_OK = 0
DIS_E_INVALID = -1
DIS_E_UNSUPPORTED = -2
DIS_E_NOMEM = -3
DIS_E_HIP = -4
DIS_E_NODEVICE = -5


class DisGeom(ctypes.Structure):
    _fields_ = [("n", c_int32 * 3), ("Nfull", c_int32 * 3), ("ncoils", c_int32), ("nstates", c_int32),
                ("nmotion", c_int32), ("nsamples", c_int32)]


# name -> (restype, argtypes); every symbol of include/disorder.h
SIGNATURES = {
    "dis_version": (c_int, []),
    "dis_status_string": (ctypes.c_char_p, [c_int]),
    "dis_ctx_create": (c_int, [c_int, ctypes.POINTER(c_void_p)]),
    "dis_ctx_destroy": (c_int, [c_void_p]),
    "dis_last_error": (ctypes.c_char_p, [c_void_p]),
    "dis_sync": (c_int, [c_void_p]),
    "dis_set_workspace_limit": (c_int, [c_void_p, c_size_t]),
    "dis_rigid_transform": (c_int, [c_void_p, c_int32_p, c_int32_p, c_double_p, c_int32, c_int32, c_int32, c_int32,
                                    c_void_p, c_void_p]),
    "dis_plan_create": (c_int, [c_void_p, ctypes.POINTER(DisGeom), c_void_p, c_int32_p, c_int32_p, c_int32_p,
                                c_float_p, ctypes.POINTER(c_void_p)]),
    "dis_plan_destroy": (c_int, [c_void_p]),
    "dis_plan_set_weights": (c_int, [c_void_p, c_float_p]),
    "dis_plan_set_fold": (c_int, [c_void_p, c_int32, c_int32]),
    "dis_plan_set_readout_roi": (c_int, [c_void_p, c_int32, c_int32]),
    "dis_plan_set_lm_filter": (c_int, [c_void_p, c_float_p, c_int32]),
    "dis_coil_gram_noalloc": (c_int, [c_void_p, c_void_p, ctypes.c_int64, c_int32, c_void_p]),
    "dis_stream_probe": (c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_int32, c_int32, c_int32,
                                 c_double_p]),
    "dis_encode": (c_int, [c_void_p, c_double_p, c_void_p, c_void_p]),
    "dis_decode": (c_int, [c_void_p, c_double_p, c_void_p, c_void_p]),
    "dis_encodedecode": (c_int, [c_void_p, c_double_p, c_void_p, c_void_p]),
    "dis_energy": (c_int, [c_void_p, c_double_p, c_void_p, c_void_p, c_float_p, c_double_p, c_double_p]),
    "dis_cg": (c_int, [c_void_p, c_double_p, c_void_p, c_float_p, c_float_p, ctypes.c_double, c_void_p, c_int32,
                       c_int32, c_int32, ctypes.c_double, c_double_p, c_int32_p, c_int32_p]),
    "dis_autotune": (c_int, [c_void_p, c_int32_p, c_int32_p, c_int32, c_int32_p, c_double_p]),
    "dis_shearlet_create": (c_int, [c_void_p, c_int32_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32_p, c_int32,
                                    ctypes.POINTER(c_void_p)]),
    "dis_shearlet_destroy": (c_int, [c_void_p]),
    "dis_shearlet_weights": (c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_double, ctypes.c_double]),
    "dis_shearlet_regularize": (c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_double, c_void_p]),
    "dis_shearlet_energy": (c_int, [c_void_p, c_void_p, c_void_p, c_double_p]),
    "dis_fidelity_trace": (c_int, [c_void_p, c_void_p, c_void_p, c_int32, c_float_p, c_void_p, c_double_p]),
    "dis_fidelity_trace_folded": (c_int, [c_void_p, c_int32_p, c_void_p, c_int32, c_int32, c_int32, c_float_p, c_void_p,
                                   c_double_p]),
    "dis_cg_cs": (c_int, [c_void_p, c_double_p, c_void_p, c_float_p, c_float_p, ctypes.c_double, c_void_p,
                          ctypes.c_double, c_void_p, c_int32, c_int32, c_int32, ctypes.c_double, c_int32_p]),
    "dis_autotune_clear": (c_int, [c_void_p]),
    "dis_set_coil_schedule": (c_int, [c_void_p, c_int32]),
    "dis_set_cache_budget": (c_int, [c_void_p, ctypes.c_int64]),
    "dis_set_pass_shape": (c_int, [c_void_p, c_int32_p, c_int32, c_int32, c_int32, c_int32]),
    "dis_cg_batched": (c_int, [c_void_p, c_int32, ctypes.POINTER(c_void_p), ctypes.POINTER(c_double_p),
                               ctypes.POINTER(c_void_p), ctypes.POINTER(c_float_p), ctypes.POINTER(c_void_p), c_int32,
                               c_int32, c_double_p, c_double_p]),
    "dis_rigid_transform_grad": (c_int, [c_void_p, c_int32_p, c_int32_p, c_double_p, c_int32, c_void_p, c_void_p,
                                         c_void_p]),
    "dis_lm_partials": (c_int, [c_void_p, c_double_p, c_void_p, c_void_p, c_double_p]),
    "dis_lm": (c_int, [c_void_p, c_double_p, c_double_p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_double, c_double_p,
                       c_void_p, c_void_p, c_double_p, c_int32_p]),
    "dis_coil_gram": (c_int, [c_void_p, c_void_p, ctypes.c_int64, c_int32, c_double_p]),
    "dis_coil_basis": (c_int, [c_double_p, c_int32, c_double_p, c_int32, c_int32_p, c_double_p, c_double_p]),
    "dis_coil_apply": (c_int, [c_void_p, c_double_p, c_int32, c_int32, c_void_p, ctypes.c_int64, c_void_p, c_int32]),
    "dis_compress_coils": (c_int, [c_void_p, c_void_p, ctypes.c_int64, c_int32, c_double_p, c_int32, c_void_p,
                                   ctypes.c_int64, c_void_p, c_void_p, c_int32, c_int32_p, c_double_p, c_double_p]),
    "dis_resample": (c_int, [c_void_p, c_int32_p, c_int32_p, ctypes.c_int64, c_void_p, c_void_p]),
    "dis_gather_samples": (c_int, [c_void_p, c_void_p, c_int32, ctypes.c_int64, c_int32_p, c_int32, ctypes.c_float,
                                   c_void_p]),
    "dis_coil_precond": (c_int, [c_void_p, c_void_p, ctypes.c_int64, c_int32, ctypes.c_float, c_float_p]),
    "dis_filter_rows": (c_int, [c_void_p, c_void_p, ctypes.c_int64, c_int32, c_int32, c_int32_p, c_int32_p, c_int32,
                                c_int32, ctypes.c_float]),
    "dis_mlp_param_count": (c_int, []),
    "dis_mlp_forward": (c_int, [c_void_p, c_float_p, c_float_p, c_int32, c_int32_p, c_int32, c_float_p]),
    "dis_mlp_backward": (c_int, [c_void_p, c_float_p, c_float_p, c_int32, c_int32_p, c_int32, c_float_p, c_float_p,
                                 c_float_p]),
    "dis_mlp_train_step": (c_int, [c_void_p, c_float_p, c_float_p, c_float_p, c_int32, ctypes.c_float, c_float_p,
                                   c_float_p, c_int32, c_int32_p, c_int32, c_float_p]),
    "dis_profile_enable": (c_int, [c_void_p, c_int32]),
    "dis_profile_read": (c_int, [c_void_p, c_int32, ctypes.POINTER(ctypes.c_int64), c_double_p, c_double_p]),
}

K_PASS, K_BCAST, K_ACCUM, K_COIL, K_FLIP, K_GRAM, K_GEMM, K_SHEAR = range(8)
K_NAMES = {K_PASS: "pass", K_BCAST: "bcast_pass", K_ACCUM: "accum_pass", K_COIL: "coil_stage", K_FLIP: "flip_pass",
           K_GRAM: "coil_gram", K_GEMM: "coil_apply", K_SHEAR: "shearlet_pass"}

_lib = None


class DisorderError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"disorder status {status}: {message}")
        self.status = status


def load():
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with make -C disorder_amd/csrc "
                          "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, ctx=None):
    if status != DIS_OK:
        lib = load()
        msg = lib.dis_last_error(ctx).decode() if ctx else lib.dis_status_string(status).decode()
        raise DisorderError(status, msg or lib.dis_status_string(status).decode())
