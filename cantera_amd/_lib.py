"""ctypes binding of libcantera_amd.so (the C ABI in include/cantera_gpu.h).

The library is built in-tree by ``python -m cantera_amd.build``.  Loading fails
loudly if it is missing: there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcantera_amd.so")
CG_FP64 = 0
CG_MIXED = 1
CG_FP32 = 2

_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

# (name, restype, argtypes) for every function declared in include/cantera_gpu.h
SIGNATURES = [
    ("cg_version", C.c_int32, []),
    ("cg_last_error", C.c_int32, [C.c_char_p, C.c_int32]),
    ("cg_device_count", C.c_int32, []),
    ("cg_dev_alloc", C.c_void_p, [C.c_int64]),
    ("cg_dev_free", C.c_int32, [C.c_void_p]),
    ("cg_memcpy_h2d", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("cg_memcpy_d2h", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("cg_device_synchronize", C.c_int32, []),
    ("cg_mech_load", C.c_int32, [C.c_char_p]),
    ("cg_mech_free", C.c_int32, [C.c_int32]),
    ("cg_mech_nspecies", C.c_int32, [C.c_int32]),
    ("cg_mech_nreactions", C.c_int32, [C.c_int32]),
    ("cg_mech_nelements", C.c_int32, [C.c_int32]),
    ("cg_mech_species_index", C.c_int32, [C.c_int32, C.c_char_p]),
    ("cg_mech_species_name", C.c_int32, [C.c_int32, C.c_int32, C.c_char_p, C.c_int32]),
    ("cg_mech_molecular_weights", C.c_int32, [C.c_int32, C.c_int32, _dp]),
    ("cg_rates_batch", C.c_int32, [C.c_int32, C.c_int64, _dp, _dp, _dp, _dp, C.c_int32]),
    ("cg_rates_batch_device", C.c_int32, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_int32, C.c_void_p]),
    ("cg_jac_batch", C.c_int32, [C.c_int32, C.c_int64, _dp, _dp, _dp, C.c_void_p, C.c_int32]),
    ("cg_jac_batch_device", C.c_int32, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_int32, C.c_void_p]),
    ("cg_jac_last_kernel_ms", C.c_double, [C.c_int32]),
    ("cg_equilibrate_batch", C.c_int32, [C.c_int32, C.c_int64, C.c_char_p, _dp, _dp, _dp, _dp, _dp, _i32p]),
    ("cg_equilibrate_batch_device", C.c_int32, [C.c_int32, C.c_int64, C.c_char_p, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                                C.c_void_p]),
    ("cg_reactors_create", C.c_int32, [C.c_int32, C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int32]),
    ("cg_reactors_create_const_volume", C.c_int32,
     [C.c_int32, C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int32]),
    ("cg_reactors_free", C.c_int32, [C.c_int32]),
    ("cg_reactors_advance", C.c_int32, [C.c_int32, _dp]),
    ("cg_reactors_advance_device", C.c_int32, [C.c_int32, C.c_void_p]),
    ("cg_reactors_advance_graph", C.c_int32, [C.c_int32, C.c_int32, C.c_void_p]),
    ("cg_reactors_autotune", C.c_int32, [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
    ("cg_reactors_get_state", C.c_int32, [C.c_int32, _dp]),
    ("cg_reactors_stats", C.c_int32, [C.c_int32, _i64p]),
    ("cg_reactors_status", C.c_int32, [C.c_int32, _i32p]),
    ("cg_reactors_create_mole", C.c_int32, [C.c_int32, C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int32]),
    ("cg_reactors_linear_stats", C.c_int32, [C.c_int32, _i64p]),
    ("cg_reactors_add_sensitivity_reaction", C.c_int32, [C.c_int32, C.c_int32]),
    ("cg_reactors_set_sensitivity_tolerances", C.c_int32, [C.c_int32, C.c_double, C.c_double]),
    ("cg_reactors_get_sensitivities", C.c_int32, [C.c_int32, C.c_int32, _dp]),
    ("cg_reactors_get_control", C.c_int32, [C.c_int32, _dp]),
    ("cg_reactors_set_options", C.c_int32, [C.c_int32, C.c_int32, C.c_int32]),
    ("cg_reactors_set_schedule", C.c_int32, [C.c_int32, C.c_int32]),
    ("cg_reactors_set_matrix_storage", C.c_int32, [C.c_int32, C.c_int32]),
    ("cg_reactors_last_kernel_ms", C.c_double, [C.c_int32]),
    ("cg_rates_last_kernel_ms", C.c_double, [C.c_int32]),
    ("cg_memcpy_d2d", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("cg_surrogate_create", C.c_int32, [C.c_int32, C.c_double, C.c_int32, C.c_int32, C.c_uint64]),
    ("cg_surrogate_free", C.c_int32, [C.c_int32]),
    ("cg_surrogate_nparams", C.c_int64, [C.c_int32]),
    ("cg_surrogate_fit_normalization", C.c_int32, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("cg_surrogate_train_epoch", C.c_int32, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                             C.c_double, C.POINTER(C.c_double)]),
    ("cg_surrogate_loss_grad", C.c_int32, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.POINTER(C.c_double), C.POINTER(C.c_float)]),
    ("cg_surrogate_predict", C.c_int32, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("cg_surrogate_predict_noalloc", C.c_int32, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p]),
    ("cg_surrogate_get_params", C.c_int32, [C.c_int32, C.POINTER(C.c_float)]),
    ("cg_surrogate_set_params", C.c_int32, [C.c_int32, C.POINTER(C.c_float)]),
    ("cg_surrogate_get_normalization", C.c_int32, [C.c_int32, C.POINTER(C.c_double)]),
    ("cg_surrogate_set_normalization", C.c_int32, [C.c_int32, C.POINTER(C.c_double)]),
    ("cg_reactors_advance_surrogate", C.c_int32, [C.c_int32, C.c_int32, _dp, C.c_double, C.c_double, C.c_double,
                                                  C.POINTER(C.c_int64)]),
    ("cg_reactors_advance_surrogate_device", C.c_int32, [C.c_int32, C.c_int32, C.c_void_p, C.c_double, C.c_double,
                                                         C.c_double]),
    ("cg_reactors_surrogate_counts", C.c_int32, [C.c_int32, C.POINTER(C.c_int64), C.c_int32]),
    ("cg_reactors_get_state_device", C.c_int32, [C.c_int32, C.c_void_p]),
    ("cg_reactors_net_production_rates_device", C.c_int32, [C.c_int32, C.c_void_p]),
]


class CanteraGpuError(RuntimeError):
    pass


def lib():
    """Load libcantera_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CanteraGpuError(f"{LIB_PATH} is not built; run `python -m cantera_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(2048)
    lib().cg_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(ret, what="call"):
    if ret is None or (isinstance(ret, int) and ret < 0):
        raise CanteraGpuError(f"{what} failed: {last_error()}")
    return ret
