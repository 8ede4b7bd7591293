"""Loader of the HIP library libprofit_amd/_lib/libprofit_hip.so.

There is no fallback: if the library is missing, loading raises.  The
environment variable LIBPROFIT_HIP_LIBRARY names another build of the same
library (for instance one made with different compile-time options).
"""
import ctypes as C
import os
import re

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LIBPROFIT_HIP_LIBRARY") or os.path.join(HERE, "_lib", "libprofit_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "profit_hip.h")

_lib = None


class LibraryMissing(RuntimeError):
    pass


def exported_symbols():
    """Names of the functions declared in include/profit_hip.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pk_[a-z0-9_]+)\s*\(", txt)) - {"pk_integration_func_t"})


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            "libprofit_hip.so is not built (%s); run `python -m libprofit_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    dbl, u32, i32 = C.c_double, C.c_uint32, C.c_int32
    sig = {
        "pk_version": (C.c_char_p, []),
        "pk_init": (C.c_int, []),
        "pk_finish": (None, []),
        "pk_last_error": (C.c_char_p, []),
        "pk_device_count": (C.c_int, [P(C.c_int)]),
        "pk_ctx_create": (C.c_int, [C.c_int, C.c_int, P(C.c_void_p)]),
        "pk_ctx_destroy": (C.c_int, [C.c_void_p]),
        "pk_ctx_precision": (C.c_int, [C.c_void_p]),
        "pk_component_defaults": (C.c_int, [C.c_int, P(abi.Component)]),
        "pk_model_defaults": (C.c_int, [u32, u32, P(abi.Model)]),
        "pk_model_output_dims": (C.c_int, [P(abi.Model), P(u32), P(u32)]),
        "pk_evaluate": (C.c_int, [C.c_void_p, P(abi.Model), C.c_void_p, C.c_size_t, P(abi.Result)]),
        "pk_convolve": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, u32, u32, C.c_void_p, u32, u32, C.c_void_p,
                                  C.c_int, C.c_void_p, C.c_size_t, P(u32), P(u32), P(i32), P(i32)]),
        "pk_convolver_padding": (C.c_int, [C.c_int, u32, u32, u32, u32, P(i32), P(i32), P(i32), P(i32)]),
        "pk_component_constants": (C.c_int, [P(abi.Model), u32, P(abi.Constants)]),
        "pk_batch_create": (C.c_int, [C.c_void_p, P(abi.Model), u32, u32, P(C.c_void_p)]),
        "pk_batch_set_components": (C.c_int, [C.c_void_p, P(abi.Component), u32]),
        "pk_batch_evaluate": (C.c_int, [C.c_void_p]),
        "pk_batch_sync": (C.c_int, [C.c_void_p, P(abi.BatchInfo)]),
        "pk_batch_image_dims": (C.c_int, [C.c_void_p, u32, P(u32), P(u32)]),
        "pk_batch_work_counts": (C.c_int, [C.c_void_p, P(abi.WorkCounts)]),
        "pk_batch_get_image": (C.c_int, [C.c_void_p, u32, C.c_void_p, C.c_size_t]),
        "pk_batch_device_images": (C.c_int, [C.c_void_p, P(C.c_void_p), P(C.c_size_t)]),
        "pk_batch_image_offset": (C.c_int, [C.c_void_p, u32, P(C.c_uint64)]),
        "pk_batch_image_origin": (C.c_int, [C.c_void_p, u32, P(C.c_int32), P(C.c_int32)]),
        "pk_evaluate_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
        "pk_evaluate_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
        "pk_evaluate_batch": (C.c_int, [C.c_void_p, C.c_void_p, u32, C.c_void_p, C.c_void_p, C.c_void_p]),
        "pk_chi2_grad": (C.c_int, [C.c_void_p, C.c_void_p, u32, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                   C.c_void_p]),
        "pk_batch_destroy": (C.c_int, [C.c_void_p]),
        "pk_batch_set_data": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
        "pk_batch_evaluate_grad": (C.c_int, [C.c_void_p]),
        "pk_batch_get_chi2_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
        "pk_batch_autotune": (C.c_int, [C.c_void_p]),
        "pk_batch_get_tuning": (C.c_int, [C.c_void_p, P(abi.Tuning)]),
        "pk_batch_set_tuning": (C.c_int, [C.c_void_p, P(abi.Tuning)]),
        "pk_clear_cache": (C.c_int, []),
        "pk_init": (C.c_int, []),
        "pk_finish": (None, []),
        "pk_qgamma": (dbl, [dbl, dbl]),
        "pk_pgamma": (dbl, [dbl, dbl]),
        "pk_gammafn": (dbl, [dbl]),
        "pk_beta": (dbl, [dbl, dbl]),
        "pk_integrate_qags": (dbl, [C.c_void_p, dbl, dbl, C.c_void_p]),
        "pk_integrate_qagi": (dbl, [C.c_void_p, dbl, C.c_void_p]),
        "pk_set_integrator": (C.c_int, [C.c_int]),
        "pk_get_integrator": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


INTEGRAND = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


class ProfitError(Exception):
    """Raised for a non-zero status; `code` is the PK_* status."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class InvalidParameter(ProfitError, ValueError):
    pass


class UnknownParameter(InvalidParameter):
    pass


def check(status):
    if status != abi.PK_OK:
        msg = load().pk_last_error().decode()
        if status == abi.PK_EUNKNOWN:
            raise UnknownParameter(status, msg)
        if status == abi.PK_EINVAL:
            raise InvalidParameter(status, msg)
        raise ProfitError(status, msg)
