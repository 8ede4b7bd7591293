"""ctypes wrapper of the FP64 CPU oracle (oracle/profit_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from libprofit_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
LIB_FAST = os.path.join(HERE, "_build", "liboracle_fast.so")

_lib = None
_fast = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def fast_lib():
    """The same oracle built at -O3 -march=native on this host (the CPU
    baseline); rebuilt on first use."""
    global _fast
    if _fast is None:
        subprocess.check_call(["make", "-s", "-C", HERE, "fast"])
        _fast = _bind(C.CDLL(LIB_FAST))
    return _fast


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _bind(C.CDLL(LIB))
    return _lib


def _bind(L):
    L.or_gammafn.restype = C.c_double
    L.or_gammafn.argtypes = [C.c_double]
    L.or_beta.restype = C.c_double
    L.or_beta.argtypes = [C.c_double, C.c_double]
    L.or_pgamma.restype = C.c_double
    L.or_pgamma.argtypes = [C.c_double, C.c_double]
    L.or_qgamma.restype = C.c_double
    L.or_qgamma.argtypes = [C.c_double, C.c_double]
    L.or_last_error.restype = C.c_char_p
    L.or_stat_refined.restype = C.c_uint64
    L.or_stat_subcells.restype = C.c_uint64
    L.or_component_eval_at.restype = C.c_double
    L.or_component_eval_at.argtypes = [C.c_void_p, C.c_uint, C.c_double, C.c_double]
    L.or_set_threads.argtypes = [C.c_int]
    L.or_set_integrator.argtypes = [C.c_int]
    L.or_quadpack_test.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_uint, C.c_void_p]
    L.or_quadpack.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.c_uint, C.c_void_p]
    return L


class OracleError(Exception):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def set_threads(n):
    lib().or_set_threads(int(n))


INTEGRATORS = {"quadpack": 0, "exact": 1}


def set_integrator(mode):
    """Luminosity integrals by libprofit's QUADPACK qags/qagiu at 1e-4
    ("quadpack", the default) or by the same algorithm to 1e-13 ("exact")."""
    lib().or_set_integrator(INTEGRATORS[mode])


def quadpack_test(fn, alpha, a, b, epsabs, epsrel, infinite=False, limit=1000):
    """GSL test-suite integrands f1, f11, f15, f455 through the oracle's
    qags (finite) / qagiu (infinite): (result, abserr, neval, status)."""
    out = (C.c_double * 4)()
    lib().or_quadpack_test(fn, int(infinite), alpha, a, b, epsabs, epsrel, limit, out)
    return out[0], out[1], int(out[2]), int(out[3])


_QFUNC = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


def quadpack(f, a, b=None, epsabs=1e-4, epsrel=1e-4, limit=100):
    """qags on [a, b] (b given) or qagiu on [a, inf) of a Python function:
    (result, abserr, neval, status)."""
    cb = _QFUNC(lambda x, _p: f(x))
    out = (C.c_double * 4)()
    lib().or_quadpack(C.cast(cb, C.c_void_p), None, int(b is None), a, 0.0 if b is None else b, epsabs, epsrel,
                      limit, out)
    return out[0], out[1], int(out[2]), int(out[3])


def gammafn(x):
    return lib().or_gammafn(x)


def beta(a, b):
    return lib().or_beta(a, b)


def pgamma(q, shape):
    return lib().or_pgamma(q, shape)


def qgamma(p, shape):
    return lib().or_qgamma(p, shape)


def output_dims(spec):
    w, h = C.c_uint(), C.c_uint()
    st = lib().or_output_dims(spec.ref, C.byref(w), C.byref(h))
    if st:
        raise OracleError(st, lib().or_last_error().decode())
    return w.value, h.value


def evaluate(spec, L=None):
    """Model::evaluate -> (image[h, w] float64, (offset_x, offset_y)).  `L`:
    another build of the oracle (fast_lib()) instead of the parity one."""
    L = L or lib()
    m = spec.m
    fs = max(1, m.finesampling)
    cap = 4 * (m.width * fs + (m.psf_width or 0) * 2 + 4) * (m.height * fs + (m.psf_height or 0) * 2 + 4) + 64
    cap = max(cap, 4 * max(m.width * fs, m.psf_width) ** 2 * 4)
    out = np.zeros(cap, dtype=np.float64)
    w, h, ox, oy = C.c_uint(), C.c_uint(), C.c_int(), C.c_int()
    st = L.or_evaluate(spec.ref, out.ctypes.data_as(C.c_void_p), C.c_size_t(cap), C.byref(w), C.byref(h),
                       C.byref(ox), C.byref(oy))
    if st:
        raise OracleError(st, L.or_last_error().decode())
    return out[: w.value * h.value].reshape(h.value, w.value).copy(), (ox.value, oy.value)


def pattern(mode):
    """Frozen refinement pattern: "record" (the next evaluations log their
    refinement pattern), "replay" (the next evaluations reuse the log in the
    same order) or "off"."""
    lib().or_pattern(C.c_int({"off": 0, "record": 1, "replay": 2}[mode]))


def pattern_ok():
    return lib().or_pattern_status() == 0


def stats():
    return lib().or_stat_refined(), lib().or_stat_subcells()


def convolve(src, krn, mask=None, crop=True, conv="brute"):
    src = np.ascontiguousarray(src, dtype=np.float64)
    krn = np.ascontiguousarray(krn, dtype=np.float64)
    t = abi.CONVOLVERS[conv] if isinstance(conv, str) else conv
    H, W = src.shape
    kh, kw = krn.shape
    mx = max(W, kw, H, kh)
    cap = 4 * (2 * mx) * (2 * mx) + 16
    out = np.zeros(cap, dtype=np.float64)
    mptr = None
    if mask is not None:
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        mptr = mask.ctypes.data_as(C.c_void_p)
    ow, oh, ox, oy = C.c_uint(), C.c_uint(), C.c_int(), C.c_int()
    st = lib().or_convolve(C.c_int(t), src.ctypes.data_as(C.c_void_p), C.c_uint(W), C.c_uint(H),
                           krn.ctypes.data_as(C.c_void_p), C.c_uint(kw), C.c_uint(kh), mptr, C.c_int(int(crop)),
                           out.ctypes.data_as(C.c_void_p), C.c_size_t(cap), C.byref(ow), C.byref(oh),
                           C.byref(ox), C.byref(oy))
    if st:
        raise OracleError(st, lib().or_last_error().decode())
    return out[: ow.value * oh.value].reshape(oh.value, ow.value).copy(), (ox.value, oy.value)


def component_setup(spec, c=0):
    out = (C.c_double * 16)()
    st = lib().or_component_setup(spec.ref, C.c_uint(c), out)
    if st:
        raise OracleError(st, lib().or_last_error().decode())
    keys = ["rscale", "ie", "acc", "rscale_switch", "rscale_max", "resolution", "bn", "flux_scale",
            "cos_ang", "sin_ang", "rescale_factor"]
    return {k: out[i] for i, k in enumerate(keys)}


def eval_at(spec, c, x, y):
    return lib().or_component_eval_at(C.cast(spec.ref, C.c_void_p), c, x, y)
