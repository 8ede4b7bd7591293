"""pk_evaluate_device_noalloc: the allocation-free single evaluation against
pk_evaluate_device (the entry point it varies) on the same models, two
calls in a row with the second model smaller, in both precisions; its size
limit; and the context's own entry points afterwards."""
import ctypes as C

import numpy as np
import pytest

from libprofit_amd import abi, workloads
from libprofit_amd._lib import ProfitError, check, load
from parity_util import TOL_FP32, TOL_FP64, check_image

pytestmark = pytest.mark.gpu


def _hip():
    return C.CDLL("libamdhip64.so.7")  # the runtime the library is linked against (already loaded)


def _device_eval(fn, ctx, spec, cap):
    hip = _hip()
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), C.c_size_t(8 * cap)) == 0
    try:
        fill = np.full(cap, -1.0)
        assert hip.hipMemcpy(dev, fill.ctypes.data_as(C.c_void_p), C.c_size_t(8 * cap), 1) == 0
        res = abi.Result()
        st = fn(ctx.handle, spec.ref, dev, C.c_size_t(cap), C.byref(res))
        got = np.zeros(cap)
        assert hip.hipMemcpy(got.ctypes.data_as(C.c_void_p), dev, C.c_size_t(8 * cap), 2) == 0
    finally:
        hip.hipFree(dev)
    return st, got, res


def _models():
    psf = workloads.moffat_psf_image(15, 3.0, 2.5)
    big = abi.ModelSpec(180, 150, [
        abi.make_component("sersic", xcen=90.3, ycen=70.6, mag=16, re=12, nser=3.5, axrat=0.6, ang=25, convolve=True),
        abi.make_component("coresersic", xcen=88.1, ycen=74.2, mag=17, re=15, rb=3, nser=3, a=2, b=1.3,
                           axrat=0.8, ang=60, convolve=True),
        abi.make_component("sky", bg=0.5)], psf=psf, magzero=30.0)
    small = abi.ModelSpec(60, 50, [
        abi.make_component("sersic", xcen=30.1, ycen=24.4, mag=15, re=5, nser=1.5, convolve=True),
        abi.make_component("sky", bg=0.2)], psf=psf, convolver="fft", crop=False, magzero=30.0)
    return big, small


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_evaluate_device(ctx64, ctx32, oracle, prec):
    print("CASE noalloc", prec, flush=True)
    L = load()
    ctx = ctx64 if prec == "fp64" else ctx32
    tol = TOL_FP64 if prec == "fp64" else TOL_FP32
    for rep in range(2):
        for spec in _models():  # larger model first, then a smaller one
            ref, off = oracle.evaluate(spec)
            cap = ref.size + 9
            st0, want, res0 = _device_eval(L.pk_evaluate_device, ctx, spec, cap)
            check(st0)
            st1, got, res1 = _device_eval(L.pk_evaluate_device_noalloc, ctx, spec, cap)
            check(st1)
            assert (res1.width, res1.height) == (res0.width, res0.height) == (ref.shape[1], ref.shape[0])
            assert (res1.offset_x, res1.offset_y) == (res0.offset_x, res0.offset_y) == off
            assert np.array_equal(got, want), (prec, rep)  # the same kernels on the same inputs
            assert np.all(got[ref.size:] == -1.0)  # nothing written past the image
            img = got[: ref.size].reshape(ref.shape)
            if prec == "fp64" and spec.m.convolver == abi.FFT:
                # FFT rounding is relative to the peak (as for pk_evaluate)
                assert np.max(np.abs(img - ref)) <= 1e-14 * np.max(np.abs(ref))
                check_image(img, spec, ref, oracle, 1e-8, (prec, rep))
            else:
                check_image(img, spec, ref, oracle, tol, (prec, rep))


def test_noalloc_dry_run_and_size_limit(ctx64):
    print("CASE noalloc limits", flush=True)
    L = load()
    dry = abi.ModelSpec(30, 20, [abi.make_component("sersic", xcen=15, ycen=10)], dry_run=True)
    st, got, res = _device_eval(L.pk_evaluate_device_noalloc, ctx64, dry, 610)
    check(st)
    assert (res.width, res.height) == (30, 20)
    assert not np.any(got[:600]) and np.all(got[600:] == -1.0)
    # 3000 x 3000 drawing pixels in FP64 need more than the fixed workspace
    huge = abi.ModelSpec(3000, 3000, [abi.make_component("sersic", xcen=1500, ycen=1500, re=30, nser=1)])
    st, _, _ = _device_eval(L.pk_evaluate_device_noalloc, ctx64, huge, 3000 * 3000)
    assert st == abi.PK_ERANGE
    assert b"allocation-free" in L.pk_last_error()
    # the context's own entry points still work after the rejected call
    spec = _models()[1]
    st0, want, _ = _device_eval(L.pk_evaluate_device, ctx64, spec, 120 * 120 + 5)
    check(st0)
    st1, got, _ = _device_eval(L.pk_evaluate_device_noalloc, ctx64, spec, 120 * 120 + 5)
    check(st1)
    assert np.array_equal(got, want)
