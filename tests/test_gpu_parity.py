"""Device parity: pk_evaluate / pk_convolve against the FP64 oracle.

FP64 mode must agree to <= 1e-12 per pixel and FP32 mode to <= 1e-5 on pixels
above 1e-6 of the peak (relative error), at sizes the oracle finishes in
seconds.
"""
import re
import sys

import numpy as np
import pytest

from libprofit_amd import abi
from libprofit_amd._lib import InvalidParameter
from libprofit_amd.model import evaluate_spec
from parity_util import TOL_FP32, TOL_FP64, sersic100_adjusted, sersic100_fixed, galaxy400_params, comp, moffat_psf, rel_errors, sersic

pytestmark = pytest.mark.gpu


def announce(name):
    print("CASE", name, flush=True)
    sys.stdout.flush()


def run_both(spec, ctx, oracle):
    got, goff, res = evaluate_spec(ctx, spec)
    ref, roff = oracle.evaluate(spec)
    assert got.shape == ref.shape
    assert goff == roff
    return got, ref, res


def test_sersic100_adjusted_fp64(ctx64, oracle):
    announce("sersic100 adjusted fp64")
    got, ref, res = run_both(sersic100_adjusted(), ctx64, oracle)
    pix, flux = rel_errors(got, ref)
    nref, nsub = oracle.stats()
    assert res.n_refined_pixels == nref
    assert res.n_subcells == nsub
    assert pix <= TOL_FP64, pix
    assert flux <= TOL_FP64, flux


def test_sersic100_fixed_fp64(ctx64, oracle):
    announce("sersic100 fixed fp64")
    got, ref, _ = run_both(sersic100_fixed(), ctx64, oracle)
    pix, flux = rel_errors(got, ref)
    assert pix <= TOL_FP64, pix


def test_sersic100_fp32(ctx32, oracle):
    for name, spec in (("sersic100_adjusted", sersic100_adjusted()), ("sersic100_fixed", sersic100_fixed())):
        announce(name + "_fp32")
        got, ref, _ = run_both(spec, ctx32, oracle)
        pix, flux = rel_errors(got, ref)
        assert pix <= TOL_FP32, (name, pix)
        assert flux <= TOL_FP32, (name, flux)


PROFILES = [
    ("sersic", dict(re=5.0, nser=2.5)),
    ("sersic", dict(re=3.0, nser=0.5)),
    ("sersic", dict(re=4.0, nser=1.0)),
    ("sersic", dict(re=6.0, nser=8.0)),
    ("moffat", dict(fwhm=4.0, con=3.0)),
    ("ferrer", dict(rout=12.0, a=1.5, b=0.5)),
    ("king", dict(rc=3.0, rt=20.0, a=2.0)),
    ("brokenexp", dict(h1=6.0, h2=3.0, rb=8.0, a=0.5)),
    ("coresersic", dict(re=8.0, rb=2.0, nser=3.0, a=2.0, b=0.8)),
]


@pytest.mark.parametrize("box", [0.0, 0.4, -0.3])
@pytest.mark.parametrize("name,params", PROFILES)
def test_profiles_fp64(ctx64, oracle, name, params, box):
    announce("profile_fp64 %s %s box=%s" % (name, params, box))
    spec = abi.ModelSpec(64, 64, [comp(name, xcen=31.3, ycen=32.6, mag=16, ang=25, axrat=0.7, box=box, **params)])
    got, ref, _ = run_both(spec, ctx64, oracle)
    pix, flux = rel_errors(got, ref)
    assert pix <= TOL_FP64, pix


# FP32 evaluation regimes: truncations (king, ferrer with small a), sharp breaks,
# b = 0 cores, compact Sersic profiles with deep refinement
PROFILES_FP32 = PROFILES + [
    ("ferrer", dict(rout=10.0, a=0.05, b=1.0)),
    ("ferrer", dict(rout=9.0, a=3.0, b=1.9)),
    ("king", dict(rc=1.0, rt=10.0, a=0.5)),
    ("brokenexp", dict(h1=4.0, h2=8.0, rb=10.0, a=5.0)),
    ("coresersic", dict(re=6.0, rb=1.0, nser=2.0, a=3.0, b=0.0)),
    ("sersic", dict(re=1.2, nser=4.0)),
]


@pytest.mark.parametrize("box", [0.0, 0.4, -0.3])
@pytest.mark.parametrize("name,params", PROFILES_FP32)
def test_profiles_fp32(ctx32, oracle, name, params, box):
    announce("profile_fp32 %s %s box=%s" % (name, params, box))
    spec = abi.ModelSpec(64, 64, [comp(name, xcen=31.3, ycen=32.6, mag=16, ang=25, axrat=0.7, box=box, **params)])
    got, ref, _ = run_both(spec, ctx32, oracle)
    pix, flux = rel_errors(got, ref)
    assert pix <= TOL_FP32, pix


def test_sky_and_mask(ctx64, oracle):
    announce("sky_mask")
    mask = np.ones((2, 2), dtype=np.uint8)
    mask[1, 1] = 0
    spec = abi.ModelSpec(2, 2, [comp("sky", bg=5.0)], mask=mask)
    got, _, _ = evaluate_spec(ctx64, spec)
    # tests/test_sky.h:48-61
    assert list(got.ravel()) == [5, 5, 5, 0]


def test_psf_profile_exact(ctx64):
    announce("psf_profile")
    # tests/test_psf.h:36-50
    spec = abi.ModelSpec(10, 10, [comp("psf", xcen=2.0, ycen=2.0, mag=0.0)], psf=np.ones((2, 2)))
    got, _, _ = evaluate_spec(ctx64, spec)
    exp = np.zeros((10, 10))
    exp[1:3, 1:3] = 0.25
    assert np.allclose(got, exp, rtol=1e-9, atol=0)


def test_mask_finesampling(ctx64, oracle):
    announce("mask_finesampling")
    rng = np.random.Generator(np.random.PCG64(3))
    mask = (rng.uniform(size=(40, 50)) > 0.3).astype(np.uint8)
    for fs in (1, 2):
        for rf in (True, False):
            spec = abi.ModelSpec(50, 40, [sersic(xcen=20.2, ycen=22.7, re=6, nser=3, axrat=0.5, ang=70)], mask=mask,
                                 finesampling=fs, return_finesampled=rf)
            got, ref, _ = run_both(spec, ctx64, oracle)
            pix, _ = rel_errors(got, ref)
            assert pix <= TOL_FP64, (fs, rf, pix)


@pytest.mark.parametrize("conv", ["brute", "brute-old", "fft"])
def test_convolved_model_fp64(ctx64, oracle, conv):
    announce("convolved %s" % conv)
    psf, _ = oracle.evaluate(moffat_psf(15, 3.0, 2.5))
    spec = abi.ModelSpec(96, 80, [sersic(xcen=47.1, ycen=41.4, re=7, nser=2, axrat=0.6, ang=10, convolve=True),
                                  comp("sky", bg=1e-9)], psf=psf, convolver=conv)
    got, ref, _ = run_both(spec, ctx64, oracle)
    pix, flux = rel_errors(got, ref)
    if conv != "fft":
        assert pix <= TOL_FP64, pix
    else:
        # FFT rounding error is relative to the peak (~log2(N) eps max|image|),
        # not to each pixel: bound it against the peak, and per pixel at 1e-8
        # above 1e-6 of the peak (the oracle's own FFT carries the same error)
        assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))
        assert pix <= 1e-8, pix


@pytest.mark.parametrize("seed", [0, 1])
def test_galaxy400_fp64_fp32(ctx64, ctx32, oracle, seed):
    announce("galaxy400 seed %d" % seed)
    psf, _ = oracle.evaluate(moffat_psf(25, 3.0, 2.5))
    p = galaxy400_params(seed)
    spec = abi.ModelSpec(400, 400, [sersic(convolve=True, **p)], psf=psf)
    ref, _ = oracle.evaluate(spec)
    g64, _, _ = evaluate_spec(ctx64, spec)
    g32, _, _ = evaluate_spec(ctx32, spec)
    pix64, flux64 = rel_errors(g64, ref)
    pix32, flux32 = rel_errors(g32, ref)
    print("galaxy400 seed", seed, "fp64", pix64, flux64, "fp32", pix32, flux32)
    assert pix64 <= TOL_FP64
    assert pix32 <= TOL_FP32
    assert flux32 <= TOL_FP32


def test_convolver_api(ctx64, oracle):
    announce("convolver_api")
    rng = np.random.Generator(np.random.PCG64(7))
    src = rng.uniform(size=(37, 41))
    krn = rng.uniform(size=(9, 6))
    mask = (rng.uniform(size=src.shape) > 0.2).astype(np.uint8)
    from libprofit_amd.model import Convolver
    for t in (abi.BRUTE_OLD, abi.BRUTE, abi.FFT):
        for crop in (True, False):
            for mk in (None, mask):
                conv = Convolver(t, context=ctx64)
                got, goff = conv.convolve(src, krn, mask=mk, crop=crop)
                ref, roff = oracle.convolve(src, krn, mask=mk, crop=crop, conv=t)
                assert got.shape == ref.shape and goff == roff, (t, crop)
                assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), (t, crop)


@pytest.mark.parametrize("sw,sh,kw,kh", [(40, 40, 9, 9), (41, 40, 8, 9), (40, 41, 9, 8), (41, 41, 8, 8),
                                         (100, 60, 25, 25), (37, 53, 13, 7)])
def test_fft_convolver_matches_direct(ctx64, oracle, sw, sh, kw, kh):
    """Device FFT convolution (FP64, 2-D radix-4/2) against the oracle's direct
    BRUTE_OLD convolution, all even/odd image and kernel combinations
    (tests/test_fft.h:45-139 compares FFT with BRUTE at 1 %): the FFT's error
    is relative to the peak, ~1e-15 of it."""
    announce("fft vs direct %dx%d %dx%d" % (sw, sh, kw, kh))
    from libprofit_amd.model import Convolver
    rng = np.random.Generator(np.random.PCG64(sw * 1000 + kw))
    src = rng.uniform(size=(sh, sw)) ** 4
    krn = rng.uniform(size=(kh, kw))
    for crop in (True, False):
        try:
            ref, roff = oracle.convolve(src, krn, crop=crop, conv="fft")
        except oracle.OracleError as e:
            # libprofit's crop leaves its 2 max(W, kw)-row buffer for tall images
            with pytest.raises(InvalidParameter, match=re.escape(str(e))):
                Convolver(abi.FFT, context=ctx64).convolve(src, krn, crop=crop)
            continue
        got, goff = Convolver(abi.FFT, context=ctx64).convolve(src, krn, crop=crop)
        assert got.shape == ref.shape and goff == roff
        exact, _ = oracle.convolve(src, krn, crop=True, conv="brute-old")
        if crop:
            assert np.max(np.abs(got - exact)) <= 1e-13 * np.max(np.abs(exact))
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_galaxy400_fft_model(ctx64, ctx32, oracle):
    """The 400x400 Sersic + 25x25 Moffat PSF model with the FFT convolver,
    against the oracle's BRUTE model: per pixel above 1e-6 of the peak."""
    announce("galaxy400 fft")
    psf, _ = oracle.evaluate(moffat_psf(25, 3.0, 2.5))
    p = galaxy400_params(2)
    ref, _ = oracle.evaluate(abi.ModelSpec(400, 400, [sersic(convolve=True, **p)], psf=psf, convolver="brute"))
    for ctx, tol in ((ctx64, 1e-8), (ctx32, TOL_FP32)):
        got, _, _ = evaluate_spec(ctx, abi.ModelSpec(400, 400, [sersic(convolve=True, **p)], psf=psf, convolver="fft"))
        pix, flux = rel_errors(got, ref)
        print("galaxy400 fft", ctx.precision, pix, flux)
        assert pix <= tol and flux <= tol


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_evaluate_into_device_memory(ctx64, ctx32, oracle, prec):
    """pk_evaluate_device writes the same image as pk_evaluate, into device
    memory (from hipMalloc of the HIP runtime the library uses)."""
    announce("evaluate_device %s" % prec)
    import ctypes as C
    from libprofit_amd._lib import check, load
    ctx = ctx64 if prec == "fp64" else ctx32
    psf, _ = oracle.evaluate(moffat_psf(15, 3.0, 2.5))
    specs = [abi.ModelSpec(96, 80, [sersic(xcen=47.1, ycen=41.4, re=7, nser=2, axrat=0.6, ang=10, convolve=True),
                                    comp("sky", bg=1e-9)], psf=psf),
             abi.ModelSpec(60, 50, [sersic(xcen=30.1, ycen=24.4, re=5, nser=1.5, convolve=True)], psf=psf,
                           convolver="fft", crop=False),
             abi.ModelSpec(30, 20, [sersic(xcen=15, ycen=10)], dry_run=True)]
    L = load()
    hip = C.CDLL("libamdhip64.so.7")  # the runtime the library is linked against (already loaded)
    for sp in specs:
        host, off, _ = evaluate_spec(ctx, sp)
        cap = host.size + 7
        dev = C.c_void_p()
        assert hip.hipMalloc(C.byref(dev), C.c_size_t(8 * cap)) == 0
        try:
            fill = np.full(cap, -1.0)
            assert hip.hipMemcpy(dev, fill.ctypes.data_as(C.c_void_p), C.c_size_t(8 * cap), 1) == 0  # H2D
            res = abi.Result()
            check(L.pk_evaluate_device(ctx.handle, sp.ref, dev, cap, C.byref(res)))
            got = np.zeros(cap)
            assert hip.hipMemcpy(got.ctypes.data_as(C.c_void_p), dev, C.c_size_t(8 * cap), 2) == 0  # D2H
        finally:
            hip.hipFree(dev)
        assert (res.width, res.height) == (host.shape[1], host.shape[0])
        assert (res.offset_x, res.offset_y) == off
        assert np.array_equal(got[: host.size].reshape(host.shape), host)
        assert np.all(got[host.size:] == -1.0)  # nothing written past the image
