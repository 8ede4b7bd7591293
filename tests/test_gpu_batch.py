"""Batched, device-resident evaluation (pk_batch_*): every image of a
heterogeneous batch against the FP64 oracle, graph replays against eager
launches, and parameter updates between evaluations."""
import numpy as np
import pytest

from libprofit_amd import abi, workloads
from libprofit_amd.model import Batch
from parity_util import TOL_FP32, TOL_FP64, check_image, rel_errors

pytestmark = pytest.mark.gpu


def small_batch(n=20, seed=3):
    return workloads.galaxy_batch(n, seed, smin=24, smax=96)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_matches_oracle(ctx64, ctx32, oracle, prec):
    print("CASE batch", prec, flush=True)
    specs = small_batch()
    ctx = ctx64 if prec == "fp64" else ctx32
    b = Batch(ctx, specs, graph=False)
    b.evaluate()
    b.sync()
    tol = TOL_FP64 if prec == "fp64" else TOL_FP32
    for i, s in enumerate(specs):
        ref, _ = oracle.evaluate(s)
        check_image(b.image(i), s, ref, oracle, tol, i)
    b.close()


def test_graph_replays_match_eager(ctx32):
    print("CASE graph replay", flush=True)
    specs = small_batch(12, 5)
    eager = Batch(ctx32, specs, graph=False)
    eager.evaluate()
    eager.sync()
    ref = [eager.image(i) for i in range(len(specs))]
    g = Batch(ctx32, specs, graph=True)
    for rep in range(3):
        g.evaluate()
        g.sync()
        for i in range(len(specs)):
            assert np.array_equal(g.image(i), ref[i]), (rep, i)
        # eager work of the program between replays
        eager.evaluate()
        eager.sync()
    g.close()
    eager.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_set_components(ctx64, ctx32, oracle, prec):
    """New parameters for every component, the sky levels included (an FP32
    job whose unconvolved part is only sky takes it as one value computed on
    the device at each evaluation): the graph replays the new images."""
    print("CASE set_components", prec, flush=True)
    specs = small_batch(6, 9)
    b = Batch(ctx64 if prec == "fp64" else ctx32, specs, graph=True)
    b.evaluate()
    b.sync()
    # move every component and brighten it, as one step of a fit would
    new_specs, comps = [], []
    for s in specs:
        cs = []
        for k in range(s.m.n_components):
            c = abi.Component.from_buffer_copy(s.m.components[k])
            if c.kind != abi.PROFILE_KINDS["sky"]:
                c.xcen += 0.37
                c.ycen -= 0.21
                c.mag -= 0.5
            else:
                c.p[0] = 1.5 * c.p[0] + 0.25  # bg
            cs.append(c)
        comps += cs
        new_specs.append(abi.ModelSpec(s.m.width, s.m.height, cs, psf=s._psf, magzero=s.m.magzero))
    b.set_components(comps)
    b.evaluate()
    b.sync()
    for i, s in enumerate(new_specs):
        ref, _ = oracle.evaluate(s)
        check_image(b.image(i), s, ref, oracle, TOL_FP64 if prec == "fp64" else TOL_FP32, i)
    b.close()


DEFAULT_TUNING = {"conv_tiles": 0, "refine_occupancy": 100, "handover_div": 4, "claim_pixels": 256,
                  "claim_cells": 256, "refine_passes": 5, "claim_spread": 0}


def images(b, n):
    return [b.image(i) for i in range(n)]


def assert_same_to_fp32_rounding(got, ref):
    # the matrix-core convolution forms sum the PSF products in another order
    for i, (g, r) in enumerate(zip(got, ref)):
        peak = np.max(np.abs(r))
        assert np.max(np.abs(g - r)) <= 2e-6 * peak, i


def test_tuning_settings_keep_the_images(ctx32):
    """The refinement settings only move work between wavefronts: the
    fixed-point sums make the images bit-identical.  The matrix-core
    convolution forms (banded tiles, strips) change the FP32 rounding only."""
    print("CASE tuning settings", flush=True)
    from libprofit_amd._lib import InvalidParameter
    specs = small_batch(16, 11)
    b = Batch(ctx32, specs, graph=True)
    t0 = b.tuning.as_dict()
    assert {k: t0[k] for k in DEFAULT_TUNING} == DEFAULT_TUNING
    b.evaluate()
    b.sync()
    ref = images(b, len(specs))
    for setting in ({"handover_div": 2, "refine_occupancy": 50, "claim_pixels": 16, "claim_cells": 4,
                     "refine_passes": 4, "claim_spread": 4},
                    {"handover_div": 16, "refine_occupancy": 75, "claim_pixels": 128, "claim_cells": 64,
                     "refine_passes": 7, "claim_spread": 64},
                    {"handover_div": 1, "refine_occupancy": 1, "claim_spread": 4096, "refine_passes": 6}):
        print("CASE tuning", setting, flush=True)
        b.tuning = {**DEFAULT_TUNING, **setting}
        for _rep in range(2):
            b.evaluate()
            b.sync()
            for i, r in enumerate(ref):
                assert np.array_equal(b.image(i), r), (setting, i)
    for tiles in (1, 2):
        b.tuning = {**DEFAULT_TUNING, "conv_tiles": tiles}
        b.evaluate()
        b.sync()
        assert_same_to_fp32_rounding(images(b, len(specs)), ref)
    for bad in ({"handover_div": 0}, {"refine_occupancy": 101}, {"conv_tiles": 3}, {"claim_pixels": 0},
                {"refine_passes": 3}, {"refine_passes": 8},
                {"claim_spread": 3}, {"claim_spread": 8192}):
        with pytest.raises(InvalidParameter):
            b.tuning = {**DEFAULT_TUNING, **bad}
    b.close()


def test_autotune_and_its_cache(ctx32, tmp_path, monkeypatch):
    print("CASE autotune", flush=True)
    from libprofit_amd import model as M
    home = tmp_path / "profit_home"
    monkeypatch.setenv("PROFIT_HOME", str(home))
    M.clear_cache()
    specs = small_batch(16, 13)
    plain = Batch(ctx32, specs, graph=False)
    plain.evaluate()
    plain.sync()
    ref = images(plain, len(specs))
    plain.close()

    b = Batch(ctx32, specs, graph=True, autotune=True)
    b.evaluate()  # tunes first, then captures and replays the graph
    b.sync()
    t = b.tuning.as_dict()
    print("CASE tuned", t, flush=True)
    assert t["ms_evaluate"] > 0
    assert t["conv_tiles"] in (0, 1, 2) and 1 <= t["refine_occupancy"] <= 100
    for _rep in range(2):
        got = images(b, len(specs))
        if t["conv_tiles"] == 0:
            for i, r in enumerate(ref):
                assert np.array_equal(got[i], r), i
        else:
            assert_same_to_fp32_rounding(got, ref)
        b.evaluate()
        b.sync()

    # the next batch of the same workload takes the cached settings
    b2 = Batch(ctx32, specs, graph=False, autotune=True)
    b2.evaluate()
    b2.sync()
    assert b2.tuning.as_dict() == t
    b2.close()

    # finish() writes the cache, init() reads it back after clear_cache()
    M.finish()
    f = home / "hip_tuning"
    text = f.read_text()
    lines = text.splitlines()
    assert len(lines) == 1 and "fp32" in lines[0] and "models16" in lines[0], text
    M.clear_cache()
    assert not f.exists()
    f.write_text(text)
    M.init()
    b3 = Batch(ctx32, specs, graph=False, autotune=True)
    b3.evaluate()
    b3.sync()
    t3 = b3.tuning.as_dict()
    assert abs(t3.pop("ms_evaluate") - t["ms_evaluate"]) <= 1e-5  # the file keeps 6 decimals
    assert t3 == {k: v for k, v in t.items() if k != "ms_evaluate"}
    b3.close()
    b.close()
    M.clear_cache()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_mixed_composition(ctx64, ctx32, oracle, prec):
    """Jobs whose convolution writes the returned image (no downsampling) next
    to jobs composed afterwards (SUM downsampling, no PSF) in one batch."""
    print("CASE mixed composition", prec, flush=True)
    psf = workloads.moffat_psf_image(9, 2.5, 3.0)
    specs = []
    for k, (fs, rf, with_psf) in enumerate([(1, True, True), (2, False, True), (2, True, True), (1, True, False),
                                            (3, False, True), (1, True, True)]):
        s = 30 + 7 * k
        comps = [abi.make_component("sersic", xcen=0.45 * s, ycen=0.55 * s, mag=15.0, re=3.0 + k, nser=2.5,
                                    axrat=0.7, ang=20.0 * k, convolve=with_psf),
                 abi.make_component("sky", bg=0.5)]
        specs.append(abi.ModelSpec(s, s, comps, psf=psf if with_psf else None, finesampling=fs,
                                   return_finesampled=rf, magzero=30.0))
    ctx = ctx64 if prec == "fp64" else ctx32
    tol = TOL_FP64 if prec == "fp64" else TOL_FP32
    refs = [oracle.evaluate(s)[0] for s in specs]
    for tiles in ((0,) if prec == "fp64" else (0, 1, 2)):
        b = Batch(ctx, specs, graph=True)
        b.tuning = {**DEFAULT_TUNING, "conv_tiles": tiles}
        for _rep in range(2):
            b.evaluate()
            b.sync()
            for i, (s, ref) in enumerate(zip(specs, refs)):
                got = b.image(i)
                assert got.shape == ref.shape, (tiles, i, got.shape, ref.shape)
                check_image(got, s, ref, oracle, tol, (tiles, i))
        b.close()


# every PSF-size class of the batch convolution (largest side <= 17, 21, ...,
# 41, 48 and above), odd sides from 15 to 41, and even, non-square, tiny and
# wide PSFs (the run-time column paths; for the strip form: every count of
# PSF column steps 1..12 and row blocks 1..3, tail rows 1..3 after one or two
# blocks, and the banded tiles above 48)
PSF_SHAPES = [(15, 15), (17, 17), (19, 19), (21, 21), (23, 23), (25, 25), (27, 27), (29, 29), (31, 31), (33, 33),
              (35, 35), (37, 37), (39, 39), (41, 41), (51, 51), (3, 3), (8, 8), (5, 9), (12, 7), (1, 5), (30, 30),
              (16, 24), (21, 31), (41, 15), (45, 47), (48, 48), (47, 33), (20, 18), (9, 34), (48, 35), (2, 17)]


@pytest.mark.parametrize("tiles", [0, 1, 2])
def test_batch_psf_shapes(ctx32, oracle, tiles):
    """The convolution forms (VALU tiles, banded MFMA tiles, MFMA strips)
    against the oracle in FP32 on stamps whose galaxy dominates its pixels,
    over every PSF-size class and the odd, even and non-square PSF shapes."""
    print("CASE psf shapes, conv_tiles", tiles, flush=True)
    specs = []
    for k, (pw, ph) in enumerate(PSF_SHAPES):
        s = 48 + 5 * k
        psf = workloads.moffat_psf_image(max(pw, ph), 2.0 + 0.1 * k, 2.5 + 0.1 * k)[:ph, :pw]
        comps = [abi.make_component("sersic", xcen=0.5 * s + 0.3, ycen=0.5 * s - 0.2, mag=16.0, re=3.0 + 0.2 * k,
                                    nser=1.5 + 0.1 * k, axrat=0.8, ang=15.0 * k, convolve=True),
                 abi.make_component("sky", bg=0.2)]
        specs.append(abi.ModelSpec(s, s, comps, psf=psf, magzero=30.0))
    b = Batch(ctx32, specs, graph=False)
    b.tuning = {**DEFAULT_TUNING, "conv_tiles": tiles}
    b.evaluate()
    b.sync()
    for i, s in enumerate(specs):
        ref, _ = oracle.evaluate(s)
        got = b.image(i)
        assert got.shape == ref.shape, (i, got.shape, ref.shape)
        check_image(got, s, ref, oracle, TOL_FP32, PSF_SHAPES[i])
    b.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("profile", workloads.RADIAL)
def test_batch_each_profile(ctx64, ctx32, oracle, prec, profile):
    """Galaxies made of one radial profile type only (1-3 components of it
    with their typical parameter ranges), so the device set-up of that
    profile (normalisation by closed form or quadrature, adjusted
    sub-pixel parameters) and its evaluation are compared visibly."""
    print("CASE one profile type", profile, prec, flush=True)
    specs = workloads.galaxy_batch(10, 7, smin=24, smax=160, profiles=(profile,))
    ctx = ctx64 if prec == "fp64" else ctx32
    tol = TOL_FP64 if prec == "fp64" else TOL_FP32
    for tiles in ((0,) if prec == "fp64" else (0, 1, 2)):
        b = Batch(ctx, specs, graph=False)
        b.tuning = {**DEFAULT_TUNING, "conv_tiles": tiles}
        b.evaluate()
        b.sync()
        for i, s in enumerate(specs):
            ref, _ = oracle.evaluate(s)
            check_image(b.image(i), s, ref, oracle, tol, (profile, tiles, i))
        b.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_masks(ctx64, ctx32, oracle, prec):
    """Calculation masks in a batch (Model::evaluate's mask handling,
    src/model.cpp:230-316): adjusted masks needing PSF padding or not,
    finesampled masks, masks in drawing dimensions, masks without a PSF."""
    print("CASE batch masks", prec, flush=True)
    rng = np.random.Generator(np.random.PCG64(21))
    psf = workloads.moffat_psf_image(9, 2.5, 3.0)

    def model(s, mask, psf_=psf, fs=1, adjust=True, rf=True):
        comps = [abi.make_component("sersic", xcen=0.48 * s, ycen=0.53 * s, mag=15.0, re=5.0, nser=2.0, axrat=0.6,
                                    ang=35.0, convolve=psf_ is not None),
                 abi.make_component("sky", bg=0.3)]
        return abi.ModelSpec(s, s, comps, psf=psf_, mask=mask, finesampling=fs, adjust_mask=adjust,
                             return_finesampled=rf, magzero=30.0)

    specs = []
    s = 50
    specs.append(model(s, (rng.uniform(size=(s, s)) > 0.3).astype(np.uint8)))          # bbox at the edges
    inner = np.zeros((s, s), np.uint8)
    inner[15:35, 12:38] = 1
    specs.append(model(s, inner))                                                      # no PSF padding needed
    specs.append(model(s, (rng.uniform(size=(s, s)) > 0.5).astype(np.uint8), fs=2))    # finesampled
    specs.append(model(s, (rng.uniform(size=(s, s)) > 0.5).astype(np.uint8), fs=2, rf=False))
    specs.append(model(s, (rng.uniform(size=(s + 8, s + 8)) > 0.4).astype(np.uint8), adjust=False))  # drawing dims
    specs.append(model(s, (rng.uniform(size=(s, s)) > 0.2).astype(np.uint8), psf_=None))  # no convolution
    specs.append(model(s, None))                                                       # no mask
    ctx = ctx64 if prec == "fp64" else ctx32
    tol = TOL_FP64 if prec == "fp64" else TOL_FP32
    refs = [oracle.evaluate(sp)[0] for sp in specs]
    # every convolution form of the precision (FP32: VALU tiles, banded MFMA
    # tiles, MFMA strips), each masking the pixels it stores
    for tiles in ((0,) if prec == "fp64" else (0, 1, 2)):
        b = Batch(ctx, specs, graph=True)
        b.tuning = {**DEFAULT_TUNING, "conv_tiles": tiles}
        for _rep in range(2):
            b.evaluate()
            b.sync()
            for i, (sp, ref) in enumerate(zip(specs, refs)):
                got = b.image(i)
                assert got.shape == ref.shape, (tiles, i, got.shape, ref.shape)
                assert np.array_equal(got == 0, ref == 0), (tiles, i)  # masked pixels are exactly 0 in both
                check_image(got, sp, ref, oracle, tol, (tiles, i))
        b.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_psf_profiles(ctx64, ctx32, oracle, prec):
    """PSF profiles in a batch (PsfProfile::evaluate, src/psf.cpp:58-145):
    convolved and not, several per model (their additions overlap), a model
    made of psf profiles only, finesampled models, and new positions through
    set_components (whose footprints change size at the image edges)."""
    print("CASE batch psf profiles", prec, flush=True)
    psf = workloads.moffat_psf_image(9, 2.5, 3.0)
    psf_even = workloads.moffat_psf_image(8, 2.0, 3.0)

    def comp(kind, convolve=True, **kw):
        return abi.make_component(kind, convolve=convolve, **kw)

    def spec(s, comps, psf_=psf, fs=1):
        return abi.ModelSpec(s, s, comps, psf=psf_, finesampling=fs, magzero=30.0)

    specs = [
        spec(40, [comp("sersic", xcen=20.3, ycen=19.6, mag=15.0, re=4.0, nser=1.5, axrat=0.7, ang=20.0),
                  comp("psf", xcen=12.4, ycen=25.1, mag=16.0),
                  comp("psf", xcen=14.0, ycen=24.0, mag=15.5, convolve=False),
                  comp("sky", bg=2e-10, convolve=False)]),
        spec(33, [comp("psf", xcen=16.5, ycen=16.5, mag=14.0),
                  comp("psf", xcen=17.2, ycen=15.9, mag=14.5),
                  comp("psf", xcen=1.0, ycen=31.5, mag=15.0, convolve=False)], psf_=psf_even),
        spec(36, [comp("sersic", xcen=18.0, ycen=18.0, mag=14.0, re=3.0, nser=3.0, axrat=0.5, ang=70.0),
                  comp("psf", xcen=9.7, ycen=8.2, mag=15.0)], fs=2),
        spec(30, [comp("sersic", xcen=15.0, ycen=15.2, mag=15.0, re=5.0, nser=1.0, axrat=0.9, ang=0.0)]),
    ]
    ctx = ctx64 if prec == "fp64" else ctx32
    tol = TOL_FP64 if prec == "fp64" else TOL_FP32
    b = Batch(ctx, specs, graph=True)

    def check(sps):
        b.evaluate()
        b.sync()
        for i, sp in enumerate(sps):
            ref, _ = oracle.evaluate(sp)
            got = b.image(i)
            assert got.shape == ref.shape, (i, got.shape, ref.shape)
            check_image(got, sp, ref, oracle, tol, i)

    check(specs)
    check(specs)
    new_specs, comps = [], []
    for sp in specs:
        cs = []
        for k in range(sp.m.n_components):
            c = abi.Component.from_buffer_copy(sp.m.components[k])
            if c.kind == abi.PROFILE_KINDS["psf"]:
                c.xcen = c.xcen * 0.2 + 0.3   # towards the corner: smaller footprints
                c.ycen += 0.45
                c.mag -= 0.3
            cs.append(c)
        comps += cs
        new_specs.append(abi.ModelSpec(sp.m.width, sp.m.height, cs, psf=sp._psf, finesampling=sp.m.finesampling,
                                       magzero=sp.m.magzero))
    b.set_components(comps)
    check(new_specs)
    b.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_fft_convolver(ctx64, ctx32, oracle, prec):
    """Models convolved by the FFT convolver in a batch, next to BRUTE and
    unconvolved ones: cropped and uncropped results (with the offset
    Model::evaluate reports, src/model.cpp:286-318), finesampled and masked
    models, a psf profile, an even PSF. FFT rounding is relative to the peak
    (~log2(N) eps max|image|), so FP64 is bounded against the peak and per
    pixel at 1e-8 above 1e-6 of the peak, as for a single evaluation."""
    print("CASE batch fft convolver", prec, flush=True)
    rng = np.random.Generator(np.random.PCG64(5))
    psf, _ = oracle.evaluate(abi.ModelSpec(15, 15, [abi.make_component("moffat", xcen=7.5, ycen=7.5, fwhm=3.0,
                                                                       con=2.5, mag=0.0)]))
    psf_even = workloads.moffat_psf_image(10, 2.5, 3.0)

    def sersic(s, **kw):
        p = dict(xcen=0.47 * s, ycen=0.52 * s, mag=15.0, re=0.1 * s, nser=2.0, axrat=0.6, ang=35.0, convolve=True)
        p.update(kw)
        return abi.make_component("sersic", **p)

    sky = abi.make_component("sky", bg=0.5)
    mz = 30.0
    specs = [
        abi.ModelSpec(60, 50, [sersic(55), sky], psf=psf, convolver="fft", magzero=mz),
        abi.ModelSpec(60, 50, [sersic(55), sky], psf=psf, convolver="fft", crop=False, magzero=mz),
        abi.ModelSpec(40, 40, [sersic(40, nser=1.0), sky], psf=psf, convolver="brute", magzero=mz),
        abi.ModelSpec(36, 36, [sersic(36, nser=3.0)], psf=psf, convolver="fft", finesampling=2,
                      return_finesampled=False, magzero=mz),
        abi.ModelSpec(48, 48, [sersic(48), sky], psf=psf, convolver="fft",
                      mask=(rng.uniform(size=(48, 48)) > 0.3).astype(np.uint8), magzero=mz),
        abi.ModelSpec(44, 44, [sersic(44, convolve=False), abi.make_component("psf", xcen=20.2, ycen=23.6, mag=15.5,
                                                                             convolve=True)],
                      psf=psf_even, convolver="fft", magzero=mz),
        abi.ModelSpec(32, 32, [sersic(32, convolve=False), sky], magzero=mz),
    ]
    ctx = ctx64 if prec == "fp64" else ctx32
    b = Batch(ctx, specs, graph=True)
    for _rep in range(2):
        b.evaluate()
        b.sync()
        for i, sp in enumerate(specs):
            ref, roff = oracle.evaluate(sp)
            got = b.image(i)
            assert got.shape == ref.shape, (i, got.shape, ref.shape)
            assert b.origin(i) == tuple(roff), (i, b.origin(i), roff)
            pix, flux = rel_errors(got, ref)
            if prec == "fp32":
                check_image(got, sp, ref, oracle, TOL_FP32, i)
            elif sp.m.convolver == abi.FFT and any(sp.m.components[k].convolve for k in range(sp.m.n_components)):
                assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref)), i
                assert pix <= 1e-8, (i, pix)
            else:
                check_image(got, sp, ref, oracle, TOL_FP64, i)
    b.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_dry_run(ctx64, ctx32, oracle, prec):
    """A dry-run model in a batch returns zeros of the drawing dimensions
    (Model::evaluate, src/model.cpp:263-267), as a single evaluation does, and
    leaves the other models' images alone."""
    print("CASE batch dry run", prec, flush=True)
    from libprofit_amd.model import evaluate_spec
    psf = workloads.moffat_psf_image(9, 2.5, 3.0)
    sers = abi.make_component("sersic", xcen=20.2, ycen=19.7, mag=15.0, re=4.0, nser=2.0, axrat=0.7, ang=30.0,
                              convolve=True)
    specs = [abi.ModelSpec(40, 40, [sers], psf=psf, magzero=30.0),
             abi.ModelSpec(40, 36, [sers], psf=psf, dry_run=True, magzero=30.0),
             abi.ModelSpec(30, 30, [abi.make_component("sky", bg=0.5)], dry_run=True, finesampling=2),
             abi.ModelSpec(40, 40, [sers], psf=psf, magzero=30.0)]
    ctx = ctx64 if prec == "fp64" else ctx32
    tol = TOL_FP64 if prec == "fp64" else TOL_FP32
    b = Batch(ctx, specs, graph=True)
    b.evaluate()
    b.sync()
    for i, sp in enumerate(specs):
        got = b.image(i)
        if sp.m.dry_run:
            single, _, _ = evaluate_spec(ctx, sp)
            assert got.shape == single.shape, (i, got.shape, single.shape)
            assert not np.any(got), i
        else:
            ref, _ = oracle.evaluate(sp)
            check_image(got, sp, ref, oracle, tol, i)
    b.close()


def test_one_shot_entry_points(ctx32, ctx64, oracle):
    """pk_evaluate_batch and pk_chi2_grad give what a pk_batch of the same
    models gives (images bit-identical; chi^2 and gradient to the FP64 sums'
    rounding), including an uncropped FFT model's offset."""
    print("CASE one-shot entry points", flush=True)
    from libprofit_amd.model import chi2_grad_once, evaluate_batch
    psf = workloads.moffat_psf_image(9, 2.5, 3.0)
    specs = small_batch(5, 17) + [abi.ModelSpec(40, 40, [abi.make_component(
        "sersic", xcen=20.2, ycen=19.7, mag=15.0, re=4.0, nser=2.0, convolve=True)], psf=psf, convolver="fft",
        crop=False, magzero=30.0)]
    for ctx in (ctx64, ctx32):
        b = Batch(ctx, specs, graph=False)
        b.evaluate()
        b.sync()
        got = evaluate_batch(ctx, specs)
        for i, (img, off) in enumerate(got):
            assert np.array_equal(img, b.image(i)), i
            assert off == b.origin(i), i
        b.close()
    # gradients: bulge + disk + sky fits
    fits, _true = workloads.fit_batch(3, seed=2)
    rng = np.random.Generator(np.random.PCG64(4))
    datas = []
    for sp in fits:
        ref, _ = oracle.evaluate(sp)
        datas.append(ref * (1 + 0.01 * rng.standard_normal(ref.shape)))
    sig = [np.full(d.shape, 0.01 * np.max(np.abs(d))) for d in datas]
    gb = Batch(ctx32, fits, graph=False, grad=True)
    gb.set_data(datas, sig)
    gb.evaluate_grad()
    c_ref, g_ref = gb.chi2_grad()
    gb.close()
    c1, g1 = chi2_grad_once(ctx32, fits, datas, sig)
    assert np.allclose(c1, c_ref, rtol=1e-12, atol=0)
    scale = np.max(np.abs(g_ref), axis=1, keepdims=True)
    assert np.all(np.abs(g1 - g_ref) <= 1e-11 * np.maximum(scale, 1e-300))


# coresersic components whose cusp (f ~ r^-b, b ~ 1.4) sits inside a pixel
# refined to 10^4 sub-cells: the FP32 refinement once carried cell sizes in
# FP32, which moved the sub-cell centres ~1e-8 px off the sub-pixel grid and
# the cusp pixel by up to 5e-4 (drawn from the 4096-galaxy workload)
CUSPS = [
    (283, dict(xcen=115.67246823524668, ycen=161.5464776775312, mag=18.380238625301047, ang=58.6569794054999,
               axrat=0.6097819655056684, re=18.046735457838764, rb=4.722559233079145, nser=4.000359166456292,
               a=1.8206748251676907, b=1.4844564321178517)),
    (73, dict(xcen=34.04079217214631, ycen=34.34578195114746, mag=19.760283438898483, ang=87.90286770678021,
              axrat=0.3904136910368339, re=14.166685941225545, rb=3.948601214478852, nser=3.475679002910674,
              a=1.1146656669163968, b=1.356464158287992)),
    (111, dict(xcen=54.744727068819024, ycen=44.46922342158413, mag=19.4157668075907, ang=28.039734483398696,
               axrat=0.9124329192959093, re=13.26279550549425, rb=4.865262260358035, nser=4.459245400184846,
               a=2.203062723288399, b=1.4102444686436317)),
]


def test_coresersic_cusp_pixels_fp32(ctx32, oracle):
    """The pixels holding a coresersic cusp, at the adjusted sub-pixel
    settings and at fixed resolutions 10 and 12 with 2 and 3 levels, in the
    FP32 batch and the single FP32 evaluation, against the oracle."""
    print("CASE coresersic cusps fp32", flush=True)
    from libprofit_amd.model import evaluate_spec
    specs = []
    for size, p in CUSPS:
        specs.append(abi.ModelSpec(size, size, [abi.make_component("coresersic", **p)], magzero=30.0))
        for res, mr in ((10, 2), (12, 3)):
            specs.append(abi.ModelSpec(size, size, [abi.make_component(
                "coresersic", adjust=False, resolution=res, max_recursions=mr, rscale_switch=1.0, rscale_max=100.0,
                **p)], magzero=30.0))
    b = Batch(ctx32, specs, graph=False)
    b.evaluate()
    b.sync()
    for i, s in enumerate(specs):
        ref, _ = oracle.evaluate(s)
        check_image(b.image(i), s, ref, oracle, TOL_FP32, ("batch", i))
        single, _, _ = evaluate_spec(ctx32, s)
        check_image(single, s, ref, oracle, TOL_FP32, ("single", i))
    b.close()
