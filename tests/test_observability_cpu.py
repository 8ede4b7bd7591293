"""The batched-path comparisons can fail: mistakes planted in the oracle (on
the CPU) move the images of the GPU tests' batches far beyond the FP32
tolerance.  Each planted mistake stands for a kernel that drops its work: the
galaxies (profile evaluation), the PSF convolution, or the sub-pixel
refinement (rough evaluation at pixel centres)."""
import numpy as np

from libprofit_amd import abi, workloads
from parity_util import TOL_FP32, galaxy_share, rel_errors, without_sky


def planted(spec, what):
    m = spec.m
    comps = []
    for k in range(m.n_components):
        c = abi.Component.from_buffer_copy(m.components[k])
        if what == "no convolution":
            c.convolve = 0
        elif what == "no refinement" and c.kind not in (abi.PROFILE_KINDS["sky"], abi.PROFILE_KINDS["psf"]):
            c.rough = 1
        comps.append(c)
    if what == "no galaxy":
        return without_sky(spec)
    return abi.ModelSpec(m.width, m.height, comps, psf=spec._psf, mask=spec._mask, magzero=m.magzero,
                         finesampling=m.finesampling, convolver=m.convolver)


def batches():
    yield "small batch", workloads.galaxy_batch(20, 3, smin=24, smax=96)
    for p in workloads.RADIAL:
        yield p, workloads.galaxy_batch(10, 7, smin=24, smax=160, profiles=(p,))
    specs = []
    for k, (pw, ph) in enumerate([(15, 15), (25, 25), (31, 31), (41, 41), (51, 51), (8, 8), (21, 31)]):
        s = 48 + 5 * k
        psf = workloads.moffat_psf_image(max(pw, ph), 2.0 + 0.1 * k, 2.5 + 0.1 * k)[:ph, :pw]
        specs.append(abi.ModelSpec(s, s, [abi.make_component("sersic", xcen=0.5 * s + 0.3, ycen=0.5 * s - 0.2,
                                                             mag=16.0, re=3.0 + 0.2 * k, nser=1.5 + 0.1 * k,
                                                             axrat=0.8, ang=15.0 * k, convolve=True),
                                          abi.make_component("sky", bg=0.2)], psf=psf, magzero=30.0))
    yield "psf shapes", specs


def test_planted_mistakes_fail_the_fp32_comparison(oracle):
    for name, specs in batches():
        refs = [oracle.evaluate(s)[0] for s in specs]
        for s, ref in zip(specs, refs):
            assert galaxy_share(s, ref, oracle) >= 1e3 * TOL_FP32, name
        for what in ("no galaxy", "no convolution", "no refinement"):
            worst = 0.0
            for s, ref in zip(specs, refs):
                bad, _ = oracle.evaluate(planted(s, what))
                worst = max(worst, rel_errors(bad, ref)[0])
            # the comparison of that batch would fail by far
            assert worst > 100 * TOL_FP32, (name, what, worst)


def test_sky_dominated_batch_hides_the_galaxy(oracle):
    """The reason for magzero 30: at magzero 0 the same galaxies are below
    one FP32 ulp of their sky, and dropping them stays within 1e-5."""
    specs = workloads.galaxy_batch(6, 3, smin=24, smax=96, magzero=0.0)
    for s in specs:
        ref, _ = oracle.evaluate(s)
        bad, _ = oracle.evaluate(without_sky(s))
        sky_only = ref - bad
        assert rel_errors(sky_only, ref)[0] < TOL_FP32
        assert galaxy_share(s, ref, oracle) < 1e-6
