"""The FP32 mode's pixel classification (coarse pass) against libprofit's FP64
one (src/radial.cpp:316-321): every pixel/component pair must land in the
same class -- beyond rscale_max, evaluated at its centre, or refined --
whatever the FP32 rounding of the profile coordinates.  The FP32 pass
decides outside guard bands around the two limits and hands the pixels
inside them to the FP64 test; these cases put pixel centres exactly on a
limit, use extreme axis ratios (wide bands, or bands covering everything)
and every radial profile.

Checks: the refined-pixel count of a single evaluation equals the oracle's
(and the FP64 mode's), the image meets the FP32 tolerance, and over a batch
the FP32 and FP64 passes count the same direct evaluations and refinements.
"""
import pytest

from libprofit_amd import abi, workloads
from libprofit_amd.model import Batch, evaluate_spec
from parity_util import TOL_FP32, check_image, comp

pytestmark = pytest.mark.gpu


def _cases():
    cases = []
    # centre on a pixel centre: pixel centres at integer offsets; r_switch = 5
    # and r_max = 10 pass exactly through centres (3,4), (5,0), (6,8), (10,0)
    for ang in (0.0, 90.0, 45.0):
        cases.append(("ties ang %g" % ang, abi.ModelSpec(
            101, 101, [comp("sersic", xcen=50.5, ycen=50.5, mag=15, re=5.0, nser=2.0, ang=ang, axrat=1.0,
                            adjust=False, rscale_switch=1.0, rscale_max=2.0, resolution=4, max_recursions=2,
                            acc=0.1)])))
    # elliptical ties: axrat 0.5 with the major axis along x (ang 90 -> x' = y)
    cases.append(("ties axrat 0.5", abi.ModelSpec(
        81, 81, [comp("sersic", xcen=40.5, ycen=40.5, mag=15, re=4.0, nser=1.5, ang=0.0, axrat=0.5,
                      adjust=False, rscale_switch=1.0, rscale_max=3.0, resolution=4, max_recursions=1, acc=0.1)])))
    # wide guard bands (axrat 0.05: bands of +-1.2 % of r^2) and bands over
    # everything (axrat 1e-3: all pixels through the FP64 test)
    for q in (0.05, 1e-3):
        cases.append(("axrat %g" % q, abi.ModelSpec(
            96, 96, [comp("sersic", xcen=47.3, ycen=48.9, mag=15, re=12.0, nser=2.5, ang=37.0, axrat=q)])))
    # every radial profile, boxy and not, adjusted limits
    prof = {
        "sersic": dict(re=7.0, nser=3.5),
        "moffat": dict(fwhm=5.0, con=2.5),
        "ferrer": dict(rout=18.0, a=1.5, b=0.5),
        "king": dict(rc=3.0, rt=20.0, a=2.0),
        "coresersic": dict(re=9.0, rb=2.0, nser=3.0, a=2.5, b=0.5),
        "brokenexp": dict(h1=6.0, h2=3.0, rb=10.0, a=0.8),
    }
    for name, p in prof.items():
        for box in (0.0, 0.4):
            cases.append(("%s box %g" % (name, box), abi.ModelSpec(
                72, 68, [comp(name, xcen=36.2, ycen=33.9, mag=15, ang=28.0, axrat=0.55, box=box, **p)])))
    # rough: never refined, only the rscale_max test
    cases.append(("rough", abi.ModelSpec(
        64, 64, [comp("sersic", xcen=31.7, ycen=32.2, mag=15, re=5.0, nser=1.0, ang=60.0, axrat=0.7, rough=True)])))
    # pixel scale far from 1
    cases.append(("scale 0.01", abi.ModelSpec(
        80, 80, [comp("sersic", xcen=0.401, ycen=0.397, mag=15, re=0.08, nser=2.0, ang=15.0, axrat=0.6)],
        scale=(0.01, 0.01))))
    return cases


CASES = _cases()


@pytest.mark.parametrize("k", range(len(CASES)), ids=[c[0] for c in CASES])
def test_fp32_classification_single(ctx32, ctx64, oracle, k):
    name, spec = CASES[k]
    print("CASE classification", name, flush=True)
    ref, _ = oracle.evaluate(spec)
    refined, _ = oracle.stats()
    img32, _, res32 = evaluate_spec(ctx32, spec)
    _, _, res64 = evaluate_spec(ctx64, spec)
    assert res64.n_refined_pixels == refined, (name, res64.n_refined_pixels, refined)
    assert res32.n_refined_pixels == refined, (name, res32.n_refined_pixels, refined)
    check_image(img32, spec, ref, oracle, TOL_FP32, name)


def test_fp32_classification_batch_counts(ctx32, ctx64):
    """Direct evaluations and refinements of the FP32 and FP64 coarse passes
    over a batch of the galaxy workload (every radial profile)."""
    print("CASE classification batch", flush=True)
    specs = workloads.galaxy_batch(48, 11, smin=24, smax=160)
    counts = []
    for ctx in (ctx32, ctx64):
        b = Batch(ctx, specs, graph=False)
        b.evaluate()
        b.sync()
        wc = b.work_counts()
        counts.append((wc.pixel_component_pairs, wc.coarse_evaluations, wc.refined_pixels))
        b.close()
    assert counts[0] == counts[1], counts
    assert counts[0][1] > 0 and counts[0][2] > 0, counts
