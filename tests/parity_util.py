"""Shared helpers for the parity tests (device result vs FP64 oracle)."""
import numpy as np

from libprofit_amd import abi

# Tolerances stated by the target: per-pixel relative error on pixels above
# 1e-6 of the peak, FP64 mode and FP32 mode.
TOL_FP64 = 1e-12
TOL_FP32 = 1e-5
PEAK_FLOOR = 1e-6


def rel_errors(got, ref, floor=PEAK_FLOOR):
    """(max per-pixel relative error on pixels above floor*peak, total-flux relative error)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    peak = np.max(np.abs(ref)) if ref.size else 0.0
    if peak == 0:
        return float(np.max(np.abs(got))) if got.size else 0.0, 0.0
    sel = np.abs(ref) > floor * peak
    pix = float(np.max(np.abs(got[sel] - ref[sel]) / np.abs(ref[sel])))
    tot_ref = ref.sum()
    flux = float(abs(got.sum() - tot_ref) / abs(tot_ref)) if tot_ref != 0 else 0.0
    return pix, flux


def sersic(**kw):
    return abi.make_component("sersic", **kw)


def comp(name, **kw):
    return abi.make_component(name, **kw)


def moffat_psf(size, fwhm, con):
    """A PSF image drawn by the model itself: a moffat profile centred on a
    size x size stamp, normalised to unit sum (the C2-C5 workloads)."""
    return abi.ModelSpec(size, size, [comp("moffat", xcen=size / 2.0, ycen=size / 2.0, fwhm=fwhm, con=con,
                                           mag=0.0)])


def sersic100_adjusted():
    return abi.ModelSpec(100, 100, [sersic(xcen=50.3, ycen=49.7, mag=15, re=10, nser=4, ang=30, axrat=0.6, box=0)])


def sersic100_fixed():
    return abi.ModelSpec(100, 100, [sersic(xcen=50.3, ycen=49.7, mag=15, re=10, nser=4, ang=30, axrat=0.6, box=0,
                                           adjust=False, rscale_switch=1.1, resolution=9, max_recursions=2,
                                           acc=0.1, rscale_max=0)])


def galaxy400_params(seed):
    """400x400 galaxy draws (numpy PCG64, in this order)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nser = rng.uniform(0.5, 8)
    re = rng.uniform(2, 50)
    axrat = rng.uniform(0.1, 1)
    ang = rng.uniform(0, 180)
    mag = rng.uniform(12, 20)
    xcen, ycen = rng.uniform(150, 250, size=2)
    return dict(nser=nser, re=re, axrat=axrat, ang=ang, mag=mag, xcen=xcen, ycen=ycen)


def without_sky(spec):
    """The same model without its sky components (the galaxy-only image)."""
    comps = [abi.Component.from_buffer_copy(spec.m.components[k]) for k in range(spec.m.n_components)]
    comps = [c for c in comps if c.kind != abi.PROFILE_KINDS["sky"]]
    m = spec.m
    return abi.ModelSpec(m.width, m.height, comps, psf=spec._psf, mask=spec._mask, scale=(m.scale_x, m.scale_y),
                         psf_scale=(m.psf_scale_x, m.psf_scale_y), magzero=m.magzero, finesampling=m.finesampling,
                         crop=m.crop, return_finesampled=m.return_finesampled, adjust_mask=m.adjust_mask,
                         dry_run=m.dry_run, convolver=m.convolver, reuse_krn_fft=m.reuse_krn_fft)


def galaxy_share(spec, ref, oracle, floor=PEAK_FLOOR):
    """Largest share of a pixel value that the non-sky profiles contribute,
    over the pixels a comparison at `floor` of the peak looks at."""
    ref = np.asarray(ref, dtype=np.float64)
    peak = np.max(np.abs(ref)) if ref.size else 0.0
    if peak == 0:
        return 0.0
    gal, _ = oracle.evaluate(without_sky(spec))
    sel = np.abs(ref) > floor * peak
    return float(np.max(np.abs(gal[sel]) / np.abs(ref[sel])))


def check_image(got, spec, ref, oracle, tol, what=""):
    """Per-pixel and total-flux error of `got` against the oracle image `ref`
    within `tol`, on a model whose galaxy is observable at that tolerance: on
    some compared pixel the non-sky profiles make up at least 1e3 * tol of
    the value, so an image that lost its galaxy (or its convolution or its
    sub-pixel refinement) fails."""
    pix, flux = rel_errors(got, ref)
    assert pix <= tol, (what, "pixel", pix)
    assert flux <= tol, (what, "flux", flux)
    if any(spec.m.components[k].kind != abi.PROFILE_KINDS["sky"] for k in range(spec.m.n_components)):
        share = galaxy_share(spec, ref, oracle)
        assert share >= 1e3 * tol, (what, "galaxy share of the pixels", share)
    return pix, flux


def stratified_sample(specs, n, seed=12345):
    """Indices of up to n models: for every PSF side its smallest and its
    largest stamp, the rest drawn at random (seeded)."""
    by_k = {}
    for i, s in enumerate(specs):
        by_k.setdefault(int(s.m.psf_width), []).append(i)
    pick = set()
    for _k, idx in sorted(by_k.items()):
        idx = sorted(idx, key=lambda i: (specs[i].m.width, i))
        pick.update((idx[0], idx[-1]))
    rng = np.random.Generator(np.random.PCG64(seed))
    for i in rng.permutation(len(specs)).tolist():
        if len(pick) >= n:
            break
        pick.add(i)
    return sorted(pick)
