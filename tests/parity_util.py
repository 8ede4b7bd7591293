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
