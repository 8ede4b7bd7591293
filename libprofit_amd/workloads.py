"""Synthetic inputs of the benchmark configurations (seeded, deterministic).

  single_sersic_psf(seed)   400x400 image, one Sersic convolved with a 25x25
                            Moffat PSF (fwhm 3, con 2.5)
  galaxy_batch(n, seed)     n square stamps with side ~U{32..512}, 1-3
                            components drawn from the six radial profiles plus a
                            sky level, each convolved with its own Moffat PSF of
                            odd side ~U{15..41}
All draws use numpy's PCG64 in the order written here.
"""
import numpy as np

from . import abi

RADIAL = ("sersic", "coresersic", "moffat", "ferrer", "king", "brokenexp")


def moffat_psf_image(size, fwhm, con):
    """A size x size Moffat PSF sampled at pixel centres, unit sum."""
    rs = fwhm / (2 * np.sqrt(2 ** (1 / con) - 1))
    c = (size - 1) / 2.0
    y, x = np.mgrid[0:size, 0:size]
    r2 = ((x - c) ** 2 + (y - c) ** 2) / rs ** 2
    img = (1 + r2) ** (-con)
    return img / img.sum()


def single_sersic_psf(seed, psf=None):
    rng = np.random.Generator(np.random.PCG64(seed))
    nser = rng.uniform(0.5, 8)
    re = rng.uniform(2, 50)
    axrat = rng.uniform(0.1, 1)
    ang = rng.uniform(0, 180)
    mag = rng.uniform(12, 20)
    xcen, ycen = rng.uniform(150, 250, size=2)
    if psf is None:
        psf = moffat_psf_image(25, 3.0, 2.5)
    comp = abi.make_component("sersic", xcen=xcen, ycen=ycen, mag=mag, re=re, nser=nser, ang=ang, axrat=axrat,
                              convolve=True)
    return abi.ModelSpec(400, 400, [comp], psf=psf)


def _profile_params(rng, name, s):
    u = rng.uniform
    if name == "sersic":
        return dict(nser=u(0.5, 6), re=u(2, max(2.0, 0.2 * s)))
    if name == "moffat":
        return dict(fwhm=u(2, 10), con=u(1.5, 10))
    if name == "ferrer":
        return dict(rout=u(5, max(5.0, 0.3 * s)), a=u(0.3, 2), b=u(0, 1.5))
    if name == "king":
        return dict(rc=u(2, 8), rt=u(15, max(15.0, 0.3 * s)), a=u(1, 3))
    if name == "brokenexp":
        h1 = u(5, 15)
        return dict(h1=h1, h2=u(2, h1), rb=u(5, 15), a=u(0.1, 1))
    if name == "coresersic":
        return dict(re=u(5, 20), rb=u(1, 5), nser=u(1, 6), a=u(1, 3), b=u(0, 1.5))
    raise ValueError(name)


def galaxy_batch(n=4096, seed=0, smin=32, smax=512, magzero=30.0, profiles=RADIAL):
    """List of abi.ModelSpec, one stamp per galaxy.  Magnitudes U[14,20] are
    counted against `magzero` (30: total fluxes 1e4 .. 2.5e6 against a sky of
    U[0,10] per pixel, so the galaxies dominate the pixels near their
    centres).  `profiles` restricts the profile draw (same draw order)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    specs = []
    for _ in range(n):
        s = int(rng.integers(smin, smax + 1))
        ncomp = int(rng.integers(1, 4))
        comps = []
        for _c in range(ncomp):
            name = profiles[int(rng.integers(0, len(profiles)))]
            xc, yc = rng.uniform(0.4 * s, 0.6 * s, size=2)
            mag = rng.uniform(14, 20)
            ang = rng.uniform(0, 180)
            axrat = rng.uniform(0.2, 1)
            params = _profile_params(rng, name, s)
            comps.append(abi.make_component(name, xcen=xc, ycen=yc, mag=mag, ang=ang, axrat=axrat, box=0.0,
                                            convolve=True, **params))
        comps.append(abi.make_component("sky", bg=rng.uniform(0, 10)))
        k = 2 * int(rng.integers(7, 21)) + 1  # odd, 15..41
        fwhm = rng.uniform(2, 4)
        con = rng.uniform(2, 5)
        specs.append(abi.ModelSpec(s, s, comps, psf=moffat_psf_image(k, fwhm, con), magzero=magzero))
    return specs


def pixel_count(specs):
    return sum(int(s.m.width) * int(s.m.height) for s in specs)


# ---- fits: chi^2 and its gradient -------------------------------------------

def bulge_disk_fit(size=256, seed=0, psf_size=31, fwhm=3.5, con=3.0, sky=None, true=None):
    """A bulge + disk fit: (true components, PSF). Defaults: the 256x256
    differentiation workload -- bulge sersic nser 4, re 8, mag 17, axrat 0.8,
    ang 40; disk sersic nser 1, re 30, axrat 0.4, mag 16, ang 40; both centred
    on (128.3, 127.6), box 0; 31x31 Moffat PSF (fwhm 3.5, con 3).  `sky`
    adds a sky component of that level."""
    c = size / 2.0
    if true is None:
        true = [dict(xcen=c + 0.3, ycen=c - 0.4, mag=17.0, re=8.0, nser=4.0, ang=40.0, axrat=0.8, box=0.0),
                dict(xcen=c + 0.3, ycen=c - 0.4, mag=16.0, re=30.0 * size / 256, nser=1.0, ang=40.0, axrat=0.4,
                     box=0.0)]
    comps = [abi.make_component("sersic", convolve=True, **p) for p in true]
    if sky is not None:
        comps.append(abi.make_component("sky", bg=sky))
    return comps, moffat_psf_image(psf_size, fwhm, con)


def perturbed(params, rng, scale=1.0):
    """Start of a fit: the true parameters moved by a few per cent (seeded)."""
    p = dict(params)
    u = lambda a: scale * rng.uniform(-a, a)
    p["xcen"] += u(0.5)
    p["ycen"] += u(0.5)
    p["mag"] += u(0.2)
    p["re"] *= 1 + u(0.08)
    p["nser"] *= 1 + u(0.08)
    p["ang"] += u(5.0)
    p["axrat"] = min(1.0, p["axrat"] * (1 + u(0.05)))
    p["box"] += u(0.1)
    return p


def fit_batch(n=1024, seed=0, size=128, psf_size=31, magzero=30.0):
    """n independent bulge + disk + sky fits of size x size images with a
    psf_size Moffat PSF, magnitudes against `magzero`.  Per fit i (PCG64 seed
    `seed + i`): true parameters drawn around a compact bulge (nser U[2.5,5],
    re U[2,6]) and an extended disk (nser U[0.6,0.95], re U[8,20]) with a
    common centre, the sky level U[1,20]; the fit's starting point
    `perturbed` from them.  Returns (start specs, true specs)."""
    start, true = [], []
    for i in range(n):
        rng = np.random.Generator(np.random.PCG64(seed + i))
        c = size / 2.0
        xc, yc = rng.uniform(c - 3, c + 3, size=2)
        ang = rng.uniform(0, 180)
        b = dict(xcen=xc, ycen=yc, mag=rng.uniform(16, 18), re=rng.uniform(2, 6), nser=rng.uniform(2.5, 5),
                 ang=ang, axrat=rng.uniform(0.6, 1.0), box=0.0)
        d = dict(xcen=xc, ycen=yc, mag=rng.uniform(15, 17), re=rng.uniform(8, 20), nser=rng.uniform(0.6, 0.95),
                 ang=ang + rng.uniform(-10, 10), axrat=rng.uniform(0.3, 0.8), box=0.0)
        sky = rng.uniform(1.0, 20.0)
        fwhm, con = rng.uniform(2.5, 4.0), rng.uniform(2.5, 4.0)
        psf = moffat_psf_image(psf_size, fwhm, con)
        tc = [abi.make_component("sersic", convolve=True, **b), abi.make_component("sersic", convolve=True, **d),
              abi.make_component("sky", bg=sky)]
        sc = [abi.make_component("sersic", convolve=True, **perturbed(b, rng)),
              abi.make_component("sersic", convolve=True, **perturbed(d, rng)),
              abi.make_component("sky", bg=sky * (1 + rng.uniform(-0.1, 0.1)))]
        true.append(abi.ModelSpec(size, size, tc, psf=psf, magzero=magzero))
        start.append(abi.ModelSpec(size, size, sc, psf=psf, magzero=magzero))
    return start, true


def noisy(image, rng, rel_sigma=0.01):
    """image + N(0, sigma), sigma = rel_sigma * peak; returns (data, sigma)."""
    sigma = rel_sigma * float(np.max(image))
    return image + rng.normal(0.0, sigma, size=image.shape), sigma
