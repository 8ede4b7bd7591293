"""CPU checks of the chi^2-gradient machinery: the set-up derivative constants
of the product against mpmath, and the oracle's finite differences with a
frozen refinement pattern against what the mathematics fixes."""
import numpy as np
import pytest

from libprofit_amd import abi
from grad_util import FitModel, chi2_of, oracle_fd_grad


def _constants(**params):
    from libprofit_amd._lib import check, load
    import ctypes as C
    spec = abi.ModelSpec(64, 64, [abi.make_component("sersic", **params)])
    out = abi.Constants()
    check(load().pk_component_constants(spec.ref, 0, C.byref(out)))
    return out


@pytest.mark.parametrize("nser,box,axrat,re", [(0.5, 0.0, 0.7, 3.0), (1.0, 0.3, 0.4, 8.0), (2.5, -0.4, 1.0, 5.0),
                                               (4.0, 0.0, 0.8, 8.0), (7.5, 1.2, 0.2, 20.0)])
def test_gradient_constants_match_mpmath(nser, box, axrat, re):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 30

    def bn(n):
        return mp.findroot(lambda x: mp.gammainc(2 * n, 0, x, regularized=True) - mp.mpf(1) / 2,
                           2 * n - mp.mpf(1) / 3)

    def log_lum(r, n):  # log of src/sersic.cpp:283-289
        b = bn(n)
        return 2 * mp.log(r) + mp.log(2 * mp.pi) + mp.log(n) + mp.loggamma(2 * n) + b - 2 * n * mp.log(b)

    def log_rbox(B):  # src/radial.cpp:150-151
        return mp.log(mp.pi * B / (2 * mp.beta(1 / B, 1 / B)))

    n, B, r = mp.mpf(nser), mp.mpf(box) + 2, mp.mpf(re)
    c = _constants(re=re, nser=nser, box=box, axrat=axrat)
    exp = {
        "dbn_dnser": mp.diff(bn, n),
        "dlogflux_dre": -mp.diff(lambda t: log_lum(t, n), r),
        "dlogflux_dnser": -mp.diff(lambda t: log_lum(r, t), n),
        "dlogflux_daxrat": -1 / mp.mpf(axrat),
        "dlogflux_dbox": mp.diff(log_rbox, B),
    }
    for k, v in exp.items():
        got = getattr(c, k)
        assert abs(got - float(v)) <= 1e-10 * max(1.0, abs(float(v))), (k, got, float(v))


def _small_fit():
    from libprofit_amd import workloads
    psf = workloads.moffat_psf_image(9, 2.5, 3.0)
    comps = [("sersic", dict(xcen=20.3, ycen=19.6, mag=15, re=3.0, nser=2.5, ang=30, axrat=0.7, box=0.0)),
             ("sersic", dict(xcen=20.1, ycen=19.9, mag=14.5, re=7.0, nser=0.8, ang=50, axrat=0.45, box=0.1)),
             ("sky", dict(bg=0.3))]
    return FitModel(40, comps, psf)


def test_frozen_pattern_replay_is_exact(oracle):
    fm = _small_fit()
    ref, _ = oracle.evaluate(fm.spec())
    oracle.pattern("record")
    rec, _ = oracle.evaluate(fm.spec())
    oracle.pattern("replay")
    rep, _ = oracle.evaluate(fm.spec())
    ok = oracle.pattern_ok()
    oracle.pattern("off")
    assert ok
    assert np.array_equal(ref, rec) and np.array_equal(ref, rep)


def test_frozen_fd_of_linear_parameters_is_exact(oracle):
    """chi^2 is quadratic in the magnitudes' fluxes and the sky level: the
    central differences must equal the closed forms
      d chi^2 / d mag_c = sum g * (-0.4 ln 10) m_c,  d chi^2 / d bg = sum g,
    with g = -2 (d - m) / sigma^2 and m_c the image of component c alone."""
    fm = _small_fit()
    theta = fm.theta()
    rng = np.random.Generator(np.random.PCG64(1))
    img, _ = oracle.evaluate(fm.spec())
    sigma = 0.01 * img.max()
    data = img + rng.normal(0, sigma, img.shape) + 0.02 * img.max()
    g_fd, c0 = oracle_fd_grad(oracle, fm, theta, data, sigma)
    assert c0 == pytest.approx(chi2_of(img, data, sigma), rel=1e-14)
    g = -2 * (data - img) / sigma ** 2
    names = fm.names()
    for ci in range(2):
        only = FitModel(fm.size, [(fm.kinds[ci], fm.base[ci])], fm.psf)
        mc, _ = oracle.evaluate(only.spec())
        terms = g * (-0.4 * np.log(10)) * mc
        exp, scale = float(np.sum(terms)), float(np.sum(np.abs(terms)))
        k = names.index("%d.mag" % ci)
        # central-difference truncation: h^2/6 of the third derivative, h = 1e-4 mag
        assert abs(g_fd[k] - exp) <= 2e-6 * scale, (names[k], g_fd[k], exp, scale)
    k = names.index("2.bg")
    assert abs(g_fd[k] - float(np.sum(g))) <= 1e-9 * float(np.sum(np.abs(g)))
