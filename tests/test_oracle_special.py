"""Pins the special functions of the oracle AND of the product's host set-up
code against 40-digit mpmath values (tests/golden/special.json), closed forms,
and libprofit's own special-case tests (tests/test_utils.h:51-100).
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "special.json")))


def product():
    from libprofit_amd import _lib
    return _lib.load()


def impls(oracle):
    P = product()
    return {
        "oracle": (oracle.qgamma, oracle.pgamma, oracle.gammafn, oracle.beta),
        "product": (P.pk_qgamma, P.pk_pgamma, P.pk_gammafn, P.pk_beta),
    }


def relerr(a, b):
    return abs(a - b) / abs(b) if b != 0 else abs(a)


@pytest.mark.parametrize("who", ["oracle", "product"])
def test_qgamma_golden(oracle, who):
    qg = impls(oracle)[who][0]
    for p, a, x in GOLD["qgamma"]:
        assert relerr(qg(p, a), x) < 2e-14, (p, a, qg(p, a), x)


@pytest.mark.parametrize("who", ["oracle", "product"])
def test_pgamma_golden(oracle, who):
    pg = impls(oracle)[who][1]
    for x, a, v in GOLD["pgamma"]:
        assert abs(pg(x, a) - v) < 2e-15 + 1e-13 * v, (x, a, pg(x, a), v)


@pytest.mark.parametrize("who", ["oracle", "product"])
def test_gamma_beta_golden(oracle, who):
    _, _, gf, bt = impls(oracle)[who]
    for x, v in GOLD["gammafn"]:
        assert relerr(gf(x), v) < 1e-13, (x, gf(x), v)
    for a, b, v in GOLD["beta"]:
        assert relerr(bt(a, b), v) < 1e-13, (a, b, bt(a, b), v)


@pytest.mark.parametrize("who", ["oracle", "product"])
def test_closed_forms(oracle, who):
    qg, pg, gf, bt = impls(oracle)[who]
    # unit exponential distribution: P(1, x) = 1 - e^-x, quantile -log(1-p)
    for p in (1e-6, 0.01, 0.5, 0.9, 0.9999):
        assert relerr(qg(p, 1.0), -math.log1p(-p)) < 1e-14
    # integer shape: P(a, x) = 1 - e^-x sum_{k<a} x^k / k!
    for a in (2, 3, 5, 8):
        for x in (0.5, 3.0, 9.0):
            v = 1 - math.exp(-x) * sum(x ** k / math.factorial(k) for k in range(a))
            assert abs(pg(x, float(a)) - v) < 1e-14
    # chi-square with 1 dof: P(1/2, x) = erf(sqrt(x))
    for x in (0.01, 0.3, 2.0, 7.0):
        assert abs(pg(x, 0.5) - math.erf(math.sqrt(x))) < 1e-14
    # Gamma at integers and half integers, beta symmetry and B(1,b)=1/b
    for n in range(1, 15):
        assert relerr(gf(float(n)), float(math.factorial(n - 1))) < 1e-14
    assert relerr(gf(0.5), math.sqrt(math.pi)) < 1e-15
    for b in (0.3, 1.0, 4.5):
        assert relerr(bt(1.0, b), 1 / b) < 1e-14
        assert bt(2.0, b) == pytest.approx(bt(b, 2.0), rel=1e-14)


@pytest.mark.parametrize("who", ["oracle", "product"])
def test_special_cases(oracle, who):
    # tests/test_utils.h:51-100
    _, _, gf, bt = impls(oracle)[who]
    for x in (-100., -50., -1., 0.):
        assert math.isnan(gf(x))
    for x in (-1000.1, -2000.5):
        assert gf(x) == 0.
    for x in (-100.1, -50.1, -1.1, 0.1, 1., 2., 3.1, 10., 50., 70., 100.):
        r = gf(x)
        assert not math.isnan(r) and not math.isinf(r)
    for x in (200., 2500., 10000.):
        assert math.isinf(gf(x))
    for x in (-100., -10., -5., -4.5, -1., -0.1):
        assert math.isnan(bt(1, x)) and math.isnan(bt(x, 1))
    assert math.isinf(bt(0, 1)) and math.isinf(bt(1, 0))
    for a in (1., 2., 3., 4., 10., 100., 200.):
        for b in (1., 2., 3., 4., 10., 100., 200.):
            r = bt(a, b)
            assert not math.isinf(r) and not math.isnan(r) and r != 0
    for x in (1000., 2000., 3000.):
        assert bt(x, x) == 0.


@pytest.fixture
def exact_integrals(oracle):
    """Luminosity integrals by the exact integrators (product and oracle)
    for the duration of a test."""
    from libprofit_amd import model as M
    M.set_integrator("exact")
    oracle.set_integrator("exact")
    yield
    M.set_integrator("quadpack")
    oracle.set_integrator("quadpack")


def test_product_integrators(exact_integrals):
    """pk_integrate_qags / pk_integrate_qagi (exact mode) against closed forms."""
    from libprofit_amd import _lib
    P = product()
    f1 = _lib.INTEGRAND(lambda x, _: x * math.exp(-x))
    assert relerr(P.pk_integrate_qagi(C.cast(f1, C.c_void_p), 0.0, None), 1.0) < 1e-12
    f2 = _lib.INTEGRAND(lambda x, _: math.sqrt(x))
    assert relerr(P.pk_integrate_qags(C.cast(f2, C.c_void_p), 0.0, 4.0, None), 16.0 / 3) < 1e-12
    f3 = _lib.INTEGRAND(lambda x, _: 1.0 / (1 + x * x))
    assert relerr(P.pk_integrate_qagi(C.cast(f3, C.c_void_p), 0.0, None), math.pi / 2) < 1e-12


def _setup(oracle, name, **kw):
    from libprofit_amd import abi
    spec = abi.ModelSpec(10, 10, [abi.make_component(name, mag=0.0, axrat=1.0, box=0.0, **kw)])
    return oracle.component_setup(spec)


def test_lumtot_golden(oracle, exact_integrals):
    """Exact integrator: Ie = 10^(-0.4 mag) / (lumtot * axrat / r_box): with mag = 0, axrat = 1,
    box = 0 (r_box = 1), 1/Ie is the profile's total luminosity."""
    for re, n, v in GOLD["sersic_lumtot"]:
        s = _setup(oracle, "sersic", re=re, nser=n)
        assert relerr(1 / s["ie"], v) < 1e-13, (re, n)
    for rc, rt, a, v in GOLD["king_lumtot"]:
        s = _setup(oracle, "king", rc=rc, rt=rt, a=a)
        assert relerr(1 / s["ie"], v) < 1e-10, (rc, rt, a, 1 / s["ie"], v)
    for h1, h2, rb, a, v in GOLD["brokenexp_lumtot"]:
        s = _setup(oracle, "brokenexp", h1=h1, h2=h2, rb=rb, a=a)
        assert relerr(1 / s["ie"], v) < 1e-10, (h1, h2, rb, a, 1 / s["ie"], v)
    for re, rb, n, a, b, v in GOLD["coresersic_lumtot"]:
        s = _setup(oracle, "coresersic", re=re, rb=rb, nser=n, a=a, b=b)
        assert relerr(1 / s["ie"], v) < 1e-9, (re, rb, n, a, b, 1 / s["ie"], v)


def test_r_box_golden(oracle):
    """For box != 0 the total luminosity is divided by r_box
    (src/radial.cpp:150-152)."""
    from libprofit_amd import abi
    for box, rb in GOLD["r_box"]:
        s0 = _setup(oracle, "moffat", fwhm=3.0, con=2.5)
        spec = abi.ModelSpec(10, 10, [abi.make_component("moffat", mag=0.0, axrat=1.0, box=box, fwhm=3.0, con=2.5)])
        sb = oracle.component_setup(spec)
        assert relerr(sb["ie"] / s0["ie"], rb) < 1e-13, box


def test_adjust_values_sersic100(oracle):
    """Auto-adjusted subsampling constants of the 100x100 Sersic (n = 4,
    re = 10; src/radial.cpp:159-186, src/sersic.cpp:257-281), computed
    independently from the golden bn."""
    from libprofit_amd import abi
    spec = abi.ModelSpec(100, 100, [abi.make_component("sersic", xcen=50.3, ycen=49.7, mag=15, re=10, nser=4, ang=30,
                                                       axrat=0.6, box=0)])
    s = oracle.component_setup(spec)
    bn = [x for p, a, x in GOLD["qgamma"] if p == 0.5 and a == 8.0][0]
    assert relerr(s["bn"], bn) < 1e-15
    # fluxfrac(f) = re * (qgamma(f, 2n)/bn)^n
    q9999 = [x for p, a, x in GOLD["qgamma"] if p == 0.9999 and a == 8.0][0]
    rmax = math.ceil(max(10 * (q9999 / bn) ** 4, 2) / 10)
    assert s["rscale_max"] == rmax == 81
    assert s["rscale_switch"] == 2.0  # clamp(ceil(fluxfrac(0.992)), 2, 20) / re
    assert s["resolution"] == 8
    assert relerr(s["acc"], max(0.1, 0.1 / 2) / 0.6) < 1e-15
    lum = [v for re, n, v in GOLD["sersic_lumtot"] if re == 10.0 and n == 4.0][0]
    assert relerr(s["ie"], 10 ** (-0.4 * 15) / (lum * 0.6)) < 1e-13
    assert s["cos_ang"] == pytest.approx(math.cos(math.radians(120)), abs=1e-15)
    assert s["sin_ang"] == pytest.approx(math.sin(math.radians(120)), abs=1e-15)


def _product_constants(name, **kw):
    import ctypes as C
    from libprofit_amd import _lib, abi
    spec = abi.ModelSpec(10, 10, [abi.make_component(name, mag=0.0, axrat=1.0, box=0.0, **kw)])
    out = abi.Constants()
    _lib.check(_lib.load().pk_component_constants(spec.ref, 0, C.byref(out)))
    return out


def test_product_lumtot_golden(exact_integrals):
    """The product's set-up (host copy of the device code, exact integrator)
    against mpmath."""
    for re, n, v in GOLD["sersic_lumtot"]:
        assert relerr(_product_constants("sersic", re=re, nser=n).lumtot, v) < 1e-13, (re, n)
    for rc, rt, a, v in GOLD["king_lumtot"]:
        assert relerr(_product_constants("king", rc=rc, rt=rt, a=a).lumtot, v) < 1e-12, (rc, rt, a)
    for h1, h2, rb, a, v in GOLD["brokenexp_lumtot"]:
        assert relerr(_product_constants("brokenexp", h1=h1, h2=h2, rb=rb, a=a).lumtot, v) < 1e-12, (h1, h2, rb, a)
    for re, rb, n, a, b, v in GOLD["coresersic_lumtot"]:
        got = _product_constants("coresersic", re=re, rb=rb, nser=n, a=a, b=b).lumtot
        assert relerr(got, v) < 1e-12, (re, rb, n, a, b, got, v)


def test_product_constants_match_oracle(oracle):
    """All adjusted set-up constants of the product equal the oracle's."""
    from libprofit_amd import abi
    rng = np.random.Generator(np.random.PCG64(1))
    cases = [("sersic", dict(re=rng.uniform(1, 40), nser=rng.uniform(0.3, 9), axrat=rng.uniform(0.1, 1)))
             for _ in range(30)]
    cases += [("moffat", dict(fwhm=rng.uniform(1, 9), con=rng.uniform(1.2, 9))) for _ in range(10)]
    cases += [("ferrer", dict(rout=rng.uniform(3, 60), a=rng.uniform(0.2, 2), b=rng.uniform(0, 1.9)))
              for _ in range(10)]
    for name, kw in cases:
        spec = abi.ModelSpec(10, 10, [abi.make_component(name, mag=17.0, **kw)], magzero=2.0)
        o = oracle.component_setup(spec)
        import ctypes as C
        from libprofit_amd import _lib
        p = abi.Constants()
        _lib.check(_lib.load().pk_component_constants(spec.ref, 0, C.byref(p)))
        for k in ("rscale", "ie", "acc", "rscale_switch", "rscale_max", "bn", "cos_ang", "sin_ang"):
            assert relerr(getattr(p, k), o[k]) < 1e-13 or abs(getattr(p, k) - o[k]) < 1e-15, (name, kw, k)
        assert p.resolution == o["resolution"], (name, kw)
