"""Pins the oracle's boxy evaluation and the structure of its sub-pixel
refinement (RadialProfile::subsample_pixel, src/radial.cpp:55-135;
SersicProfile::subsampling_params, src/sersic.cpp:331-347;
include/profit/radial.h:93-102) with checks the mathematics fixes:

- rough images of boxy profiles against the boxy radius
  (|x'|^(b+2) + |y'|^(b+2))^(1/(b+2)) and the r_box normalisation, for
  every specialised Sersic form (src/sersic.cpp:225-255) and box < 0, > 0;
- total flux of boxy profiles against the magnitude;
- acc = infinity: no sub-cell is refined, every pixel inside rscale_switch
  is the res x res midpoint rule;
- acc = 0: every sub-cell is refined down to max_recursions, every pixel is
  the res^(k+1) x res^(k+1) midpoint rule, and the number of evaluations has
  the closed form sum_{l<k} 2 res^(2(l+1)) + res^(2(k+1)) per pixel (a test
  level <= max_recursions or a missing test evaluation changes it);
- mirror symmetry: a profile whose axes lie along the pixel grid, centred on
  the image, gives a mirror-symmetric image only because the accuracy test
  uses the point (|x'|, |y'| + |dy'|); the same refinement with the test
  point (x', y' + dy') (planted here in a small pure-Python refinement)
  breaks the symmetry;
- the Sersic centre-pixel override (nser > 1 and adjust: resolution 8, 10
  levels) applies exactly to the pixels whose centres lie within one pixel of
  the profile centre, brings them closer to the pixel integral, and does not
  apply for nser <= 1.
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

from libprofit_amd import abi
from parity_util import comp, rel_errors


def ev(oracle, spec):
    return oracle.evaluate(spec)[0]


def prof_coords(X, Y, xc, yc, ang, q):
    th = np.radians(ang + 90)
    x, y = X - xc, Y - yc
    return x * np.cos(th) + y * np.sin(th), (-x * np.sin(th) + y * np.cos(th)) / q


def boxy_r(xp, yp, box):
    b = box + 2
    return (np.abs(xp) ** b + np.abs(yp) ** b) ** (1 / b)


def sersic_f(r, re, n):
    bn = special.gammaincinv(2 * n, 0.5)
    return np.exp(-bn * ((r / re) ** (1 / n) - 1))


def sersic_lum(re, n):
    bn = special.gammaincinv(2 * n, 0.5)
    return re ** 2 * 2 * np.pi * n * special.gamma(2 * n) * np.exp(bn) / bn ** (2 * n)


def r_box(box):
    b = box + 2
    return np.pi * b / (2 * special.beta(1 / b, 1 / b))


# nser * (box + 2) hits each specialised form of the Sersic evaluation:
# 0.5, 1, 2, 3, 4, 8, 16 and the general power, with round and boxy radii
SERSIC_FORMS = [(0.25, 0.0), (0.5, 0.0), (1.0, 0.0), (1.5, 0.0), (2.0, 0.0), (4.0, 0.0), (8.0, 0.0), (2.7, 0.0),
                (0.5, -1.0), (0.4, 0.5), (1.0, 0.0), (2.0, -0.5), (1.6, 0.5), (3.2, 0.5), (8.0, -1.0), (1.3, 0.4),
                (2.1, -0.3)]


@pytest.mark.parametrize("n,box", SERSIC_FORMS)
def test_boxy_sersic_rough_image(oracle, n, box):
    W, H, xc, yc, ang, q, mag, re = 31, 27, 15.3, 12.8, 35.0, 0.7, 14.0, 4.0
    img = ev(oracle, abi.ModelSpec(W, H, [comp("sersic", xcen=xc, ycen=yc, ang=ang, axrat=q, mag=mag, box=box, re=re,
                                               nser=n, rough=True, rscale_max=1e6)], scale=(0.9, 1.1)))
    sx, sy = 0.9, 1.1
    X, Y = np.meshgrid((np.arange(W) + 0.5) * sx, (np.arange(H) + 0.5) * sy)
    xp, yp = prof_coords(X, Y, xc, yc, ang, q)
    f = sersic_f(boxy_r(xp, yp, box), re, n)
    exp = f * 10 ** (-0.4 * mag) / (sersic_lum(re, n) * q / r_box(box)) * sx * sy
    assert rel_errors(img, exp, floor=1e-12)[0] < 1e-10


@pytest.mark.parametrize("box", [-0.3, 0.4])
def test_boxy_moffat_rough_image(oracle, box):
    W, H, xc, yc, ang, q, mag, fwhm, con = 31, 27, 15.3, 12.8, 35.0, 0.7, 14.0, 5.0, 3.0
    img = ev(oracle, abi.ModelSpec(W, H, [comp("moffat", xcen=xc, ycen=yc, ang=ang, axrat=q, mag=mag, box=box,
                                               fwhm=fwhm, con=con, rough=True, rscale_max=1e6)]))
    X, Y = np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5)
    xp, yp = prof_coords(X, Y, xc, yc, ang, q)
    rs = fwhm / (2 * np.sqrt(2 ** (1 / con) - 1))
    f = (1 + (boxy_r(xp, yp, box) / rs) ** 2) ** (-con)
    lum = np.pi * rs ** 2 / (con - 1)
    exp = f * 10 ** (-0.4 * mag) / (lum * q / r_box(box))
    assert rel_errors(img, exp, floor=1e-12)[0] < 1e-10


@pytest.mark.parametrize("box", [-0.3, 0.4])
def test_boxy_total_flux(oracle, box):
    for name, p in (("sersic", dict(re=3.0, nser=1.0)), ("moffat", dict(fwhm=3.0, con=4.0))):
        spec = abi.ModelSpec(120, 120, [comp(name, xcen=60.2, ycen=59.9, mag=10.0, axrat=0.8, ang=10, box=box, **p)],
                             magzero=2.5)
        tot = ev(oracle, spec).sum()
        exp = 10 ** (-0.4 * (10.0 - 2.5))
        assert abs(tot - exp) < 2e-3 * exp, (name, box, tot, exp)


def _grid_average(W, H, k, f):
    """Per pixel: the mean of f over a k x k grid of sub-cell centres."""
    t = (np.arange(k) + 0.5) / k
    out = np.zeros((H, W))
    for j in range(H):
        for i in range(W):
            X, Y = np.meshgrid(i + t, j + t)
            out[j, i] = f(X, Y).mean()
    return out


REFINE = dict(xcen=4.6, ycen=4.3, mag=0.0, re=3.0, nser=1.0, ang=30.0, axrat=0.7, adjust=False, rscale_switch=100.0,
              rscale_max=0.0)


def _sersic_image(X, Y):
    p = REFINE
    xp, yp = prof_coords(X, Y, p["xcen"], p["ycen"], p["ang"], p["axrat"])
    return sersic_f(np.hypot(xp, yp), p["re"], p["nser"]) / (sersic_lum(p["re"], p["nser"]) * p["axrat"])


@pytest.mark.parametrize("res", [3, 4, 5])
def test_no_refinement_is_the_midpoint_rule(oracle, res):
    W = H = 9
    spec = abi.ModelSpec(W, H, [comp("sersic", acc=1e300, resolution=res, max_recursions=3, **REFINE)])
    img = ev(oracle, spec)
    refined, evals = oracle.stats()
    assert refined == W * H
    assert evals == W * H * 2 * res * res  # value + test point per sub-cell, no recursion
    exp = _grid_average(W, H, res, _sersic_image)
    assert rel_errors(img, exp, floor=0)[0] < 1e-12


@pytest.mark.parametrize("res,mr", [(2, 1), (3, 1), (2, 2), (3, 2), (2, 3)])
def test_full_refinement_is_the_fine_midpoint_rule(oracle, res, mr):
    W = H = 9
    spec = abi.ModelSpec(W, H, [comp("sersic", acc=0.0, resolution=res, max_recursions=mr, **REFINE)])
    img = ev(oracle, spec)
    refined, evals = oracle.stats()
    assert refined == W * H
    per_pixel = sum(2 * res ** (2 * (lv + 1)) for lv in range(mr)) + res ** (2 * (mr + 1))
    assert evals == W * H * per_pixel, (evals, W * H * per_pixel)
    exp = _grid_average(W, H, res ** (mr + 1), _sersic_image)
    assert rel_errors(img, exp, floor=0)[0] < 1e-12


def _py_refine(f, x0, x1, y0, y1, level, mr, res, acc, sin_a, cos_a, q, xc, yc, abs_test):
    """subsample_pixel in pure Python; abs_test=False plants a mistake: the
    accuracy test point (x', y' + dy') instead of (|x'|, |y'| + |dy'|)."""
    xb, yb = (x1 - x0) / res, (y1 - y0) / res
    tot = 0.0
    for i in range(res):
        for j in range(res):
            x, y = x0 + (i + 0.5) * xb, y0 + (j + 0.5) * yb
            xp = (x - xc) * cos_a + (y - yc) * sin_a
            yp = (-(x - xc) * sin_a + (y - yc) * cos_a) / q
            v = f(xp, yp)
            if level < mr:
                dyp = (-xb * sin_a + yb * cos_a) / q
                t = f(abs(xp), abs(yp) + abs(dyp)) if abs_test else f(xp, yp + dyp)
                if abs(t / v - 1) > acc:
                    tot += _py_refine(f, x - xb / 2, x + xb / 2, y - yb / 2, y + yb / 2, level + 1, mr, res, acc,
                                      sin_a, cos_a, q, xc, yc, abs_test)
                    continue
            tot += v
    return tot / (res * res)


def test_refinement_mirror_symmetry(oracle):
    W = H = 9
    p = dict(xcen=4.5, ycen=4.5, mag=0.0, re=2.0, nser=2.0, ang=0.0, axrat=0.6, adjust=False, acc=0.05,
             resolution=3, max_recursions=2, rscale_switch=100.0, rscale_max=0.0)
    img = ev(oracle, abi.ModelSpec(W, H, [comp("sersic", **p)]))
    for mirrored in (img[::-1, :], img[:, ::-1], img[::-1, ::-1]):
        assert np.max(np.abs(mirrored - img) / img) < 1e-12
    # the planted mistake breaks the symmetry of a mirrored pixel pair by far
    bn = special.gammaincinv(2 * 2.0, 0.5)
    f = lambda xp, yp: math.exp(-bn * (math.hypot(xp, yp) / 2.0) ** 0.5)  # noqa: E731
    th = math.radians(90.0)
    args = (0, 2, 3, 0.05, math.sin(th), math.cos(th), 0.6, 4.5, 4.5)
    for abs_test, expect_sym in ((True, True), (False, False)):
        a = _py_refine(f, 3.0, 4.0, 2.0, 3.0, *args, abs_test)
        b = _py_refine(f, 5.0, 6.0, 2.0, 3.0, *args, abs_test)  # the left-right mirror of pixel (3, 2)
        assert (abs(a / b - 1) < 1e-12) == expect_sym, (abs_test, a, b)


def _sersic_pixel_integral(i, j, xc, yc, re, n, ang, q):
    """Integral of the Sersic profile over pixel (i, j), split at the
    profile centre (its cusp) when the centre lies inside."""
    th = math.radians(ang + 90)
    bn = special.gammaincinv(2 * n, 0.5)

    def f(y, x):
        xp = (x - xc) * math.cos(th) + (y - yc) * math.sin(th)
        yp = (-(x - xc) * math.sin(th) + (y - yc) * math.cos(th)) / q
        return math.exp(-bn * ((math.hypot(xp, yp) / re) ** (1 / n) - 1))

    xs = [i, i + 1] if not (i < xc < i + 1) else [i, xc, i + 1]
    ys = [j, j + 1] if not (j < yc < j + 1) else [j, yc, j + 1]
    tot = 0.0
    for a, b in zip(xs[:-1], xs[1:]):
        for c, d in zip(ys[:-1], ys[1:]):
            tot += integrate.dblquad(f, a, b, c, d, epsabs=0, epsrel=1e-11)[0]
    return tot


@pytest.mark.parametrize("n", [2.0, 1.5, 0.8])
def test_sersic_centre_pixel_override(oracle, n):
    # re = 1: the adjusted resolution is 16, so the override's resolution 8
    # changes every pixel it applies to
    W = H = 20
    xc, yc, re, ang, q = 10.3, 9.7, 1.0, 30.0, 0.6
    spec = abi.ModelSpec(W, H, [comp("sersic", xcen=xc, ycen=yc, mag=0.0, re=re, nser=n, ang=ang, axrat=q)])
    img = ev(oracle, spec)
    s = oracle.component_setup(spec)
    # the same adjusted sub-pixel settings without the override
    plain = abi.ModelSpec(W, H, [comp("sersic", xcen=xc, ycen=yc, mag=0.0, re=re, nser=n, ang=ang, axrat=q,
                                      adjust=False, acc=s["acc"], rscale_switch=s["rscale_switch"],
                                      resolution=int(s["resolution"]), max_recursions=2,
                                      rscale_max=s["rscale_max"])])
    base = ev(oracle, plain)
    differ = {(i, j) for j in range(H) for i in range(W) if img[j, i] != base[j, i]}
    if n <= 1:
        assert not differ  # no override for nser <= 1
        return
    # pixels whose centre is within one pixel scale of the profile centre
    window = {(i, j) for j in range(H) for i in range(W) if abs(i + 0.5 - xc) < 1 and abs(j + 0.5 - yc) < 1}
    assert differ == window, (differ, window)
    assert s["resolution"] != 8
    # the pixel holding the centre comes closer to its integral
    i, j = int(xc), int(yc)
    exact = s["flux_scale"] * _sersic_pixel_integral(i, j, xc, yc, re, n, ang, q)
    print("centre pixel: override %.3e, without %.3e (relative error)"
          % (abs(img[j, i] / exact - 1), abs(base[j, i] / exact - 1)))
    assert abs(img[j, i] - exact) < abs(base[j, i] - exact), (img[j, i], base[j, i], exact)
