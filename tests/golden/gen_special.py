"""Writes tests/golden/special.json: high-precision (mpmath, 40 digits)
values of the special functions and normalisation integrals the image path
depends on.

libprofit takes them from GSL or Rmath (src/utils.cpp:101-259), neither of
which is available; these values come from their mathematical definitions:
  qgamma(p, a)  root x of P(a, x) = p                 (gsl_cdf_gamma_Pinv)
  pgamma(x, a)  regularised lower incomplete gamma    (gsl_cdf_gamma_P)
  gammafn(x)    Gamma(x)                              (gsl_sf_gamma)
  beta(a, b)    Gamma(a)Gamma(b)/Gamma(a+b)           (gsl_sf_beta)
  lumtot        2 pi int_0^R r f(r) dr of the profiles whose libprofit
                normalisation is a QUADPACK integral (src/king.cpp:87-97,
                src/brokenexponential.cpp:115-127, src/coresersic.cpp:92-104)
  sersic lumtot closed form (src/sersic.cpp:283-289) and r_box
                (src/radial.cpp:150-151).
Run: python tests/golden/gen_special.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 40
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "special.json")


def f(x):
    return float(x)


def qgamma(p, a):
    p, a = mp.mpf(p), mp.mpf(a)
    g = lambda x: mp.gammainc(a, 0, x, regularized=True) - p  # noqa: E731
    # bracket then solve
    lo, hi = mp.mpf(0), a + 10 * mp.sqrt(a) + 50
    x0 = a * (1 - 1 / (9 * a)) ** 3 if a >= 1 else (p * mp.gamma(a + 1)) ** (1 / a)
    try:
        return mp.findroot(g, x0, tol=mp.mpf(10) ** -35)
    except Exception:
        return mp.findroot(g, (lo + mp.mpf(10) ** -30, hi), solver="anderson")


def king_lum(rc, rt, a):
    rc, rt, a = mp.mpf(rc), mp.mpf(rt), mp.mpf(a)
    t = 1 / (1 + (rt / rc) ** 2) ** (1 / a)
    fr = lambda r: r * (1 / (1 + (r / rc) ** 2) ** (1 / a) - t) ** a  # noqa: E731
    return 2 * mp.pi * mp.quad(fr, [0, rt / 2, rt])


def bexp_lum(h1, h2, rb, a):
    h1, h2, rb, a = map(mp.mpf, (h1, h2, rb, a))
    fr = lambda r: r * mp.exp(-r / h1 + (1 / h1 - 1 / h2) * mp.log(1 + mp.exp(a * (r - rb))) / a)  # noqa: E731
    return 2 * mp.pi * mp.quad(fr, [0, rb, 2 * rb + 10 * max(h1, h2), mp.inf])


def csersic_lum(re, rb, n, a, b):
    re, rb, n, a, b = map(mp.mpf, (re, rb, n, a, b))
    bn = qgamma(0.5, 2 * n)
    fr = lambda r: r * (1 + (r / rb) ** (-a)) ** (b / a) * mp.exp(-bn * ((r ** a + rb ** a) / re ** a) ** (1 / (n * a)))  # noqa: E731
    return 2 * mp.pi * mp.quad(fr, [0, rb / 100, rb, re, 4 * re, 40 * re, mp.inf])


def sersic_lum(re, n):
    re, n = mp.mpf(re), mp.mpf(n)
    bn = qgamma(0.5, 2 * n)
    return re ** 2 * 2 * mp.pi * n * mp.gamma(2 * n) * mp.exp(bn) / bn ** (2 * n)


def r_box(box):
    b2 = mp.mpf(box) + 2
    return mp.pi * b2 / (2 * mp.beta(1 / b2, 1 / b2))


def main():
    out = {}
    ns = [0.3, 0.5, 0.75, 1.0, 1.5, 2.0, 2.5, 3.0, 4.0, 5.0, 6.0, 8.0, 10.0, 16.0]
    out["qgamma"] = [[0.5, 2 * n, f(qgamma(0.5, 2 * n))] for n in ns]
    for p in (0.9999, 0.968, 0.5, 0.1, 1e-3, 0.999999):
        for a in (0.6, 1.0, 2.0, 5.0, 8.0, 16.0):
            out["qgamma"].append([p, a, f(qgamma(p, a))])
    out["pgamma"] = [[x, a, f(mp.gammainc(a, 0, x, regularized=True))]
                     for a in (0.6, 1.0, 2.0, 5.0, 8.0, 16.0) for x in (1e-3, 0.1, 1.0, 5.0, 15.0, 40.0)]
    out["gammafn"] = [[x, f(mp.gamma(x))] for x in (-100.1, -50.1, -1.1, 0.1, 0.5, 1.0, 2.0, 3.1, 10.0, 16.0, 50.0,
                                                   70.0, 100.0, 170.5)]
    out["beta"] = [[a, b, f(mp.beta(a, b))] for a in (0.25, 0.5, 1.0, 2.0, 3.0, 10.0, 100.0)
                   for b in (0.25, 0.5, 1.0, 2.0, 4.0, 10.0)]
    out["sersic_lumtot"] = [[re, n, f(sersic_lum(re, n))] for re in (1.0, 10.0) for n in ns]
    out["r_box"] = [[b, f(r_box(b))] for b in (-1.5, -0.3, 0.0, 0.4, 1.0, 2.0)]
    out["king_lumtot"] = [[rc, rt, a, f(king_lum(rc, rt, a))]
                          for (rc, rt, a) in ((1, 3, 2), (3, 20, 2), (2, 15, 1), (5, 40, 3), (2.5, 30, 1.5))]
    out["brokenexp_lumtot"] = [[h1, h2, rb, a, f(bexp_lum(h1, h2, rb, a))]
                               for (h1, h2, rb, a) in ((1, 1, 1, 1), (6, 3, 8, 0.5), (10, 2, 12, 0.3), (5, 5, 5, 1))]
    out["coresersic_lumtot"] = [[re, rb, n, a, b, f(csersic_lum(re, rb, n, a, b))]
                                for (re, rb, n, a, b) in ((1, 1, 4, 1, 1), (8, 2, 3, 2, 0.8), (15, 3, 1.5, 2.5, 0.2),
                                                          (10, 4, 5, 1.5, 1.4))]
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
