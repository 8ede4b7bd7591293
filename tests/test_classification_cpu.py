"""The error bound behind the FP32 pixel classification's guard bands.

The coarse pass forms the profile coordinates in FP32 from the FP64 offsets
dx, dy to the centre:
    x' = fma(dy, sin, dx cos),  y' = fma(dy, cos, -dx sin) / axrat,
    r^2 = fma(x', x', y' y'),
each operand rounded to FP32 (u = 2^-24).  Its guard bands assume
|r^2_fp32 - r^2| <= 31 u M r^2 with M = max(axrat^2, axrat^-2), and use a
relative half-width of 3e-5 M.  This emulates the FP32 operations (products
of two FP32 values are exact in FP64, so fma = one rounding of the FP64 sum
up to 2^-53) over random and adversarial inputs -- near-cancelling rotations,
extreme axis ratios, offsets from 1e-6 to 1e4 pixels -- and checks the bound
against the exact r^2 (FP64 of the unrounded inputs, Fraction for the worst
cases).
"""
from fractions import Fraction

import numpy as np

U = 2.0 ** -24
f32 = np.float32


def fma32(a, b, c):
    return f32(np.float64(a) * np.float64(b) + np.float64(c))


def r2_fp32(dx, dy, ang_rad, axrat):
    c, s, iq = f32(np.cos(ang_rad)), f32(np.sin(ang_rad)), f32(1.0 / axrat)
    dxf, dyf = f32(dx), f32(dy)
    dxc = f32(dxf * c)
    ndxs = f32(-dxf * s)
    xf = fma32(dyf, s, dxc)
    yf = f32(fma32(dyf, c, ndxs) * iq)
    return float(fma32(xf, xf, f32(yf * yf)))


def r2_exact(dx, dy, ang_rad, axrat):
    C, S = Fraction(np.cos(ang_rad)), Fraction(np.sin(ang_rad))
    X, Y = Fraction(dx), Fraction(dy)
    xp = X * C + Y * S
    yp = (Y * C - X * S) / Fraction(axrat)
    return xp * xp + yp * yp


def bound(r2, axrat):
    M = max(axrat ** 2, axrat ** -2)
    return 31 * U * M * r2


def test_band_bound_random():
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(20000):
        scale = 10 ** rng.uniform(-6, 4)
        dx, dy = rng.uniform(-1, 1, size=2) * scale
        ang = rng.uniform(0, 2 * np.pi)
        q = 10 ** rng.uniform(-3, 0)
        ex = r2_exact(dx, dy, ang, q)
        if ex == 0:
            continue
        err = abs(Fraction(r2_fp32(dx, dy, ang, q)) - ex)
        ratio = float(err / Fraction(bound(float(ex), q)))
        worst = max(worst, ratio)
        assert ratio <= 1.0, (dx, dy, ang, q, ratio)
    assert worst > 0.01  # the check is not vacuous


def test_band_bound_cancellation():
    """Pixels on the minor axis of a thin, rotated profile: x' cancels, y'
    carries the whole radius scaled by 1/axrat; and the reverse."""
    rng = np.random.default_rng(8)
    for _ in range(5000):
        ang = rng.uniform(0, 2 * np.pi)
        q = 10 ** rng.uniform(-3, 0)
        t = 10 ** rng.uniform(-2, 3)
        # along the profile's y' axis: (dx, dy) parallel to (-sin, cos)
        for dx, dy in ((-np.sin(ang) * t, np.cos(ang) * t), (np.cos(ang) * t, np.sin(ang) * t)):
            dx *= 1 + rng.uniform(-1e-7, 1e-7)
            ex = r2_exact(dx, dy, ang, q)
            err = abs(Fraction(r2_fp32(dx, dy, ang, q)) - ex)
            assert err <= Fraction(bound(float(ex), q)), (dx, dy, ang, q)


def test_band_width_covers_bound():
    """The band's relative half-width 3e-5 M exceeds the relative error bound
    31 u M (1.85e-6 M) with margin, and 3e-5 M >= 3e-5 covers the FP32
    rounding of the limit itself (u = 6e-8) and libprofit's own FP64 band
    (1e-10)."""
    for q in (1e-4, 1e-3, 0.05, 0.5, 1.0):
        M = max(q ** 2, q ** -2)
        assert 3e-5 * M > 10 * 31 * U * M
        assert 3e-5 * M > U + 1e-10
