// FP64 special functions and quadrature used by the per-component set-up.
//
// They take the place of libprofit's math back-end entry points
// (include/profit/utils.h:52-113, implemented with GSL or Rmath in
// src/utils.cpp:101-259).  Every function is __host__ __device__: the host
// copy backs the exported pk_qgamma/... symbols and the single-model path,
// the device copy runs inside the batched set-up kernel, so both paths produce
// the same constants.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#ifndef PK_PI
#define PK_PI 3.14159265358979323846
#endif

namespace pk {

#define PK_HD __host__ __device__ inline

// Gamma function with libprofit's error semantics (src/utils.cpp:113-129):
// NaN at 0 and the negative integers, 0 on underflow, +Inf on overflow.
PK_HD double gammafn(double x)
{
	if (x != x) return NAN;
	if (x <= 0 && x == floor(x)) return NAN;
	double g = tgamma(x);
	if (isinf(g)) return x > 0 ? INFINITY : NAN;
	return g;
}

// Beta function with R semantics (include/profit/utils.h:79-92).
PK_HD double beta(double a, double b)
{
	if (a < 0 || b < 0) return NAN;
	if (a == 0 || b == 0) return INFINITY;
	if (a + b < 170.0) return tgamma(a) * tgamma(b) / tgamma(a + b);
	return exp(lgamma(a) + lgamma(b) - lgamma(a + b));
}

// Regularised incomplete gamma P(shape, x) (upper = false) or
// Q(shape, x) = 1 - P (upper = true): power series below x = shape + 1,
// Lentz continued fraction above, so that neither is formed as a difference of
// nearly equal numbers.
PK_HD double incgamma(double x, double a, bool upper)
{
	if (x != x || a != a || a <= 0) return NAN;
	if (x <= 0) return upper ? 1 : 0;
	if (isinf(x)) return upper ? 0 : 1;
	double lpre = a * log(x) - x - lgamma(a);
	if (x < a + 1) {
		double term = 1.0 / a, sum = term;
		for (int n = 1; n < 100000; n++) {
			term *= x / (a + n);
			sum += term;
			if (fabs(term) < fabs(sum) * 1e-17) break;
		}
		double pv = sum * exp(lpre);
		return upper ? 1 - pv : pv;
	}
	const double tiny = 1e-300;
	double b = x + 1 - a, c = 1 / tiny, d = 1 / b, h = d;
	for (int i = 1; i < 100000; i++) {
		double an = -i * (i - a);
		b += 2;
		d = an * d + b;
		if (fabs(d) < tiny) d = tiny;
		c = b + an / c;
		if (fabs(c) < tiny) c = tiny;
		d = 1 / d;
		double del = d * c;
		h *= del;
		if (fabs(del - 1) < 1e-17) break;
	}
	double qv = exp(lpre) * h;
	return upper ? qv : 1 - qv;
}

// Regularised lower incomplete gamma P(shape, x) (gsl_cdf_gamma_P(x, shape, 1)).
PK_HD double pgamma(double x, double a) { return incgamma(x, a, false); }

// Starting point for the gamma quantile: inverse normal by a rational
// approximation (Acklam), only used to seed the Halley iteration.
PK_HD double norm_quantile_seed(double p)
{
	const double a0 = -3.969683028665376e+01, a1 = 2.209460984245205e+02, a2 = -2.759285104469687e+02,
	             a3 = 1.383577518672690e+02, a4 = -3.066479806614716e+01, a5 = 2.506628277459239e+00;
	const double b0 = -5.447609879822406e+01, b1 = 1.615858368580409e+02, b2 = -1.556989798598866e+02,
	             b3 = 6.680131188771972e+01, b4 = -1.328068155288572e+01;
	const double c0 = -7.784894002430293e-03, c1 = -3.223964580411365e-01, c2 = -2.400758277161838e+00,
	             c3 = -2.549732539343734e+00, c4 = 4.374664141464968e+00, c5 = 2.938163982698783e+00;
	const double d0 = 7.784695709041462e-03, d1 = 3.224671290700398e-01, d2 = 2.445134137142996e+00,
	             d3 = 3.754408661907416e+00;
	if (p < 0.02425 || p > 1 - 0.02425) {
		double q = sqrt(-2 * log(p < 0.5 ? p : 1 - p));
		double v = (((((c0 * q + c1) * q + c2) * q + c3) * q + c4) * q + c5) /
		           ((((d0 * q + d1) * q + d2) * q + d3) * q + 1);
		return p < 0.5 ? v : -v;
	}
	double q = p - 0.5, r = q * q;
	return (((((a0 * r + a1) * r + a2) * r + a3) * r + a4) * r + a5) * q /
	       (((((b0 * r + b1) * r + b2) * r + b3) * r + b4) * r + 1);
}

// Quantile of the unit-scale gamma distribution: x with P(shape, x) = p.
// Halley steps on P from a Wilson-Hilferty (shape >= 1) or small-x seed.
PK_HD double qgamma(double p, double a)
{
	if (p != p || a != a || a <= 0 || p < 0 || p > 1) return NAN;
	if (p == 0) return 0;
	if (p == 1) return INFINITY;
	double x;
	if (a < 1) {
		x = pow(p * tgamma(a + 1), 1 / a);
		if (!(x > 0)) x = 1e-300;
	} else {
		double z = norm_quantile_seed(p);
		double t = 1 / (9 * a);
		double w = 1 - t + z * sqrt(t);
		x = a * w * w * w;
		if (!(x > 0)) x = a * 1e-3;
	}
	double lg = lgamma(a);
	const bool upper = p > 0.5;
	const double q = 1 - p;  // exact for p > 0.5
	double hprev = INFINITY;
	for (int it = 0; it < 200; it++) {
		double f = upper ? q - incgamma(x, a, true) : incgamma(x, a, false) - p;
		double pdf = exp((a - 1) * log(x) - x - lg);
		if (pdf == 0) break;
		double u = f / pdf;
		double h = u / (1 - 0.5 * u * ((a - 1) / x - 1));
		if (!isfinite(h)) h = u;
		double xn = x - h;
		if (xn <= 0) xn = x / 2;
		if (fabs(xn - x) <= 4e-16 * x) {
			x = xn;
			break;
		}
		// Halley steps shrink cubically; a step no smaller than half the last
		// one at the 1e-13 level is rounding noise of P (the device exp/log
		// can leave it a few ulps above 4e-16 forever): x is converged
		if (fabs(h) >= 0.5 * fabs(hprev) && fabs(h) <= 1e-13 * x) break;
		hprev = h;
		x = xn;
	}
	return x;
}

// Digamma psi(x) = d lgamma / dx for x > 0: the recurrence psi(x) = psi(x+1) - 1/x
// up to x >= 10, then the asymptotic series with Bernoulli coefficients
// (Abramowitz & Stegun 6.3.18), accurate to ~1e-15 there.
PK_HD double digamma(double x)
{
	if (!(x > 0)) return NAN;
	double acc = 0;
	while (x < 10) {
		acc -= 1 / x;
		x += 1;
	}
	const double r = 1 / (x * x);
	const double series =
	    r * (1. / 12 - r * (1. / 120 - r * (1. / 252 - r * (1. / 240 - r * (1. / 132 - r * (691. / 32760 - r / 12))))));
	return acc + log(x) - 0.5 / x - series;
}

// d bn / d nser of the Sersic constant bn(n) = qgamma(0.5, 2n)
// (src/sersic.cpp:297): Richardson-extrapolated central differences of
// qgamma, steps n/500 and n/1000 (truncation O(h^4), ~1e-12 relative).
PK_HD double sersic_dbn_dn(double n)
{
	const double h = n * 2e-3;
	const double d1 = (qgamma(0.5, 2 * (n + h)) - qgamma(0.5, 2 * (n - h))) / (2 * h);
	const double d2 = (qgamma(0.5, 2 * (n + h / 2)) - qgamma(0.5, 2 * (n - h / 2))) / h;
	return (4 * d2 - d1) / 3;
}

// ---- double-exponential quadrature ----------------------------------------
// libprofit integrates the luminosity of the king, brokenexp and coresersic
// profiles with QUADPACK (qags on [0, rt], qagi on [0, inf), tolerance 1e-4,
// src/utils.cpp:155-186).  Here: tanh-sinh on finite intervals and exp-sinh on
// [a, inf) (Takahasi & Mori 1974), which converge exponentially also for the
// integrable endpoint singularities these integrands have (coresersic behaves
// like r^(1-b) at 0, king like (rt-r)^a at rt), need no work arrays and run
// unchanged inside a device thread.  Levels halve the step until two
// successive estimates agree to 1e-14.

// The quadratures below spread their nodes over the lanes of a `Lanes`
// policy: SerialLanes (one thread: host code) or a wavefront (device set-up
// kernel, in pk_device.h).  lane()/count() select a lane's share of the
// nodes; sum() adds the lanes' partial sums and gives every lane the same
// total.
struct SerialLanes {
	PK_HD int lane() const { return 0; }
	PK_HD int count() const { return 1; }
	PK_HD double sum(double v) const { return v; }
	PK_HD double pick(double v, int) const { return v; }
};

// tanh-sinh on [a, b]: x = a + L q(t) / b - L q(t) with q = e/(1+e),
// e = exp(-pi sinh t), weight L pi cosh(t) e / (1+e)^2.
template <typename F, typename LN>
PK_HD double integrate_ts(const F &f, double a, double b, const LN &ln)
{
	const double L = b - a;
	if (!(L > 0)) return 0;
	const double pi = PK_PI;
	double h = 1.0;
	auto add_pair = [&](double t) -> double {
		const double e = exp(-pi * sinh(t));
		const double q = e / (1 + e);
		const double w = L * pi * cosh(t) * e / ((1 + e) * (1 + e));
		if (!(w > 1e-300) || !(q * L > 0)) return 0;
		double v = w * (f(a + L * q) + f(b - L * q));
		return isfinite(v) ? v : 0;
	};
	// level 0: t = 0 (weight L pi / 4) and the integers 1..7
	double part = ln.lane() == 0 ? 0.25 * L * pi * f(a + 0.5 * L) : 0.0;
	for (int k = 1 + ln.lane(); k <= 7; k += ln.count()) part += add_pair((double)k);
	double sum = ln.sum(part);
	double est = h * sum;
	for (int level = 1; level <= 12; level++) {
		h *= 0.5;
		// new nodes t = h (2i + 1) <= 7
		part = 0;
		for (int i = ln.lane();; i += ln.count()) {
			const double t = h * (2 * i + 1);
			if (t > 7.0) break;
			part += add_pair(t);
		}
		sum += ln.sum(part);
		double nest = h * sum;
		if (level >= 3 && fabs(nest - est) <= 1e-14 * fabs(nest)) return nest;
		est = nest;
	}
	return est;
}

// exp-sinh on [a, inf): x = a + exp(pi/2 sinh t), weight pi/2 cosh(t) (x - a).
template <typename F, typename LN>
PK_HD double integrate_es(const F &f, double a, const LN &ln)
{
	const double hp = 0.5 * PK_PI;
	auto term = [&](double t) -> double {
		const double u = hp * sinh(t);
		if (u > 700 || u < -700) return 0;
		const double d = exp(u);
		double v = hp * cosh(t) * d * f(a + d);
		return isfinite(v) ? v : 0;
	};
	double h = 1.0;
	double part = ln.lane() == 0 ? term(0.0) : 0.0;
	for (int k = 1 + ln.lane(); k <= 6; k += ln.count()) part += term((double)k) + term(-(double)k);
	double sum = ln.sum(part);
	double est = h * sum;
	for (int level = 1; level <= 12; level++) {
		h *= 0.5;
		part = 0;
		for (int i = ln.lane();; i += ln.count()) {
			const double t = h * (2 * i + 1);
			if (t > 6.0) break;
			part += term(t) + term(-t);
		}
		sum += ln.sum(part);
		double nest = h * sum;
		if (level >= 3 && fabs(nest - est) <= 1e-14 * fabs(nest)) return nest;
		est = nest;
	}
	return est;
}

template <typename F, typename LN = SerialLanes>
PK_HD double integrate_adaptive(const F &f, double a, double b, const LN &ln = LN())
{
	return integrate_ts(f, a, b, ln);
}

template <typename F, typename LN = SerialLanes>
PK_HD double integrate_to_inf(const F &f, double a, const LN &ln = LN())
{
	return integrate_es(f, a, ln);
}

}  // namespace pk
