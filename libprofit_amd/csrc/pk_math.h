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
struct QagWork;
struct SerialLanes {
	QagWork *ws = nullptr;  // interval list of the QUADPACK integrator
	PK_HD int lane() const { return 0; }
	PK_HD int count() const { return 1; }
	PK_HD double sum(double v) const { return v; }
	PK_HD double pick(double v, int) const { return v; }
	PK_HD QagWork *work() const { return ws; }
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

// ---- QUADPACK qags / qagiu as GSL runs them --------------------------------
// libprofit's luminosity integrals call gsl_integration_qags on [a, b] and
// gsl_integration_qagiu on [a, inf) with epsabs = epsrel = 1e-4 and a limit
// of 100 intervals (src/utils.cpp:155-186).  The algorithm is QUADPACK's
// dqagse (Piessens, de Doncker-Kapenga, Ueberhuber, Kahaner 1983): a
// Gauss-Kronrod rule (21 points; 15 points on the transformed integrand
// f(a + (1-t)/t)/t^2 over (0, 1] for qagiu), bisection of the interval with
// the largest error estimate, the error rescaling
// min(1, (200 err/resasc)^1.5) resasc, and Wynn's epsilon algorithm over the
// sequence of area estimates.  It follows GSL's sequence of operations, so
// the result carries the same quadrature error libprofit's does.
//
// The rule's nodes are spread over the lanes of the Lanes policy (on the
// device: the 21 nodes of each of the two halves of a bisected interval in
// lanes 0-20 and 32-52) and gathered back, so every lane forms the sums in
// GSL's order and runs the same bookkeeping on the same values.  The
// interval list lives in a QagWork the policy provides (LDS on the device).

constexpr int QAG_LIMIT = 100;

struct QagWork {
	double alist[QAG_LIMIT], blist[QAG_LIMIT], rlist[QAG_LIMIT], elist[QAG_LIMIT];
	int32_t order[QAG_LIMIT], level[QAG_LIMIT];
	double epstab[52], res3la[3];
};

// Kronrod nodes, Gauss weights and Kronrod weights of the 21- and 15-point
// rules (published constants)
template <int NPTS>
struct GkRule;
template <>
struct GkRule<21> {
	static constexpr int nk = 11;
	PK_HD static double x(int i)
	{
		constexpr double v[11] = {0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
		                          0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
		                          0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
		                          0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
		                          0.294392862701460198131126603103866, 0.148874338981631210884826001129720,
		                          0.000000000000000000000000000000000};
		return v[i];
	}
	PK_HD static double wg(int i)
	{
		constexpr double v[5] = {0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
		                         0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
		                         0.295524224714752870173892994651338};
		return v[i];
	}
	PK_HD static double wk(int i)
	{
		constexpr double v[11] = {0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
		                          0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
		                          0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
		                          0.123491976262065851077958109831074, 0.134709217311473325928054001771707,
		                          0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
		                          0.149445554002916905664936468389821};
		return v[i];
	}
};
template <>
struct GkRule<15> {
	static constexpr int nk = 8;
	PK_HD static double x(int i)
	{
		constexpr double v[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
		                         0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
		                         0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
		                         0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
		return v[i];
	}
	PK_HD static double wg(int i)
	{
		constexpr double v[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
		                         0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
		return v[i];
	}
	PK_HD static double wk(int i)
	{
		constexpr double v[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
		                         0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
		                         0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
		                         0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
		return v[i];
	}
};

constexpr double QAG_EPS = 2.2204460492503131e-16;
constexpr double QAG_DMIN = 2.2250738585072014e-308;
constexpr double QAG_DMAX = 1.7976931348623157e+308;

PK_HD double qag_rescale_error(double err, double result_abs, double result_asc)
{
	err = fabs(err);
	if (result_asc != 0 && err != 0) {
		double scale = pow(200 * err / result_asc, 1.5);
		err = scale < 1 ? result_asc * scale : result_asc;
	}
	if (result_abs > QAG_DMIN / (50 * QAG_EPS)) {
		double min_err = 50 * QAG_EPS * result_abs;
		if (min_err > err) err = min_err;
	}
	return err;
}

// abscissa of node j of the rule on an interval: j = 0 the centre, j = 2k+1
// centre - h x_k, j = 2k+2 centre + h x_k
template <int NPTS>
PK_HD double qk_node(double center, double half_length, int j)
{
#pragma clang fp contract(off)
	if (j == 0) return center;
	const double abscissa = half_length * GkRule<NPTS>::x((j - 1) >> 1);
	return (j & 1) ? center - abscissa : center + abscissa;
}

struct QkOut {
	double result, abserr, resabs, resasc;
};

// the rule's sums from its node values fv[j] (node order of qk_node), in
// the order of GSL's gsl_integration_qk
template <int NPTS>
PK_HD QkOut qk_sums(const double *fv, double half_length)
{
#pragma clang fp contract(off)
	using R = GkRule<NPTS>;
	constexpr int n = R::nk;
	const double f_center = fv[0];
	double result_gauss = 0, result_kronrod = f_center * R::wk(n - 1);
	double result_abs = fabs(result_kronrod);
	if (n % 2 == 0) result_gauss = f_center * R::wg(n / 2 - 1);
	for (int j = 0; j < (n - 1) / 2; j++) {
		const int jtw = j * 2 + 1;
		const double fval1 = fv[2 * jtw + 1], fval2 = fv[2 * jtw + 2];
		const double fsum = fval1 + fval2;
		result_gauss += R::wg(j) * fsum;
		result_kronrod += R::wk(jtw) * fsum;
		result_abs += R::wk(jtw) * (fabs(fval1) + fabs(fval2));
	}
	for (int j = 0; j < n / 2; j++) {
		const int jtwm1 = j * 2;
		const double fval1 = fv[2 * jtwm1 + 1], fval2 = fv[2 * jtwm1 + 2];
		result_kronrod += R::wk(jtwm1) * (fval1 + fval2);
		result_abs += R::wk(jtwm1) * (fabs(fval1) + fabs(fval2));
	}
	const double mean = result_kronrod * 0.5;
	double result_asc = R::wk(n - 1) * fabs(f_center - mean);
	for (int j = 0; j < n - 1; j++) result_asc += R::wk(j) * (fabs(fv[2 * j + 1] - mean) + fabs(fv[2 * j + 2] - mean));
	const double err = (result_kronrod - result_gauss) * half_length;
	const double ahl = fabs(half_length);
	QkOut o;
	o.result = result_kronrod * half_length;
	o.resabs = result_abs * ahl;
	o.resasc = result_asc * ahl;
	o.abserr = qag_rescale_error(err, o.resabs, o.resasc);
	return o;
}

// the rule on `nint` (1 or 2) intervals [a_i, b_i] at once
template <int NPTS, typename F, typename LN>
PK_HD void qk_intervals(const F &f, const double *a, const double *b, int nint, QkOut *out, const LN &ln)
{
	double fv[2][NPTS];
	double center[2], half[2];
	for (int i = 0; i < nint; i++) {
		center[i] = 0.5 * (a[i] + b[i]);
		half[i] = 0.5 * (b[i] - a[i]);
	}
	if (ln.count() >= 64) {
		const int l = ln.lane(), i = l >> 5, j = l & 31;
		double v = 0;
		if (i < nint && j < NPTS) v = f(qk_node<NPTS>(center[i], half[i], j));
		for (int ii = 0; ii < nint; ii++)
			for (int jj = 0; jj < NPTS; jj++) fv[ii][jj] = ln.pick(v, ii * 32 + jj);
	} else {
		for (int ii = 0; ii < nint; ii++)
			for (int jj = 0; jj < NPTS; jj++) fv[ii][jj] = f(qk_node<NPTS>(center[ii], half[ii], jj));
	}
	for (int i = 0; i < nint; i++) out[i] = qk_sums<NPTS>(fv[i], half[i]);
}

struct QagList {
	QagWork *w;
	int limit, size, nrmax, i, maximum_level;
};

// keeps order[] descending in error over the intervals that can still be
// bisected (QUADPACK dqpsrt)
PK_HD void qag_sort(QagList &L)
{
	QagWork &w = *L.w;
	const int last = L.size - 1, limit = L.limit;
	int i_nrmax = L.nrmax, i_maxerr = w.order[i_nrmax];
	if (last < 2) {
		w.order[0] = 0;
		w.order[1] = 1;
		L.i = i_maxerr;
		return;
	}
	const double errmax = w.elist[i_maxerr];
	while (i_nrmax > 0 && errmax > w.elist[w.order[i_nrmax - 1]]) {
		w.order[i_nrmax] = w.order[i_nrmax - 1];
		i_nrmax--;
	}
	const int top = last < (limit / 2 + 2) ? last : limit - last + 1;
	int i = i_nrmax + 1;
	while (i < top && errmax < w.elist[w.order[i]]) {
		w.order[i - 1] = w.order[i];
		i++;
	}
	w.order[i - 1] = i_maxerr;
	const double errmin = w.elist[last];
	int k = top - 1;
	while (k > i - 2 && errmin >= w.elist[w.order[k]]) {
		w.order[k + 1] = w.order[k];
		k--;
	}
	w.order[k + 1] = last;
	L.i = w.order[i_nrmax];
	L.nrmax = i_nrmax;
}

PK_HD void qag_update(QagList &L, double a1, double b1, const QkOut &q1, double a2, double b2, const QkOut &q2)
{
	QagWork &w = *L.w;
	const int i_max = L.i, i_new = L.size, new_level = w.level[i_max] + 1;
	if (q2.abserr > q1.abserr) {
		w.alist[i_max] = a2;
		w.rlist[i_max] = q2.result;
		w.elist[i_max] = q2.abserr;
		w.level[i_max] = new_level;
		w.alist[i_new] = a1;
		w.blist[i_new] = b1;
		w.rlist[i_new] = q1.result;
		w.elist[i_new] = q1.abserr;
		w.level[i_new] = new_level;
	} else {
		w.blist[i_max] = b1;
		w.rlist[i_max] = q1.result;
		w.elist[i_max] = q1.abserr;
		w.level[i_max] = new_level;
		w.alist[i_new] = a2;
		w.blist[i_new] = b2;
		w.rlist[i_new] = q2.result;
		w.elist[i_new] = q2.abserr;
		w.level[i_new] = new_level;
	}
	L.size++;
	if (new_level > L.maximum_level) L.maximum_level = new_level;
	qag_sort(L);
}

PK_HD bool qag_increase_nrmax(QagList &L)
{
	const int last = L.size - 1;
	const int jupbnd = last > 1 + L.limit / 2 ? L.limit + 1 - last : last;
	for (int k = L.nrmax; k <= jupbnd; k++) {
		const int i_max = L.w->order[L.nrmax];
		L.i = i_max;
		if (L.w->level[i_max] < L.maximum_level) return true;
		L.nrmax++;
	}
	return false;
}

// Wynn's epsilon algorithm (QUADPACK dqelg as GSL places its nres update)
struct QagTable {
	int n, nres;
};

PK_HD void qag_qelg(QagTable &t, QagWork &w, double *result, double *abserr)
{
#pragma clang fp contract(off)
	double *epstab = w.epstab, *res3la = w.res3la;
	const int n = t.n - 1;
	const double current = epstab[n];
	double absolute = QAG_DMAX, relative = 5 * QAG_EPS * fabs(current);
	const int newelm = n / 2, n_orig = n, nres_orig = t.nres;
	int n_final = n;
	*result = current;
	*abserr = QAG_DMAX;
	if (n < 2) {
		*abserr = fmax(absolute, relative);
		return;
	}
	epstab[n + 2] = epstab[n];
	epstab[n] = QAG_DMAX;
	for (int i = 0; i < newelm; i++) {
		double res = epstab[n - 2 * i + 2];
		const double e0 = epstab[n - 2 * i - 2], e1 = epstab[n - 2 * i - 1], e2 = res;
		const double e1abs = fabs(e1), delta2 = e2 - e1, err2 = fabs(delta2);
		const double tol2 = fmax(fabs(e2), e1abs) * QAG_EPS;
		const double delta3 = e1 - e0, err3 = fabs(delta3), tol3 = fmax(e1abs, fabs(e0)) * QAG_EPS;
		if (err2 <= tol2 && err3 <= tol3) {
			*result = res;
			absolute = err2 + err3;
			relative = 5 * QAG_EPS * fabs(res);
			*abserr = fmax(absolute, relative);
			return;
		}
		const double e3 = epstab[n - 2 * i];
		epstab[n - 2 * i] = e1;
		const double delta1 = e1 - e3, err1 = fabs(delta1), tol1 = fmax(e1abs, fabs(e3)) * QAG_EPS;
		if (err1 <= tol1 || err2 <= tol2 || err3 <= tol3) {
			n_final = 2 * i;
			break;
		}
		const double ss = (1 / delta1 + 1 / delta2) - 1 / delta3;
		if (fabs(ss * e1) <= 0.0001) {
			n_final = 2 * i;
			break;
		}
		res = e1 + 1 / ss;
		epstab[n - 2 * i] = res;
		const double error = err2 + fabs(res - e2) + err3;
		if (error <= *abserr) {
			*abserr = error;
			*result = res;
		}
	}
	if (n_final == 49) n_final = 2 * (49 / 2);
	if (n_orig % 2 == 1) {
		for (int i = 0; i <= newelm; i++) epstab[1 + i * 2] = epstab[i * 2 + 3];
	} else {
		for (int i = 0; i <= newelm; i++) epstab[i * 2] = epstab[i * 2 + 2];
	}
	if (n_orig != n_final)
		for (int i = 0; i <= n_final; i++) epstab[i] = epstab[n_orig - n_final + i];
	t.n = n_final + 1;
	if (nres_orig < 3) {
		res3la[nres_orig] = *result;
		*abserr = QAG_DMAX;
	} else {
		*abserr = fabs(*result - res3la[2]) + fabs(*result - res3la[1]) + fabs(*result - res3la[0]);
		res3la[0] = res3la[1];
		res3la[1] = res3la[2];
		res3la[2] = *result;
	}
	t.nres = nres_orig + 1;
	*abserr = fmax(*abserr, 5 * QAG_EPS * fabs(*result));
}

// qags with the NPTS-point rule on [a, b]; returns the result GSL returns
template <int NPTS, typename F, typename LN>
PK_HD double qags(const F &f, double a, double b, double epsabs, double epsrel, int limit, QagWork &w,
                  const LN &ln, double *abserr_out = nullptr)
{
#pragma clang fp contract(off)
	if (limit > QAG_LIMIT) limit = QAG_LIMIT;
	QagList L{&w, limit, 0, 0, 0, 0};
	w.alist[0] = a;
	w.blist[0] = b;
	w.rlist[0] = 0;
	w.elist[0] = 0;
	w.order[0] = 0;
	w.level[0] = 0;
	double abserr_dummy;
	double &abserr = abserr_out ? *abserr_out : abserr_dummy;
	if (epsabs <= 0 && (epsrel < 50 * QAG_EPS || epsrel < 0.5e-28)) {
		abserr = 0;
		return 0;
	}
	QkOut q0;
	qk_intervals<NPTS>(f, &a, &b, 1, &q0, ln);
	L.size = 1;
	w.rlist[0] = q0.result;
	w.elist[0] = q0.abserr;
	double tolerance = fmax(epsabs, epsrel * fabs(q0.result));
	if ((q0.abserr <= 100 * QAG_EPS * q0.resabs && q0.abserr > tolerance) ||
	    (q0.abserr <= tolerance && q0.abserr != q0.resasc) || q0.abserr == 0.0 || limit == 1) {
		abserr = q0.abserr;
		return q0.result;
	}
	QagTable table{0, 0};
	w.epstab[table.n++] = q0.result;
	double area = q0.result, errsum = q0.abserr, res_ext = q0.result, err_ext = QAG_DMAX;
	double ertest = 0, error_over_large_intervals = 0, reseps = 0, abseps = 0, correc = 0;
	int ktmin = 0, iteration = 1;
	int roundoff_type1 = 0, roundoff_type2 = 0, roundoff_type3 = 0, error_type = 0, error_type2 = 0;
	const bool positive_integrand = fabs(q0.result) >= (1 - 50 * QAG_EPS) * q0.resabs;
	bool extrapolate = false, disallow_extrapolation = false, sum_list = false;
	do {
		const double a_i = w.alist[L.i], b_i = w.blist[L.i], r_i = w.rlist[L.i], e_i = w.elist[L.i];
		const int current_level = w.level[L.i] + 1;
		const double a1 = a_i, b1 = 0.5 * (a_i + b_i), a2 = b1, b2 = b_i;
		iteration++;
		QkOut q[2];
		const double aa[2] = {a1, a2}, bb[2] = {b1, b2};
		qk_intervals<NPTS>(f, aa, bb, 2, q, ln);
		const double area12 = q[0].result + q[1].result, error12 = q[0].abserr + q[1].abserr;
		const double last_e_i = e_i;
		errsum = errsum + error12 - e_i;
		area = area + area12 - r_i;
		tolerance = fmax(epsabs, epsrel * fabs(area));
		if (q[0].resasc != q[0].abserr && q[1].resasc != q[1].abserr) {
			const double delta = r_i - area12;
			if (fabs(delta) <= 1.0e-5 * fabs(area12) && error12 >= 0.99 * e_i) {
				if (!extrapolate)
					roundoff_type1++;
				else
					roundoff_type2++;
			}
			if (iteration > 10 && error12 > e_i) roundoff_type3++;
		}
		if (roundoff_type1 + roundoff_type2 >= 10 || roundoff_type3 >= 20) error_type = 2;
		if (roundoff_type2 >= 5) error_type2 = 1;
		{
			const double tmp = (1 + 100 * QAG_EPS) * (fabs(a2) + 1000 * QAG_DMIN);
			if (fabs(a1) <= tmp && fabs(b2) <= tmp) error_type = 4;
		}
		qag_update(L, a1, b1, q[0], a2, b2, q[1]);
		if (errsum <= tolerance) {
			sum_list = true;
			break;
		}
		if (error_type) break;
		if (iteration >= limit - 1) {
			error_type = 1;
			break;
		}
		if (iteration == 2) {
			error_over_large_intervals = errsum;
			ertest = tolerance;
			w.epstab[table.n++] = area;
			continue;
		}
		if (disallow_extrapolation) continue;
		error_over_large_intervals += -last_e_i;
		if (current_level < L.maximum_level) error_over_large_intervals += error12;
		if (!extrapolate) {
			if (w.level[L.i] < L.maximum_level) continue;
			extrapolate = true;
			L.nrmax = 1;
		}
		if (!error_type2 && error_over_large_intervals > ertest)
			if (qag_increase_nrmax(L)) continue;
		w.epstab[table.n++] = area;
		qag_qelg(table, w, &reseps, &abseps);
		ktmin++;
		if (ktmin > 5 && err_ext < 0.001 * errsum) error_type = 5;
		if (abseps < err_ext) {
			ktmin = 0;
			err_ext = abseps;
			res_ext = reseps;
			correc = error_over_large_intervals;
			ertest = fmax(epsabs, epsrel * fabs(reseps));
			if (err_ext <= ertest) break;
		}
		if (table.n == 1) disallow_extrapolation = true;
		if (error_type == 5) break;
		L.nrmax = 0;
		L.i = w.order[0];
		extrapolate = false;
		error_over_large_intervals = errsum;
	} while (iteration < limit);

	if (!sum_list) {
		if (err_ext == QAG_DMAX) {
			sum_list = true;
		} else {
			if (error_type || error_type2) {
				if (error_type2) err_ext += correc;
				if (res_ext != 0.0 && area != 0.0) {
					if (err_ext / fabs(res_ext) > errsum / fabs(area)) sum_list = true;
				} else if (err_ext > errsum) {
					sum_list = true;
				}
			}
		}
	}
	if (sum_list) {
		double sum = 0;
		for (int k = 0; k < L.size; k++) sum += w.rlist[k];
		abserr = errsum;
		return sum;
	}
	abserr = err_ext;
	return res_ext;
}

// qagiu: qags with the 15-point rule on x = a + (1 - t) / t over (0, 1]
template <typename F, typename LN>
PK_HD double qagiu(const F &f, double a, double epsabs, double epsrel, int limit, QagWork &w, const LN &ln,
                   double *abserr_out = nullptr)
{
	auto g = [&](double t) -> double {
#pragma clang fp contract(off)
		const double x = a + (1 - t) / t;
		const double y = f(x);
		return (y / t) / t;
	};
	return qags<15>(g, 0.0, 1.0, epsabs, epsrel, limit, w, ln, abserr_out);
}

// The luminosity integrals with the integrator `mode`
// (PK_INTEGRATOR_QUADPACK: libprofit's qags / qagiu at 1e-4, limit 100;
// PK_INTEGRATOR_EXACT: double-exponential quadrature to 1e-14).
template <typename F, typename LN = SerialLanes>
PK_HD double integrate_adaptive(const F &f, double a, double b, int mode, const LN &ln = LN())
{
	if (mode == 1) return integrate_ts(f, a, b, ln);
	return qags<21>(f, a, b, 1e-4, 1e-4, 100, *ln.work(), ln);
}

template <typename F, typename LN = SerialLanes>
PK_HD double integrate_to_inf(const F &f, double a, int mode, const LN &ln = LN())
{
	if (mode == 1) return integrate_es(f, a, ln);
	return qagiu(f, a, 1e-4, 1e-4, 100, *ln.work(), ln);
}

}  // namespace pk
