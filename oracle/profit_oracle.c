/*
 * profit_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * An FP64 CPU restatement of libprofit 1.9.3's image-generation path, used as
 * the parity oracle of the HIP implementation and as the CPU baseline of the
 * benchmark.  Nothing in the product links, loads or calls this file; only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg do.
 *
 * It follows the reference's CPU semantics function by function (citations
 * are reference file:line):
 *   - special functions: own implementations of the math back-end entry points
 *     qgamma/pgamma/gammafn/beta/integrate_qag{s,i} (src/utils.cpp:101-259),
 *     which libprofit takes from GSL or Rmath (absent here);
 *   - radial profiles: src/radial.cpp:47-344, src/sersic.cpp:79-360,
 *     src/moffat.cpp, src/ferrer.cpp, src/king.cpp,
 *     src/brokenexponential.cpp, src/coresersic.cpp;
 *   - sky / psf / null: src/sky.cpp:42-63, src/psf.cpp:58-145,
 *     include/profit/null.h;
 *   - Model::evaluate: src/model.cpp:152-360;
 *   - convolvers: src/convolve.cpp:55-312 (BRUTE_OLD, BRUTE with the
 *     summation order of dot_product<AVX>, include/profit/dot_product.h:
 *     272-384, FFT with an own radix-2 transform in place of FFTW).
 *
 * Compiled with -ffp-contract=off so that products and sums round like the
 * reference's default x86-64 build (no FMA contraction).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/profit_hip.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static int or_threads = 1;
static char or_err[512];

void or_set_threads(int n) { or_threads = n < 1 ? 1 : n; }
const char *or_last_error(void) { return or_err; }

static int fail(int code, const char *msg)
{
	snprintf(or_err, sizeof(or_err), "%s", msg);
	return code;
}

/* ======================================================================== */
/* Special functions (src/utils.cpp:101-259 semantics)                       */
/* ======================================================================== */

/* gsl_sf_gamma_e semantics as wrapped by src/utils.cpp:113-129: NaN at the
 * poles, 0 on underflow, +Inf on overflow for x > 0. */
double or_gammafn(double x)
{
	if (isnan(x)) return NAN;
	if (x <= 0 && x == floor(x)) return NAN;
	double g = tgamma(x);
	if (isinf(g)) return x > 0 ? INFINITY : NAN;
	return g;
}

/* R semantics for beta (include/profit/utils.h:79-92, src/utils.cpp:131-151). */
double or_beta(double a, double b)
{
	if (a < 0 || b < 0) return NAN;
	if (a == 0 || b == 0) return INFINITY;
	if (a + b < 170.0) {
		return tgamma(a) * tgamma(b) / tgamma(a + b);
	}
	return exp(lgamma(a) + lgamma(b) - lgamma(a + b));
}

/* log of x^a e^-x / Gamma(a) */
static double or_lprefactor(double a, double x)
{
	return a * log(x) - x - lgamma(a);
}

/* Regularised incomplete gamma: P(a, x) (upper = 0) or Q(a, x) = 1 - P
 * (upper = 1), each computed from the representation in which it is not the
 * difference of two nearly equal numbers: power series below x = a + 1,
 * continued fraction (modified Lentz) above. */
static double or_incgamma(double x, double a, int upper)
{
	if (isnan(x) || isnan(a) || a <= 0) return NAN;
	if (x <= 0) return upper ? 1 : 0;
	if (isinf(x)) return upper ? 0 : 1;
	if (x < a + 1) {
		double term = 1.0 / a, sum = term;
		for (int n = 1; n < 100000; n++) {
			term *= x / (a + n);
			sum += term;
			if (fabs(term) < fabs(sum) * 1e-17) break;
		}
		double pv = sum * exp(or_lprefactor(a, x));
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
	double q = exp(or_lprefactor(a, x)) * h;
	return upper ? q : 1 - q;
}

/* gsl_cdf_gamma_P(x, a, 1) */
double or_pgamma(double x, double a) { return or_incgamma(x, a, 0); }

/* Standard normal quantile, used only as a starting point (Acklam's rational
 * approximation, relative error ~1e-9, refined by the Halley steps below). */
static double or_norm_quantile(double p)
{
	static const double a[] = {-3.969683028665376e+01, 2.209460984245205e+02,
	                           -2.759285104469687e+02, 1.383577518672690e+02,
	                           -3.066479806614716e+01, 2.506628277459239e+00};
	static const double b[] = {-5.447609879822406e+01, 1.615858368580409e+02,
	                           -1.556989798598866e+02, 6.680131188771972e+01,
	                           -1.328068155288572e+01};
	static const double c[] = {-7.784894002430293e-03, -3.223964580411365e-01,
	                           -2.400758277161838e+00, -2.549732539343734e+00,
	                           4.374664141464968e+00, 2.938163982698783e+00};
	static const double d[] = {7.784695709041462e-03, 3.224671290700398e-01,
	                           2.445134137142996e+00, 3.754408661907416e+00};
	double q, r;
	if (p < 0.02425) {
		q = sqrt(-2 * log(p));
		return (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
		       ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
	}
	if (p > 1 - 0.02425) {
		q = sqrt(-2 * log(1 - p));
		return -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
		       ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
	}
	q = p - 0.5;
	r = q * q;
	return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
	       (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1);
}

/* Quantile of the unit-scale gamma distribution (gsl_cdf_gamma_Pinv(p, a, 1)):
 * the x with P(a, x) = p, by Halley iteration on P. */
double or_qgamma(double p, double a)
{
	if (isnan(p) || isnan(a) || a <= 0 || p < 0 || p > 1) return NAN;
	if (p == 0) return 0;
	if (p == 1) return INFINITY;
	double x;
	if (a < 1) {
		x = pow(p * tgamma(a + 1), 1 / a);
		if (!(x > 0)) x = 1e-300;
	} else {
		double z = or_norm_quantile(p);
		double t = 1 / (9 * a);
		x = a * pow(1 - t + z * sqrt(t), 3);
		if (!(x > 0)) x = a * 1e-3;
	}
	double lg = lgamma(a);
	const int upper = p > 0.5;
	const double q = 1 - p; /* exact for p > 0.5 */
	for (int it = 0; it < 200; it++) {
		double f = upper ? q - or_incgamma(x, a, 1) : or_incgamma(x, a, 0) - p;
		double lpdf = (a - 1) * log(x) - x - lg;
		double pdf = exp(lpdf);
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
		x = xn;
	}
	return x;
}

/* --- QUADPACK integration as GSL runs it --------------------------------- */
/*
 * libprofit integrates the king, brokenexp and coresersic luminosities with
 * gsl_integration_qags (finite interval) and gsl_integration_qagiu ([a, inf)),
 * epsabs = epsrel = 1e-4, limit 100 (src/utils.cpp:155-186).  GSL is absent
 * here, so this is a restatement of the published algorithm: QUADPACK's
 * dqagse (Piessens, de Doncker-Kapenga, Ueberhuber, Kahaner 1983) as GSL
 * 1.16 implements it in integration/qags.c -- 21-point Gauss-Kronrod rule on
 * finite intervals, the 15-point rule on the transformed integrand
 * f(a + (1-t)/t) / t^2 over (0, 1] for [a, inf), bisection of the interval
 * with the largest error, Wynn's epsilon algorithm over the sequence of
 * area estimates (qelg), GSL's error rescaling and roundoff tests.  It is
 * pinned to GSL's own test-suite results (tests/test_oracle_quadpack.py).
 */
typedef double (*or_func)(double, void *);

/* Kronrod nodes (xgk), Gauss weights (wg) and Kronrod weights (wgk) of the
 * 21- and 15-point rules (published constants; the tests check their
 * exactness on polynomials). */
static const double or_xgk21[11] = {
	0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
	0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
	0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
	0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
	0.294392862701460198131126603103866, 0.148874338981631210884826001129720,
	0.000000000000000000000000000000000};
static const double or_wg21[5] = {
	0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
	0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
	0.295524224714752870173892994651338};
static const double or_wgk21[11] = {
	0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
	0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
	0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
	0.123491976262065851077958109831074, 0.134709217311473325928054001771707,
	0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
	0.149445554002916905664936468389821};
static const double or_xgk15[8] = {
	0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
	0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
	0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
	0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double or_wg15[4] = {
	0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
	0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
static const double or_wgk15[8] = {
	0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
	0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
	0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
	0.204432940075298892414161999234649, 0.209482141084727828012999174891714};

#define OR_EPS 2.2204460492503131e-16
#define OR_DMIN 2.2250738585072014e-308
#define OR_DMAX 1.7976931348623157e+308

static double or_rescale_error(double err, double result_abs, double result_asc)
{
	err = fabs(err);
	if (result_asc != 0 && err != 0) {
		double scale = pow(200 * err / result_asc, 1.5);
		err = scale < 1 ? result_asc * scale : result_asc;
	}
	if (result_abs > OR_DMIN / (50 * OR_EPS)) {
		double min_err = 50 * OR_EPS * result_abs;
		if (min_err > err) err = min_err;
	}
	return err;
}

/* one Gauss-Kronrod rule with n Kronrod nodes per half (n = 11: 21 points,
 * n = 8: 15 points); sums in the order of GSL's gsl_integration_qk */
static void or_qk(int n, const double *xgk, const double *wg, const double *wgk, or_func f, void *ctx, double a,
                  double b, double *result, double *abserr, double *resabs, double *resasc)
{
	double fv1[11], fv2[11];
	const double center = 0.5 * (a + b);
	const double half_length = 0.5 * (b - a);
	const double abs_half_length = fabs(half_length);
	const double f_center = f(center, ctx);
	double result_gauss = 0, result_kronrod = f_center * wgk[n - 1];
	double result_abs = fabs(result_kronrod), result_asc, mean, err;
	if (n % 2 == 0) result_gauss = f_center * wg[n / 2 - 1];
	for (int j = 0; j < (n - 1) / 2; j++) {
		const int jtw = j * 2 + 1;
		const double abscissa = half_length * xgk[jtw];
		const double fval1 = f(center - abscissa, ctx), fval2 = f(center + abscissa, ctx);
		const double fsum = fval1 + fval2;
		fv1[jtw] = fval1;
		fv2[jtw] = fval2;
		result_gauss += wg[j] * fsum;
		result_kronrod += wgk[jtw] * fsum;
		result_abs += wgk[jtw] * (fabs(fval1) + fabs(fval2));
	}
	for (int j = 0; j < n / 2; j++) {
		const int jtwm1 = j * 2;
		const double abscissa = half_length * xgk[jtwm1];
		const double fval1 = f(center - abscissa, ctx), fval2 = f(center + abscissa, ctx);
		fv1[jtwm1] = fval1;
		fv2[jtwm1] = fval2;
		result_kronrod += wgk[jtwm1] * (fval1 + fval2);
		result_abs += wgk[jtwm1] * (fabs(fval1) + fabs(fval2));
	}
	mean = result_kronrod * 0.5;
	result_asc = wgk[n - 1] * fabs(f_center - mean);
	for (int j = 0; j < n - 1; j++) result_asc += wgk[j] * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
	err = (result_kronrod - result_gauss) * half_length;
	*result = result_kronrod * half_length;
	*resabs = result_abs * abs_half_length;
	*resasc = result_asc * abs_half_length;
	*abserr = or_rescale_error(err, *resabs, *resasc);
}

/* the interval list of a qags run (GSL's gsl_integration_workspace) */
#define OR_QLIMIT 1000
typedef struct {
	size_t limit, size, nrmax, i, maximum_level;
	double alist[OR_QLIMIT], blist[OR_QLIMIT], rlist[OR_QLIMIT], elist[OR_QLIMIT];
	size_t order[OR_QLIMIT], level[OR_QLIMIT];
} or_qws;

/* keeps order[] descending in error for the intervals that can still be
 * bisected (QUADPACK dqpsrt) */
static void or_qpsrt(or_qws *w)
{
	const size_t last = w->size - 1, limit = w->limit;
	double *elist = w->elist;
	size_t *order = w->order;
	size_t i_nrmax = w->nrmax, i_maxerr = order[i_nrmax];
	double errmax, errmin;
	long i, k, top;
	if (last < 2) {
		order[0] = 0;
		order[1] = 1;
		w->i = i_maxerr;
		return;
	}
	errmax = elist[i_maxerr];
	while (i_nrmax > 0 && errmax > elist[order[i_nrmax - 1]]) {
		order[i_nrmax] = order[i_nrmax - 1];
		i_nrmax--;
	}
	top = last < (limit / 2 + 2) ? (long)last : (long)(limit - last + 1);
	i = (long)i_nrmax + 1;
	while (i < top && errmax < elist[order[i]]) {
		order[i - 1] = order[i];
		i++;
	}
	order[i - 1] = i_maxerr;
	errmin = elist[last];
	k = top - 1;
	while (k > i - 2 && errmin >= elist[order[k]]) {
		order[k + 1] = order[k];
		k--;
	}
	order[k + 1] = last;
	w->i = order[i_nrmax];
	w->nrmax = i_nrmax;
}

static void or_qupdate(or_qws *w, double a1, double b1, double area1, double error1, double a2, double b2,
                       double area2, double error2)
{
	const size_t i_max = w->i, i_new = w->size, new_level = w->level[i_max] + 1;
	if (error2 > error1) {
		w->alist[i_max] = a2;
		w->rlist[i_max] = area2;
		w->elist[i_max] = error2;
		w->level[i_max] = new_level;
		w->alist[i_new] = a1;
		w->blist[i_new] = b1;
		w->rlist[i_new] = area1;
		w->elist[i_new] = error1;
		w->level[i_new] = new_level;
	} else {
		w->blist[i_max] = b1;
		w->rlist[i_max] = area1;
		w->elist[i_max] = error1;
		w->level[i_max] = new_level;
		w->alist[i_new] = a2;
		w->blist[i_new] = b2;
		w->rlist[i_new] = area2;
		w->elist[i_new] = error2;
		w->level[i_new] = new_level;
	}
	w->size++;
	if (new_level > w->maximum_level) w->maximum_level = new_level;
	or_qpsrt(w);
}

static int or_increase_nrmax(or_qws *w)
{
	const size_t last = w->size - 1;
	const long jupbnd = last > (1 + w->limit / 2) ? (long)(w->limit + 1 - last) : (long)last;
	for (long k = (long)w->nrmax; k <= jupbnd; k++) {
		size_t i_max = w->order[w->nrmax];
		w->i = i_max;
		if (w->level[i_max] < w->maximum_level) return 1;
		w->nrmax++;
	}
	return 0;
}

/* Wynn's epsilon algorithm over the sequence of area estimates (QUADPACK
 * dqelg, with GSL's placement of the nres update) */
typedef struct {
	size_t n, nres;
	double rlist2[52], res3la[3];
} or_qtab;

static void or_qelg(or_qtab *t, double *result, double *abserr)
{
	double *epstab = t->rlist2, *res3la = t->res3la;
	const size_t n = t->n - 1;
	const double current = epstab[n];
	double absolute = OR_DMAX, relative = 5 * OR_EPS * fabs(current);
	const size_t newelm = n / 2, n_orig = n, nres_orig = t->nres;
	size_t n_final = n, i;
	*result = current;
	*abserr = OR_DMAX;
	if (n < 2) {
		*result = current;
		*abserr = fmax(absolute, relative);
		return;
	}
	epstab[n + 2] = epstab[n];
	epstab[n] = OR_DMAX;
	for (i = 0; i < newelm; i++) {
		double res = epstab[n - 2 * i + 2];
		double e0 = epstab[n - 2 * i - 2], e1 = epstab[n - 2 * i - 1], e2 = res;
		double e1abs = fabs(e1), delta2 = e2 - e1, err2 = fabs(delta2);
		double tol2 = fmax(fabs(e2), e1abs) * OR_EPS;
		double delta3 = e1 - e0, err3 = fabs(delta3), tol3 = fmax(e1abs, fabs(e0)) * OR_EPS;
		double e3, delta1, err1, tol1, ss;
		if (err2 <= tol2 && err3 <= tol3) {
			*result = res;
			absolute = err2 + err3;
			relative = 5 * OR_EPS * fabs(res);
			*abserr = fmax(absolute, relative);
			return;
		}
		e3 = epstab[n - 2 * i];
		epstab[n - 2 * i] = e1;
		delta1 = e1 - e3;
		err1 = fabs(delta1);
		tol1 = fmax(e1abs, fabs(e3)) * OR_EPS;
		if (err1 <= tol1 || err2 <= tol2 || err3 <= tol3) {
			n_final = 2 * i;
			break;
		}
		ss = (1 / delta1 + 1 / delta2) - 1 / delta3;
		if (fabs(ss * e1) <= 0.0001) {
			n_final = 2 * i;
			break;
		}
		res = e1 + 1 / ss;
		epstab[n - 2 * i] = res;
		{
			const double error = err2 + fabs(res - e2) + err3;
			if (error <= *abserr) {
				*abserr = error;
				*result = res;
			}
		}
	}
	if (n_final == 49) n_final = 2 * (49 / 2);
	if (n_orig % 2 == 1) {
		for (i = 0; i <= newelm; i++) epstab[1 + i * 2] = epstab[i * 2 + 3];
	} else {
		for (i = 0; i <= newelm; i++) epstab[i * 2] = epstab[i * 2 + 2];
	}
	if (n_orig != n_final)
		for (i = 0; i <= n_final; i++) epstab[i] = epstab[n_orig - n_final + i];
	t->n = n_final + 1;
	if (nres_orig < 3) {
		res3la[nres_orig] = *result;
		*abserr = OR_DMAX;
	} else {
		*abserr = fabs(*result - res3la[2]) + fabs(*result - res3la[1]) + fabs(*result - res3la[0]);
		res3la[0] = res3la[1];
		res3la[1] = res3la[2];
		res3la[2] = *result;
	}
	t->nres = nres_orig + 1;
	*abserr = fmax(*abserr, 5 * OR_EPS * fabs(*result));
}

/* qags with the rule of `npts` (21 or 15) points.  Returns GSL's status
 * class (0 success, > 0 one of its error codes); result/abserr/neval as GSL
 * reports them. */
static int or_qags(or_func f, void *ctx, double a, double b, double epsabs, double epsrel, size_t limit, int npts,
                   double *result, double *abserr, size_t *neval)
{
	static __thread or_qws ws;
	or_qws *w = &ws;
	const int nk = npts == 21 ? 11 : 8;
	const double *xgk = npts == 21 ? or_xgk21 : or_xgk15, *wg = npts == 21 ? or_wg21 : or_wg15,
	             *wgk = npts == 21 ? or_wgk21 : or_wgk15;
	double area, errsum, res_ext, err_ext, result0, abserr0, resabs0, resasc0, tolerance;
	double ertest = 0, error_over_large_intervals = 0, reseps = 0, abseps = 0, correc = 0;
	size_t ktmin = 0, iteration = 0;
	int roundoff_type1 = 0, roundoff_type2 = 0, roundoff_type3 = 0, error_type = 0, error_type2 = 0;
	int positive_integrand, extrapolate = 0, disallow_extrapolation = 0;
	or_qtab table;

	if (limit > OR_QLIMIT) limit = OR_QLIMIT;
	w->limit = limit;
	w->size = 0;
	w->nrmax = 0;
	w->i = 0;
	w->alist[0] = a;
	w->blist[0] = b;
	w->rlist[0] = 0;
	w->elist[0] = 0;
	w->order[0] = 0;
	w->level[0] = 0;
	w->maximum_level = 0;
	*result = 0;
	*abserr = 0;
	*neval = 0;
	if (epsabs <= 0 && (epsrel < 50 * OR_EPS || epsrel < 0.5e-28)) return 4; /* GSL_EBADTOL */

	or_qk(nk, xgk, wg, wgk, f, ctx, a, b, &result0, &abserr0, &resabs0, &resasc0);
	*neval = (size_t)npts;
	w->size = 1;
	w->rlist[0] = result0;
	w->elist[0] = abserr0;
	tolerance = fmax(epsabs, epsrel * fabs(result0));
	if (abserr0 <= 100 * OR_EPS * resabs0 && abserr0 > tolerance) {
		*result = result0;
		*abserr = abserr0;
		return 18; /* GSL_EROUND */
	} else if ((abserr0 <= tolerance && abserr0 != resasc0) || abserr0 == 0.0) {
		*result = result0;
		*abserr = abserr0;
		return 0;
	} else if (limit == 1) {
		*result = result0;
		*abserr = abserr0;
		return 11; /* GSL_EMAXITER */
	}

	table.n = 0;
	table.nres = 0;
	table.rlist2[table.n++] = result0;
	area = result0;
	errsum = abserr0;
	res_ext = result0;
	err_ext = OR_DMAX;
	positive_integrand = fabs(result0) >= (1 - 50 * OR_EPS) * resabs0;
	iteration = 1;

	do {
		size_t current_level;
		double a1, b1, a2, b2, a_i, b_i, r_i, e_i;
		double area1 = 0, area2 = 0, area12 = 0, error1 = 0, error2 = 0, error12 = 0;
		double resasc1, resasc2, resabs1, resabs2, last_e_i;

		a_i = w->alist[w->i];
		b_i = w->blist[w->i];
		r_i = w->rlist[w->i];
		e_i = w->elist[w->i];
		current_level = w->level[w->i] + 1;
		a1 = a_i;
		b1 = 0.5 * (a_i + b_i);
		a2 = b1;
		b2 = b_i;
		iteration++;
		or_qk(nk, xgk, wg, wgk, f, ctx, a1, b1, &area1, &error1, &resabs1, &resasc1);
		or_qk(nk, xgk, wg, wgk, f, ctx, a2, b2, &area2, &error2, &resabs2, &resasc2);
		*neval += 2 * (size_t)npts;
		area12 = area1 + area2;
		error12 = error1 + error2;
		last_e_i = e_i;
		errsum = errsum + error12 - e_i;
		area = area + area12 - r_i;
		tolerance = fmax(epsabs, epsrel * fabs(area));
		if (resasc1 != error1 && resasc2 != error2) {
			double delta = r_i - area12;
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
			/* bad integrand behaviour at a point of the range */
			double tmp = (1 + 100 * OR_EPS) * (fabs(a2) + 1000 * OR_DMIN);
			if (fabs(a1) <= tmp && fabs(b2) <= tmp) error_type = 4;
		}
		or_qupdate(w, a1, b1, area1, error1, a2, b2, area2, error2);
		if (errsum <= tolerance) goto compute_result;
		if (error_type) break;
		if (iteration >= limit - 1) {
			error_type = 1;
			break;
		}
		if (iteration == 2) {
			error_over_large_intervals = errsum;
			ertest = tolerance;
			table.rlist2[table.n++] = area;
			continue;
		}
		if (disallow_extrapolation) continue;
		error_over_large_intervals += -last_e_i;
		if (current_level < w->maximum_level) error_over_large_intervals += error12;
		if (!extrapolate) {
			/* the interval to be bisected next is not the smallest */
			if (w->level[w->i] < w->maximum_level) continue;
			extrapolate = 1;
			w->nrmax = 1;
		}
		if (!error_type2 && error_over_large_intervals > ertest)
			if (or_increase_nrmax(w)) continue;
		table.rlist2[table.n++] = area;
		or_qelg(&table, &reseps, &abseps);
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
		if (table.n == 1) disallow_extrapolation = 1;
		if (error_type == 5) break;
		w->nrmax = 0;
		w->i = w->order[0];
		extrapolate = 0;
		error_over_large_intervals = errsum;
	} while (iteration < limit);

	*result = res_ext;
	*abserr = err_ext;
	if (err_ext == OR_DMAX) goto compute_result;
	if (error_type || error_type2) {
		if (error_type2) err_ext += correc;
		if (error_type == 0) error_type = 3;
		if (res_ext != 0.0 && area != 0.0) {
			if (err_ext / fabs(res_ext) > errsum / fabs(area)) goto compute_result;
		} else if (err_ext > errsum) {
			goto compute_result;
		} else if (area == 0.0) {
			goto return_error;
		}
	}
	{
		double max_area = fmax(fabs(res_ext), fabs(area));
		if (!positive_integrand && max_area < 0.01 * resabs0) goto return_error;
	}
	{
		double ratio = res_ext / area;
		if (ratio < 0.01 || ratio > 100.0 || errsum > fabs(area)) error_type = 6;
	}
	goto return_error;

compute_result : {
	double sum = 0;
	for (size_t k = 0; k < w->size; k++) sum += w->rlist[k];
	*result = sum;
	*abserr = errsum;
}

return_error:
	if (error_type > 2) error_type--;
	return error_type;
}

struct or_inf_ctx {
	or_func f;
	void *ctx;
	double a;
};

/* gsl_integration_qagiu's transformation x = a + (1 - t) / t */
static double or_iu_transform(double t, void *p)
{
	struct or_inf_ctx *c = (struct or_inf_ctx *)p;
	double x = c->a + (1 - t) / t;
	double y = c->f(x, c->ctx);
	return (y / t) / t;
}

static int or_qagiu(or_func f, void *ctx, double a, double epsabs, double epsrel, size_t limit, double *result,
                    double *abserr, size_t *neval)
{
	struct or_inf_ctx c = {f, ctx, a};
	return or_qags(or_iu_transform, &c, 0.0, 1.0, epsabs, epsrel, limit, 15, result, abserr, neval);
}

/* Integrator of the luminosity integrals: PK_INTEGRATOR_QUADPACK is
 * libprofit's (qags / qagiu at epsabs = epsrel = 1e-4, limit 100);
 * PK_INTEGRATOR_EXACT the same algorithm run to epsrel 1e-13 with a limit
 * of 1000 intervals, so the integral itself. */
static int or_integrator = PK_INTEGRATOR_QUADPACK;
void or_set_integrator(int mode) { or_integrator = mode; }

double or_integrate_qags(or_func f, double a, double b, void *ctx)
{
	double r, e;
	size_t n;
	if (or_integrator == PK_INTEGRATOR_EXACT)
		or_qags(f, ctx, a, b, 0, 1e-13, 1000, 21, &r, &e, &n);
	else
		or_qags(f, ctx, a, b, 1e-4, 1e-4, 100, 21, &r, &e, &n);
	return r;
}

double or_integrate_qagi(or_func f, double a, void *ctx)
{
	double r, e;
	size_t n;
	if (or_integrator == PK_INTEGRATOR_EXACT)
		or_qagiu(f, ctx, a, 0, 1e-13, 1000, &r, &e, &n);
	else
		or_qagiu(f, ctx, a, 1e-4, 1e-4, 100, &r, &e, &n);
	return r;
}

/* The integrands of GSL's own test suite for qags / qagiu
 * (integration/tests.c: f1, f11, f15, f455) and a runner returning
 * {result, abserr, neval, status} for pinning the restatement. */
static double or_t_f1(double x, void *p) { double al = *(double *)p; return pow(x, al) * log(1 / x); }
static double or_t_f11(double x, void *p) { double al = *(double *)p; return pow(log(1 / x), al - 1); }
static double or_t_f15(double x, void *p) { double al = *(double *)p; return x * x * exp(-pow(2.0, -al) * x); }
static double or_t_f455(double x, void *p) { (void)p; return log(x) / (1 + 100 * x * x); }

int or_quadpack_test(int fn, int infinite, double alpha, double a, double b, double epsabs, double epsrel,
                     unsigned limit, double *out)
{
	or_func f = fn == 1 ? or_t_f1 : fn == 11 ? or_t_f11 : fn == 15 ? or_t_f15 : or_t_f455;
	double r, e;
	size_t n;
	int st = infinite ? or_qagiu(f, &alpha, a, epsabs, epsrel, limit, &r, &e, &n)
	                  : or_qags(f, &alpha, a, b, epsabs, epsrel, limit, 21, &r, &e, &n);
	out[0] = r;
	out[1] = e;
	out[2] = (double)n;
	out[3] = st;
	return 0;
}

/* qags / qagiu as libprofit calls them, on an integrand given by ctypes */
int or_quadpack(or_func f, void *ctx, int infinite, double a, double b, double epsabs, double epsrel,
                unsigned limit, double *out)
{
	double r, e;
	size_t n;
	int st = infinite ? or_qagiu(f, ctx, a, epsabs, epsrel, limit, &r, &e, &n)
	                  : or_qags(f, ctx, a, b, epsabs, epsrel, limit, 21, &r, &e, &n);
	out[0] = r;
	out[1] = e;
	out[2] = (double)n;
	out[3] = st;
	return 0;
}

/* ======================================================================== */
/* Profiles                                                                  */
/* ======================================================================== */

enum { SV_GENERAL = 0, SV_HALF, SV_ONE, SV_TWO, SV_THREE, SV_FOUR, SV_EIGHT, SV_SIXTEEN };

typedef struct {
	int kind, convolve;
	double xcen, ycen, mag, ang, axrat, box;
	int rough, adjust, rescale_flux;
	double acc, rscale_switch, rscale_max;
	unsigned resolution, max_recursions;
	/* derived */
	double rscale, ie, cos_ang, sin_ang, _xcen, _ycen, flux_scale;
	double img_scale_x, img_scale_y;
	/* sersic / coresersic */
	double re, nser, bn, rescale_factor;
	int sv, boxy;
	/* moffat */
	double fwhm, con;
	/* ferrer */
	double rout, fa, fb;
	/* king */
	double rc, rt, ka;
	/* brokenexp */
	double h1, h2, rb, ba;
	/* coresersic */
	double crb, ca, cb;
	/* sky */
	double bg;
	/* psf profile */
	double pmag;
} or_prof;

static int or_almost(double x, double y) { return fabs(x - y) < 1e-10; }

/* boxy radius (include/profit/radial.h:93-102) */
static inline double or_boxy_r(const or_prof *p, double x, double y)
{
	if (p->box == 0) return sqrt(x * x + y * y);
	double b2 = p->box + 2.;
	return pow(pow(fabs(x), b2) + pow(fabs(y), b2), 1. / b2);
}

/* Sersic evaluation (src/sersic.cpp:79-205), all specialisations. */
static inline double or_sersic_at(const or_prof *p, double x, double y)
{
	double B = p->box + 2, base, invexp, rf;
	if (p->boxy) {
		base = pow(fabs(x / p->re), B) + pow(fabs(y / p->re), B);
		invexp = p->nser * B;
		switch (p->sv) {
		case SV_HALF: rf = base * base; break;
		case SV_ONE: rf = base; break;
		case SV_TWO: rf = sqrt(base); break;
		case SV_THREE: rf = cbrt(base); break;
		case SV_FOUR: rf = sqrt(sqrt(base)); break;
		case SV_EIGHT: rf = sqrt(sqrt(sqrt(base))); break;
		case SV_SIXTEEN: rf = sqrt(sqrt(sqrt(sqrt(base)))); break;
		default: rf = pow(base, 1 / invexp);
		}
	} else {
		base = (x * x + y * y) / (p->re * p->re);
		invexp = p->nser;
		switch (p->sv) {
		case SV_HALF: rf = base; break;
		case SV_ONE: rf = sqrt(base); break;
		case SV_TWO: rf = sqrt(sqrt(base)); break;
		case SV_THREE: rf = cbrt(sqrt(base)); break;
		case SV_FOUR: rf = sqrt(sqrt(sqrt(base))); break;
		case SV_EIGHT: rf = sqrt(sqrt(sqrt(sqrt(base)))); break;
		case SV_SIXTEEN: rf = sqrt(sqrt(sqrt(sqrt(sqrt(base))))); break;
		default: rf = pow(sqrt(base), 1 / invexp);
		}
	}
	return exp(-p->bn * (rf - 1));
}

static inline double or_broken_exponential(double r, double h1, double h2, double rb, double a)
{
	/* src/brokenexponential.cpp:39-76 */
	double base = r - rb;
	double expo = 1 / h1 - 1 / h2;
	if (a * base < 40) base = log(1 + exp(a * base)) / a;
	return exp(-r / h1 + expo * base);
}

static inline double or_king_r(const or_prof *p, double r)
{
	/* src/king.cpp:51-59 */
	if (r < p->rt)
		return pow(1 / pow(1 + pow(r / p->rc, 2), 1 / p->ka) - 1 / pow(1 + pow(p->rt / p->rc, 2), 1 / p->ka), p->ka);
	return 0;
}

static inline double or_coresersic_r(const or_prof *p, double r)
{
	/* src/coresersic.cpp:51-60 */
	return pow(1 + pow(r / p->crb, -p->ca), p->cb / p->ca) *
	       exp(-p->bn * pow((pow(r, p->ca) + pow(p->crb, p->ca)) / pow(p->re, p->ca), 1 / (p->nser * p->ca)));
}

static inline double or_evaluate_at(const or_prof *p, double x, double y)
{
	switch (p->kind) {
	case PK_PROFILE_SERSIC: return or_sersic_at(p, x, y);
	case PK_PROFILE_MOFFAT: {
		/* src/moffat.cpp:52-58 */
		double r = or_boxy_r(p, x, y);
		double rf = r / p->rscale;
		return pow(1 + rf * rf, -p->con);
	}
	case PK_PROFILE_FERRER: {
		/* src/ferrer.cpp:50-61 */
		double r = or_boxy_r(p, x, y);
		double rf = r / p->rscale;
		if (rf < 1) return pow(1 - pow(rf, 2 - p->fb), p->fa);
		return 0;
	}
	case PK_PROFILE_KING: return or_king_r(p, or_boxy_r(p, x, y));
	case PK_PROFILE_BROKENEXP:
		return or_broken_exponential(or_boxy_r(p, x, y), p->h1, p->h2, p->rb, p->ba);
	case PK_PROFILE_CORESERSIC: return or_coresersic_r(p, or_boxy_r(p, x, y));
	}
	return 0;
}

/* integrands for lumtot (src/king.cpp:76-97, src/brokenexponential.cpp:107-127,
 * src/coresersic.cpp:82-104) */
static double or_king_int(double r, void *c) { return r * or_king_r((const or_prof *)c, r); }
static double or_bexp_int(double r, void *c)
{
	const or_prof *p = (const or_prof *)c;
	return r * or_broken_exponential(r, p->h1, p->h2, p->rb, p->ba);
}
static double or_csersic_int(double r, void *c) { return r * or_coresersic_r((const or_prof *)c, r); }

/* ---- validation (RadialProfile::validate and per-profile) ---------------- */
static int or_validate(const pk_component *c, int has_psf)
{
	switch (c->kind) {
	case PK_PROFILE_NULL:
	case PK_PROFILE_SKY: return PK_OK;
	case PK_PROFILE_PSF:
		if (!has_psf) return fail(PK_EINVAL, "No psf present in the model, cannot produce a psf profile");
		return PK_OK;
	default: break;
	}
	/* src/radial.cpp:204-214 */
	if (c->axrat <= 0) return fail(PK_EINVAL, "axrat <= 0, must have axrat > 0");
	if (c->axrat > 1) return fail(PK_EINVAL, "axrat > 1, must have axrat <= 1");
	if (c->box <= -2) return fail(PK_EINVAL, "box <= -2, must have box > -2");
	const double *p = c->p;
	switch (c->kind) {
	case PK_PROFILE_SERSIC:
		if (p[PK_SERSIC_RE] <= 0) return fail(PK_EINVAL, "re <= 0, must have re > 0");
		if (p[PK_SERSIC_NSER] <= 0) return fail(PK_EINVAL, "nser <= 0, must have nser > 0");
		break;
	case PK_PROFILE_MOFFAT:
		if (p[PK_MOFFAT_FWHM] <= 0) return fail(PK_EINVAL, "fwhm <= 0, must have fwhm > 0");
		if (p[PK_MOFFAT_CON] < 0) return fail(PK_EINVAL, "con < 0, must have con >= 0");
		break;
	case PK_PROFILE_FERRER:
		if (p[PK_FERRER_ROUT] <= 0) return fail(PK_EINVAL, "rout <= 0, must have rout >= 0");
		if (p[PK_FERRER_A] < 0) return fail(PK_EINVAL, "a < 0, must have a >= 0");
		if (p[PK_FERRER_B] > 2) return fail(PK_EINVAL, "b > 2, must have b <= 2");
		break;
	case PK_PROFILE_KING:
		if (p[PK_KING_RC] <= 0) return fail(PK_EINVAL, "rc <= 0, must have rc > 0");
		if (p[PK_KING_RT] <= 0) return fail(PK_EINVAL, "rt <= 0, must have rt > 0");
		if (p[PK_KING_A] < 0) return fail(PK_EINVAL, "a < 0, must have a >=0");
		break;
	case PK_PROFILE_BROKENEXP:
		if (p[PK_BROKENEXP_H1] <= 0) return fail(PK_EINVAL, "h1 <= 0, must have h1 > 0");
		if (p[PK_BROKENEXP_H2] <= 0) return fail(PK_EINVAL, "h2 <= 0, must have h2 > 0");
		if (p[PK_BROKENEXP_RB] <= 0) return fail(PK_EINVAL, "rb <= 0, must have rb > 0");
		break;
	case PK_PROFILE_CORESERSIC:
		if (p[PK_CORESERSIC_RE] <= 0) return fail(PK_EINVAL, "re <= 0, must have re > 0");
		if (p[PK_CORESERSIC_RB] <= 0) return fail(PK_EINVAL, "rb <= 0, must have rb > 0");
		if (p[PK_CORESERSIC_NSER] <= 0) return fail(PK_EINVAL, "nser <= 0, must have nser > 0");
		if (p[PK_CORESERSIC_A] <= 0) return fail(PK_EINVAL, "a <= 0, must have a > 0");
		if (p[PK_CORESERSIC_B] > 1.999) return fail(PK_EINVAL, "b > 1.999, must have b < 1.999");
		break;
	default: return fail(PK_EINVAL, "Unknown profile kind");
	}
	return PK_OK;
}

/* ---- per-profile set-up (RadialProfile::initial_calculations and the
 *      profile overrides) --------------------------------------------------- */
static double or_sersic_fluxfrac(const or_prof *p, double fraction)
{
	/* src/sersic.cpp:257-260 */
	double ratio = or_qgamma(fraction, 2 * p->nser) / p->bn;
	return p->re * pow(ratio, p->nser);
}

static double or_moffat_fluxfrac(const or_prof *p, double fraction)
{
	/* src/moffat.cpp:72-75 */
	return p->rscale * sqrt(pow(1 - fraction, 1 / (1 - p->con)) - 1);
}

static void or_setup(or_prof *p, const pk_component *c, const pk_model *m, double sx, double sy,
                     double pad_x, double pad_y, unsigned fs)
{
	memset(p, 0, sizeof(*p));
	p->kind = c->kind;
	p->convolve = c->convolve;
	p->xcen = c->xcen;
	p->ycen = c->ycen;
	p->mag = c->mag;
	p->ang = c->ang;
	p->axrat = c->axrat;
	p->box = c->box;
	p->rough = c->rough;
	p->adjust = c->adjust;
	p->rescale_flux = c->rescale_flux;
	p->acc = c->acc;
	p->rscale_switch = c->rscale_switch;
	p->rscale_max = c->rscale_max;
	p->resolution = c->resolution;
	p->max_recursions = c->max_recursions;
	p->img_scale_x = m->scale_x;
	p->img_scale_y = m->scale_y;
	if (c->kind == PK_PROFILE_SKY) {
		/* SkyProfile::adjust_for_finesampling (src/sky.cpp:42-45) */
		p->bg = c->p[PK_SKY_BG] / (fs * fs);
		return;
	}
	if (c->kind == PK_PROFILE_PSF) {
		p->pmag = c->mag;
		return;
	}
	if (c->kind == PK_PROFILE_NULL) return;

	const double *q = c->p;
	switch (c->kind) {
	case PK_PROFILE_SERSIC:
		p->re = q[PK_SERSIC_RE];
		p->nser = q[PK_SERSIC_NSER];
		p->bn = or_qgamma(0.5, 2 * p->nser); /* src/sersic.cpp:297 */
		p->rscale = p->re;
		{
			/* specialisation choice (src/sersic.cpp:225-255) */
			double inv = p->nser;
			p->boxy = p->box != 0;
			if (p->boxy) inv *= (p->box + 2);
			p->sv = or_almost(inv, 0.5) ? SV_HALF : or_almost(inv, 1) ? SV_ONE : or_almost(inv, 2) ? SV_TWO :
			        or_almost(inv, 3) ? SV_THREE : or_almost(inv, 4) ? SV_FOUR : or_almost(inv, 8) ? SV_EIGHT :
			        or_almost(inv, 16) ? SV_SIXTEEN : SV_GENERAL;
		}
		break;
	case PK_PROFILE_MOFFAT:
		p->fwhm = q[PK_MOFFAT_FWHM];
		p->con = q[PK_MOFFAT_CON];
		p->rscale = p->fwhm / (2 * sqrt(pow(2, (1 / p->con)) - 1)); /* src/moffat.cpp:81-83 */
		break;
	case PK_PROFILE_FERRER:
		p->rout = q[PK_FERRER_ROUT];
		p->fa = q[PK_FERRER_A];
		p->fb = q[PK_FERRER_B];
		p->rscale = p->rout;
		break;
	case PK_PROFILE_KING:
		p->rc = q[PK_KING_RC];
		p->rt = q[PK_KING_RT];
		p->ka = q[PK_KING_A];
		p->rscale = p->rt;
		break;
	case PK_PROFILE_BROKENEXP:
		p->h1 = q[PK_BROKENEXP_H1];
		p->h2 = q[PK_BROKENEXP_H2];
		p->rb = q[PK_BROKENEXP_RB];
		p->ba = q[PK_BROKENEXP_A];
		p->rscale = p->h1 * 4;
		break;
	case PK_PROFILE_CORESERSIC:
		p->re = q[PK_CORESERSIC_RE];
		p->crb = q[PK_CORESERSIC_RB];
		p->nser = q[PK_CORESERSIC_NSER];
		p->ca = q[PK_CORESERSIC_A];
		p->cb = q[PK_CORESERSIC_B];
		p->bn = or_qgamma(0.5, 2 * p->nser);
		p->rscale = p->re;
		break;
	}

	/* lumtot (get_lumtot overrides) */
	double lum = 0;
	switch (c->kind) {
	case PK_PROFILE_SERSIC:
		/* src/sersic.cpp:283-289 */
		lum = pow(p->re, 2) * 2 * M_PI * p->nser * or_gammafn(2 * p->nser) * exp(p->bn) / pow(p->bn, 2 * p->nser);
		break;
	case PK_PROFILE_MOFFAT: lum = pow(p->rscale, 2) * M_PI / (p->con - 1); break; /* src/moffat.cpp:77-79 */
	case PK_PROFILE_FERRER:
		lum = pow(p->rout, 2) * M_PI * (p->fa * or_beta(p->fa, 1 + 2 / (2 - p->fb))); /* src/ferrer.cpp:85-101 */
		break;
	case PK_PROFILE_KING: lum = 2 * M_PI * or_integrate_qags(or_king_int, 0, p->rt, p); break;
	case PK_PROFILE_BROKENEXP: lum = 2 * M_PI * or_integrate_qagi(or_bexp_int, 0, p); break;
	case PK_PROFILE_CORESERSIC: lum = 2 * M_PI * or_integrate_qagi(or_csersic_int, 0, p); break;
	}

	/* RadialProfile::initial_calculations (src/radial.cpp:137-199) */
	double b2 = p->box + 2;
	double r_box = M_PI * b2 / (2 * or_beta(1 / b2, 1 / b2));
	double lumtot = lum * p->axrat / r_box;
	p->ie = pow(10, -0.4 * (p->mag - m->magzero)) / lumtot;

	if (p->adjust) {
		double rsw = 0, rmax = 0, acc = p->acc;
		switch (c->kind) {
		case PK_PROFILE_SERSIC:
			rsw = ceil(or_sersic_fluxfrac(p, 1. - p->nser * p->nser / 2e3));
			rsw = fmax(fmin(rsw, 20.), 2.);
			rsw = rsw / p->re;
			break;
		case PK_PROFILE_MOFFAT:
			rsw = fmax(or_moffat_fluxfrac(p, 0.9999), 1.823 * p->fwhm);
			rsw = fmax(fmin(rsw, 20.), 2.) / p->rscale;
			break;
		case PK_PROFILE_FERRER: rsw = 0.5; break;
		case PK_PROFILE_KING: rsw = 0.5; break;
		case PK_PROFILE_BROKENEXP: rsw = 1; break;
		case PK_PROFILE_CORESERSIC: rsw = 1; break;
		}
		p->rscale_switch = rsw;
		double resd = ceil(160 / (p->rscale_switch * p->rscale));
		unsigned res = (unsigned)resd;
		res += res % 2;
		if (res > 16) res = 16;
		if (res < 4) res = 4;
		p->resolution = res;
		if (p->rscale_max == 0) {
			switch (c->kind) {
			case PK_PROFILE_SERSIC: rmax = ceil(fmax(or_sersic_fluxfrac(p, 0.9999), 2.) / p->re); break;
			case PK_PROFILE_MOFFAT: rmax = ceil(fmax(or_moffat_fluxfrac(p, 0.9999), 2.) / p->rscale); break;
			case PK_PROFILE_FERRER: rmax = 1; break;
			case PK_PROFILE_KING: rmax = 1.5; break;
			case PK_PROFILE_BROKENEXP: rmax = 100; break;
			case PK_PROFILE_CORESERSIC: rmax = 100; break;
			}
			p->rscale_max = rmax;
		}
		switch (c->kind) {
		case PK_PROFILE_SERSIC:
			acc /= sqrt(p->nser); /* src/sersic.cpp:278-281 */
			acc = fmax(0.1, acc) / p->axrat;
			break;
		case PK_PROFILE_MOFFAT: acc = 0.1 / p->axrat; break; /* src/moffat.cpp:97-99 */
		default: break;
		}
		p->acc = acc;
	}

	double angrad = fmod(p->ang + 90, 360.) * M_PI / 180.;
	p->cos_ang = cos(angrad);
	p->sin_ang = sin(angrad);

	p->rescale_factor = 1;
	if (c->kind == PK_PROFILE_SERSIC && p->adjust && p->rescale_flux) {
		/* src/sersic.cpp:303-311 */
		double flux_r = p->bn * pow(p->rscale_max / p->re, 1 / p->nser);
		flux_r = or_pgamma(flux_r, 2 * p->nser);
		p->rescale_factor = 1 / flux_r;
	}

	/* RadialProfile::evaluate (src/radial.cpp:254-255) */
	p->_xcen = p->xcen + pad_x * sx;
	p->_ycen = p->ycen + pad_y * sy;
	/* get_pixel_scale (src/radial.cpp:219-222, src/sersic.cpp:319-325) */
	p->flux_scale = sx * sy * p->ie;
	if (c->kind == PK_PROFILE_SERSIC && p->rescale_flux) p->flux_scale *= p->rescale_factor;
}

static inline void or_to_prof(const or_prof *p, double x, double y, double *xp, double *yp)
{
	/* src/radial.cpp:47-53 */
	x -= p->_xcen;
	y -= p->_ycen;
	*xp = x * p->cos_ang + y * p->sin_ang;
	*yp = -x * p->sin_ang + y * p->cos_ang;
	*yp /= p->axrat;
}

typedef struct {
	uint64_t n_subcells;
} or_stats;

/* ---- frozen refinement pattern (finite differences of the chi^2 gradient) --
 * Mode 1 records, for every radial component and pixel in evaluation order,
 * the pixel's class (0 beyond rscale_max, 1 evaluated at its centre, 2
 * refined), the resolution and recursion depth it used, and every accuracy
 * decision of subsample_pixel; mode 2 replays them, so that an evaluation at
 * other parameters keeps the recorded refinement pattern.  Both run on one
 * thread (the log is sequential). */
static int or_pat_mode = 0;
static uint32_t *or_pat = NULL;
static size_t or_pat_len = 0, or_pat_cap = 0, or_pat_pos = 0;
static int or_pat_bad = 0;

void or_pattern(int mode)
{
	or_pat_mode = mode;
	if (mode == 1) or_pat_len = 0;
	or_pat_pos = 0;
	or_pat_bad = 0;
}

/* 0: the replay consumed exactly the recorded log */
int or_pattern_status(void) { return or_pat_bad || (or_pat_mode == 2 && or_pat_pos != or_pat_len); }

static inline uint32_t or_pat_io(uint32_t v)
{
	if (or_pat_mode == 1) {
		if (or_pat_len == or_pat_cap) {
			or_pat_cap = or_pat_cap ? 2 * or_pat_cap : 1 << 16;
			or_pat = (uint32_t *)realloc(or_pat, or_pat_cap * sizeof(uint32_t));
		}
		or_pat[or_pat_len++] = v;
		return v;
	}
	if (or_pat_mode == 2) {
		if (or_pat_pos >= or_pat_len) {
			or_pat_bad = 1;
			return v;
		}
		return or_pat[or_pat_pos++];
	}
	return v;
}

/* RadialProfile::subsample_pixel (src/radial.cpp:55-135) */
static double or_subsample(const or_prof *p, double x0, double x1, double y0, double y1, unsigned level,
                           unsigned max_rec, unsigned res, uint64_t *nev)
{
	double xbin = (x1 - x0) / res;
	double ybin = (y1 - y0) / res;
	double half_xbin = xbin / 2.;
	double half_ybin = ybin / 2.;
	double total = 0, xp, yp;
	int recurse = res > 1 && level < max_rec;
	double *pts = NULL;
	size_t npts = 0, cap = 0;
	double x = x0;
	if (recurse) {
		for (unsigned i = 0; i < res; i++) {
			x += half_xbin;
			double y = y0;
			for (unsigned j = 0; j < res; j++) {
				y += half_ybin;
				or_to_prof(p, x, y, &xp, &yp);
				double subval = or_evaluate_at(p, xp, yp);
				double dyp = (-xbin * p->sin_ang + ybin * p->cos_ang) / p->axrat;
				double testval = or_evaluate_at(p, fabs(xp), fabs(yp) + fabs(dyp));
				*nev += 2;
				if (or_pat_io(fabs(testval / subval - 1.0) > p->acc)) {
					if (npts == cap) {
						cap = cap ? 2 * cap : 16;
						pts = (double *)realloc(pts, cap * 2 * sizeof(double));
					}
					pts[2 * npts] = x;
					pts[2 * npts + 1] = y;
					npts++;
				} else {
					total += subval;
				}
				y += half_ybin;
			}
			x += half_xbin;
		}
	} else {
		for (unsigned i = 0; i < res; i++) {
			x += half_xbin;
			double y = y0;
			for (unsigned j = 0; j < res; j++) {
				y += half_ybin;
				or_to_prof(p, x, y, &xp, &yp);
				total += or_evaluate_at(p, xp, yp);
				*nev += 1;
				y += half_ybin;
			}
			x += half_xbin;
		}
	}
	for (size_t k = 0; k < npts; k++) {
		double cx = pts[2 * k], cy = pts[2 * k + 1];
		total += or_subsample(p, cx - half_xbin, cx + half_xbin, cy - half_ybin, cy + half_ybin, level + 1,
		                      max_rec, res, nev);
	}
	free(pts);
	return total / (res * res);
}

/* RadialProfile::evaluate_cpu (src/radial.cpp:291-344) */
static void or_radial_evaluate(const or_prof *p, double *image, unsigned W, unsigned H, const uint8_t *mask,
                               double sx, double sy, uint64_t *n_refined, uint64_t *n_sub)
{
	double half_xbin = sx / 2., half_ybin = sy / 2.;
	uint64_t nref = 0, nsub = 0;
	long total = (long)W * H;
	const int nthreads = or_pat_mode ? 1 : or_threads;
#pragma omp parallel for schedule(dynamic, 10) num_threads(nthreads) reduction(+ : nref, nsub)
	for (long idx = 0; idx < total; idx++) {
		unsigned i = (unsigned)(idx % W), j = (unsigned)(idx / W);
		if (mask && !mask[idx]) continue;
		double xp, yp;
		double y = half_ybin + j * sy;
		double x = half_xbin + i * sx;
		or_to_prof(p, x, y, &xp, &yp);
		double r = sqrt(xp * xp + yp * yp);
		double val;
		const uint32_t cls = or_pat_io((p->rscale_max > 0 && r / p->rscale > p->rscale_max) ? 0u
		                               : (p->rough || r / p->rscale > p->rscale_switch) ? 1u
		                                                                                : 2u);
		if (cls == 0) {
			val = 0.;
		} else if (cls == 1) {
			val = or_evaluate_at(p, xp, yp);
		} else {
			unsigned res = p->resolution, mr = p->max_recursions;
			if (p->kind == PK_PROFILE_SERSIC) {
				/* SersicProfile::subsampling_params (src/sersic.cpp:331-347) */
				int center = fabs(x - p->_xcen) < p->img_scale_x && fabs(y - p->_ycen) < p->img_scale_y;
				if (p->adjust && center && p->nser > 1) {
					res = 8;
					mr = 10;
				}
			}
			res = or_pat_io(res);
			mr = or_pat_io(mr);
			uint64_t ne = 0;
			val = or_subsample(p, x - half_xbin, x + half_xbin, y - half_ybin, y + half_ybin, 0, mr, res, &ne);
			nref += 1;
			nsub += ne;
		}
		image[idx] += p->flux_scale * val;
	}
	*n_refined += nref;
	*n_sub += nsub;
}

static unsigned or_bind(double value, unsigned max)
{
	int iv = (int)floor(value);
	if (iv < 0) return 0;
	unsigned u = (unsigned)iv;
	return u < max ? u : max;
}

/* PsfProfile::evaluate (src/psf.cpp:58-145), including its mixed x/y offsets. */
static void or_psf_evaluate(const pk_component *c, const double *psf, unsigned pw, unsigned ph, double psx,
                            double psy, double *image, unsigned W, unsigned H, double sx, double sy, double offx,
                            double offy, double magzero)
{
	double scale = pow(10, -0.4 * (c->mag - magzero));
	double origin_x = c->xcen + offx * sx - pw * psx / 2;
	double end_x = c->xcen + offy * sy + pw * psx / 2;
	double origin_y = c->ycen + offx * sx - ph * psy / 2;
	double end_y = c->ycen + offy * sy + ph * psy / 2;
	unsigned x0 = or_bind(origin_x / sx, W - 1), y0 = or_bind(origin_y / sy, H - 1);
	unsigned x1 = or_bind(end_x / sx, W - 1), y1 = or_bind(end_y / sy, H - 1);
	for (unsigned py = y0; py <= y1; py++) {
		double y = py * sy;
		for (unsigned px = x0; px <= x1; px++) {
			double x = px * sx;
			int qx0 = (int)or_bind(floor((x - origin_x) / psx), pw - 1);
			int qx1 = (int)or_bind(floor((x - origin_x + sx) / psx), ph - 1);
			int qy0 = (int)or_bind(floor((y - origin_y) / psy), pw - 1);
			int qy1 = (int)or_bind(floor((y - origin_y + sy) / psy), ph - 1);
			double val = 0;
			for (int qy = qy0; qy <= qy1; qy++) {
				double ppy = qy * psy + origin_y;
				for (int qx = qx0; qx <= qx1; qx++) {
					double ppx = qx * psx + origin_x;
					double ix = fmin(x + sx, ppx + psx) - fmax(x, ppx);
					double iy = fmin(y + sy, ppy + psy) - fmax(y, ppy);
					val += psf[qx + qy * pw] * (ix * iy) / (psx * psy);
				}
			}
			image[px + py * W] += val * scale;
		}
	}
}

/* ======================================================================== */
/* Convolution                                                               */
/* ======================================================================== */

/* BruteForceConvolver::convolve_impl (src/convolve.cpp:101-156) */
static void or_brute_old(const double *src, unsigned W, unsigned H, const double *krn, unsigned kw, unsigned kh,
                         const uint8_t *mask, double *out)
{
	unsigned hw = kw / 2, hh = kh / 2;
	long total = (long)W * H;
#pragma omp parallel for schedule(dynamic, 10) num_threads(or_threads)
	for (long idx = 0; idx < total; idx++) {
		int i = (int)(idx % W), j = (int)(idx / W);
		if (mask && !mask[idx]) {
			out[idx] = 0;
			continue;
		}
		double pixel = 0;
		const double *kp = krn + (size_t)kw * kh - 1;
		for (unsigned l = 0; l < kh; l++) {
			int sj = j + (int)l - (int)hh;
			for (unsigned k = 0; k < kw; k++) {
				int si = i + (int)k - (int)hw;
				if (si >= 0 && si < (int)W && sj >= 0 && sj < (int)H) pixel += src[si + (size_t)sj * W] * *kp;
				kp--;
			}
		}
		out[idx] = pixel;
	}
}

/* dot_product<AVX> (include/profit/dot_product.h:272-384), the product
 * create_convolver(BRUTE) picks on an AVX host (src/convolve.cpp:527-531):
 * four lane accumulators; 8-element batches add (s[0:4]k[0:4] +
 * s[4:8]k[4:8]) into them; a remainder of 4, 2 and 1 element(s) adds into
 * lanes 0-3, 0-1 and 0; the lanes are then summed as (l2 + l3) + (l0 + l1).
 * Written with GCC vector types, so an -O3 -march=native build runs it in
 * vector registers. */
typedef double or_v4d __attribute__((vector_size(32)));

static inline or_v4d or_ld4(const double *p)
{
	or_v4d v;
	memcpy(&v, p, sizeof(v));
	return v;
}

static inline double or_dot_avx(const double *s, const double *k, size_t n)
{
	or_v4d acc = {0, 0, 0, 0};
	size_t nb = n / 8;
	for (size_t b = 0; b < nb; b++) {
		const double *a = s + 8 * b, *c = k + 8 * b;
		acc = (or_ld4(a) * or_ld4(c) + or_ld4(a + 4) * or_ld4(c + 4)) + acc;
	}
	size_t rem = n % 8;
	const double *sr = s + n - rem, *kr = k + n - rem;
	if (rem >= 4) {
		acc = or_ld4(sr) * or_ld4(kr) + acc;
		sr += 4;
		kr += 4;
		rem -= 4;
	}
	if (rem >= 2) {
		or_v4d t = {sr[0] * kr[0], sr[1] * kr[1], 0, 0};
		acc = t + acc;
		sr += 2;
		kr += 2;
		rem -= 2;
	}
	if (rem == 1) {
		or_v4d t = {sr[0] * kr[0], 0, 0, 0};
		acc = t + acc;
	}
	return (acc[2] + acc[3]) + (acc[0] + acc[1]);
}

/* AssociativeBruteForceConvolver::convolve_impl (src/convolve.cpp:158-231) */
static void or_brute(const double *src, unsigned W, unsigned H, const double *krn, unsigned kw, unsigned kh,
                     const uint8_t *mask, double *out)
{
	unsigned hw = kw / 2, hh = kh / 2;
	size_t kn = (size_t)kw * kh;
	double *ik = (double *)malloc(kn * sizeof(double));
	for (size_t t = 0; t < kn; t++) ik[t] = krn[kn - 1 - t];
	long total = (long)W * H;
	size_t skoff = hw + (size_t)hh * W;
#pragma omp parallel for schedule(dynamic, 10) num_threads(or_threads)
	for (long idx = 0; idx < total; idx++) {
		unsigned i = (unsigned)(idx % W), j = (unsigned)(idx / W);
		if (mask && !mask[idx]) {
			out[idx] = 0;
			continue;
		}
		size_t so = (size_t)idx - skoff, ko = 0;
		unsigned lmin = 0, lmax = kh, kmin = 0, kmax = kw;
		if (j < hh) lmin = hh - j;
		else if (j + hh >= H) lmax = H + hh - j;
		if (i < hw) kmin = hw - i;
		else if (i + hw >= W) kmax = W + hw - i;
		so += kmin + (size_t)lmin * W;
		ko += kmin + (size_t)lmin * kw;
		double pixel = 0;
		for (unsigned l = 0; l < lmax - lmin; l++) {
			pixel += or_dot_avx(src + so, ik + ko, kmax - kmin);
			so += W;
			ko += kw;
		}
		out[idx] = pixel;
	}
	free(ik);
}

/* ---- own FFT (radix-2, complex, in place) for the FFT convolver ---------- */
static void or_fft1(double *re, double *im, size_t n, size_t stride, int inverse)
{
	/* bit reversal */
	for (size_t i = 1, j = 0; i < n; i++) {
		size_t bit = n >> 1;
		for (; j & bit; bit >>= 1) j ^= bit;
		j ^= bit;
		if (i < j) {
			double t = re[i * stride];
			re[i * stride] = re[j * stride];
			re[j * stride] = t;
			t = im[i * stride];
			im[i * stride] = im[j * stride];
			im[j * stride] = t;
		}
	}
	for (size_t len = 2; len <= n; len <<= 1) {
		double ang = 2 * M_PI / len * (inverse ? 1 : -1);
		for (size_t i = 0; i < n; i += len) {
			for (size_t k = 0; k < len / 2; k++) {
				double wr = cos(ang * k), wi = sin(ang * k);
				size_t a = (i + k) * stride, b = (i + k + len / 2) * stride;
				double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
				re[b] = re[a] - xr;
				im[b] = im[a] - xi;
				re[a] += xr;
				im[a] += xi;
			}
		}
	}
}

static void or_fft2(double *re, double *im, size_t nx, size_t ny, int inverse)
{
#pragma omp parallel for num_threads(or_threads)
	for (long y = 0; y < (long)ny; y++) or_fft1(re + y * nx, im + y * nx, nx, 1, inverse);
#pragma omp parallel for num_threads(or_threads)
	for (long x = 0; x < (long)nx; x++) or_fft1(re + x, im + x, ny, nx, inverse);
}

static size_t or_pow2(size_t n)
{
	size_t p = 1;
	while (p < n) p <<= 1;
	return p;
}

/* full linear convolution L (size (W+kw-1) x (H+kh-1)) by FFT */
static void or_linear_conv_fft(const double *src, unsigned W, unsigned H, const double *krn, unsigned kw,
                               unsigned kh, double *L)
{
	size_t lw = W + kw - 1, lh = H + kh - 1;
	size_t nx = or_pow2(lw), ny = or_pow2(lh), n = nx * ny;
	double *ar = calloc(n, sizeof(double)), *ai = calloc(n, sizeof(double));
	double *br = calloc(n, sizeof(double)), *bi = calloc(n, sizeof(double));
	for (unsigned j = 0; j < H; j++)
		for (unsigned i = 0; i < W; i++) ar[i + j * nx] = src[i + (size_t)j * W];
	for (unsigned j = 0; j < kh; j++)
		for (unsigned i = 0; i < kw; i++) br[i + j * nx] = krn[i + (size_t)j * kw];
	or_fft2(ar, ai, nx, ny, 0);
	or_fft2(br, bi, nx, ny, 0);
	for (size_t t = 0; t < n; t++) {
		double r = ar[t] * br[t] - ai[t] * bi[t];
		double im = ar[t] * bi[t] + ai[t] * br[t];
		ar[t] = r;
		ai[t] = im;
	}
	or_fft2(ar, ai, nx, ny, 1);
	for (size_t j = 0; j < lh; j++)
		for (size_t i = 0; i < lw; i++) L[i + j * lw] = ar[i + j * nx] / (double)n;
	free(ar);
	free(ai);
	free(br);
	free(bi);
}

/* FFTConvolver::offset_after_convolution (src/convolve.cpp:270-281) */
static void or_fft_offset(unsigned sw, unsigned sh, unsigned kw, unsigned kh, int *ox, int *oy)
{
	*ox = (int)(sw / 2);
	*oy = (int)(sh / 2);
	if (sw % 2 == 0 || kw % 2 == 0) *ox -= 1;
	if (sh % 2 == 0 || kh % 2 == 0) *oy -= 1;
}

/* Convolver::padding (src/convolve.cpp:50-53, 263-268) */
int or_convolver_padding(int type, unsigned sw, unsigned sh, unsigned kw, unsigned kh, int *lo_x, int *lo_y,
                         int *hi_x, int *hi_y)
{
	if (type != PK_CONV_FFT) {
		*lo_x = *lo_y = *hi_x = *hi_y = 0;
		return PK_OK;
	}
	/* max() of include/profit/image.h:204-207 uses x for both components */
	unsigned mx = sw > kw ? sw : kw;
	unsigned ex = 2 * mx, ey = 2 * mx;
	int ox, oy;
	or_fft_offset(sw, sh, kw, kh, &ox, &oy);
	*lo_x = ox;
	*lo_y = oy;
	*hi_x = (int)(ex - ex / 2) - ox;
	*hi_y = (int)(ey - ey / 2) - oy;
	return PK_OK;
}

/* convolve_impl for one convolver type; src dims >= krn dims. Returns a
 * malloc'd image in *out with its dims and offset. */
static int or_convolve_impl(int type, const double *src, unsigned W, unsigned H, const double *krn, unsigned kw,
                            unsigned kh, const uint8_t *mask, int crop, double **out, unsigned *ow, unsigned *oh,
                            int *offx, int *offy)
{
	if (type == PK_CONV_BRUTE_OLD || type == PK_CONV_BRUTE || type == PK_CONV_OPENCL) {
		double *o = (double *)malloc((size_t)W * H * sizeof(double));
		if (type == PK_CONV_BRUTE_OLD) or_brute_old(src, W, H, krn, kw, kh, mask, o);
		else or_brute(src, W, H, krn, kw, kh, mask, o);
		*out = o;
		*ow = W;
		*oh = H;
		return PK_OK;
	}
	if (type != PK_CONV_FFT) return fail(PK_EINVAL, "Unsupported convolver type");
	/* FFTConvolver::convolve_impl (src/convolve.cpp:283-312): the 1-D
	 * transform over the flattened 2*max(src,krn) buffer equals the 2-D linear
	 * convolution placed with the kernel at (src-krn)/2. */
	unsigned mx = W > kw ? W : kw;
	unsigned ex = 2 * mx, ey = 2 * mx;
	size_t lw = W + kw - 1, lh = H + kh - 1;
	double *L = (double *)malloc(lw * lh * sizeof(double));
	or_linear_conv_fft(src, W, H, krn, kw, kh, L);
	unsigned kx0 = (W - kw) / 2, ky0 = (H - kh) / 2;
	int ox, oy;
	or_fft_offset(W, H, kw, kh, &ox, &oy);
	/* Image::crop (include/profit/image.h:433-448) of mask_and_crop throws
	 * when the cropped block leaves the extended image: with ey = 2 max(W, kw)
	 * (the x component, include/profit/image.h:204-207) this happens for images
	 * taller than about 4/3 of their width */
	if (crop && ((unsigned)ox + W > ex || (unsigned)oy + H > ey)) {
		free(L);
		return fail(PK_EINVAL, "start_y + new_height should be <= image.height");
	}
	/* the 1-D transform of length ex * ey is circular: the rows of the
	 * placed linear convolution beyond ey wrap around to row 0 (no row
	 * overflows in x, since ex >= W + kw - 1 + (W - kw) / 2) */
	double *ext = (double *)calloc((size_t)ex * ey, sizeof(double));
	for (size_t j = 0; j < lh; j++)
		for (size_t i = 0; i < lw; i++) {
			size_t X = i + kx0, Y = (j + ky0) % ey;
			if (X < ex) ext[X + Y * ex] += L[i + j * lw];
		}
	free(L);
	/* Convolver::mask_and_crop (src/convolve.cpp:81-99) */
	if (!crop) {
		*offx = ox;
		*offy = oy;
		if (mask) {
			for (unsigned j = 0; j < ey; j++)
				for (unsigned i = 0; i < ex; i++) {
					int si = (int)i - ox, sj = (int)j - oy;
					int m = si >= 0 && sj >= 0 && si < (int)W && sj < (int)H && mask[si + (size_t)sj * W];
					if (!m) ext[i + (size_t)j * ex] = 0;
				}
		}
		*out = ext;
		*ow = ex;
		*oh = ey;
		return PK_OK;
	}
	double *o = (double *)malloc((size_t)W * H * sizeof(double));
	for (unsigned j = 0; j < H; j++)
		for (unsigned i = 0; i < W; i++) {
			double v = ext[(i + ox) + (size_t)(j + oy) * ex];
			if (mask && !mask[i + (size_t)j * W]) v = 0;
			o[i + (size_t)j * W] = v;
		}
	free(ext);
	*out = o;
	*ow = W;
	*oh = H;
	return PK_OK;
}

/* Convolver::convolve (src/convolve.cpp:55-79). */
static int or_convolve_alloc(int type, const double *src, unsigned W, unsigned H, const double *krn, unsigned kw,
                             unsigned kh, const uint8_t *mask, int crop, double **out, unsigned *ow, unsigned *oh,
                             int *offx, int *offy)
{
	*offx = *offy = 0;
	if (!(W >= kw && H >= kh)) {
		unsigned m = W > kw ? W : kw; /* max() quirk: x for both */
		unsigned nw = m, nh = m;
		if (nw < W) return fail(PK_EINVAL, "new_width should be >= width");
		if (nh < H) return fail(PK_EINVAL, "new_height should be >= height");
		double *es = (double *)calloc((size_t)nw * nh, sizeof(double));
		uint8_t *em = NULL;
		for (unsigned j = 0; j < H; j++)
			for (unsigned i = 0; i < W; i++) es[i + (size_t)j * nw] = src[i + (size_t)j * W];
		if (mask) {
			em = (uint8_t *)calloc((size_t)nw * nh, 1);
			for (unsigned j = 0; j < H; j++)
				for (unsigned i = 0; i < W; i++) em[i + (size_t)j * nw] = mask[i + (size_t)j * W];
		}
		double *r;
		unsigned rw, rh;
		int st = or_convolve_impl(type, es, nw, nh, krn, kw, kh, em, crop, &r, &rw, &rh, offx, offy);
		free(es);
		free(em);
		if (st) return st;
		unsigned cw = rw - (nw - W), ch = rh - (nh - H);
		double *c = (double *)malloc((size_t)cw * ch * sizeof(double));
		for (unsigned j = 0; j < ch; j++)
			for (unsigned i = 0; i < cw; i++) c[i + (size_t)j * cw] = r[i + (size_t)j * rw];
		free(r);
		*out = c;
		*ow = cw;
		*oh = ch;
		return PK_OK;
	}
	return or_convolve_impl(type, src, W, H, krn, kw, kh, mask, crop, out, ow, oh, offx, offy);
}

int or_convolve(int type, const double *src, unsigned W, unsigned H, const double *krn, unsigned kw, unsigned kh,
                const uint8_t *mask, int crop, double *out, size_t cap, unsigned *ow, unsigned *oh, int *offx,
                int *offy)
{
	double *r;
	int st = or_convolve_alloc(type, src, W, H, krn, kw, kh, mask, crop, &r, ow, oh, offx, offy);
	if (st) return st;
	size_t n = (size_t)(*ow) * (*oh);
	if (n > cap) {
		free(r);
		return fail(PK_ERANGE, "output buffer too small");
	}
	memcpy(out, r, n * sizeof(double));
	free(r);
	return PK_OK;
}

/* ======================================================================== */
/* Mask helpers (src/image.cpp:112-162, include/profit/image.h:399-501)      */
/* ======================================================================== */

static uint8_t *or_mask_upsample(const uint8_t *m, unsigned w, unsigned h, unsigned f)
{
	uint8_t *o = (uint8_t *)malloc((size_t)w * f * h * f);
	for (unsigned r = 0; r < h * f; r++)
		for (unsigned c = 0; c < w * f; c++) o[c + (size_t)r * w * f] = m[c / f + (size_t)(r / f) * w] ? 1 : 0;
	return o;
}

static uint8_t *or_mask_extend(const uint8_t *m, unsigned w, unsigned h, unsigned nw, unsigned nh, unsigned sx,
                               unsigned sy)
{
	uint8_t *o = (uint8_t *)calloc((size_t)nw * nh, 1);
	for (unsigned j = 0; j < h; j++)
		for (unsigned i = 0; i < w; i++) o[(i + sx) + (size_t)(j + sy) * nw] = m[i + (size_t)j * w] ? 1 : 0;
	return o;
}

/* Mask::expand_by (src/image.cpp:112-162) */
static uint8_t *or_mask_expand(const uint8_t *m, unsigned w, unsigned h, unsigned px, unsigned py)
{
	uint8_t *o = (uint8_t *)malloc((size_t)w * h);
	for (size_t t = 0; t < (size_t)w * h; t++) o[t] = m[t] ? 1 : 0;
	for (unsigned j = 0; j < h; j++)
		for (unsigned i = 0; i < w; i++) {
			if (o[i + (size_t)j * w]) continue;
			int jl = (int)j - (int)py, jh = (int)j + (int)py, il = (int)i - (int)px, ih = (int)i + (int)px;
			if (jl < 0) jl = 0;
			if (il < 0) il = 0;
			if (jh > (int)h - 1) jh = (int)h - 1;
			if (ih > (int)w - 1) ih = (int)w - 1;
			int hit = 0;
			for (int y = jl; y <= jh && !hit; y++)
				for (int x = il; x <= ih; x++)
					if (m[x + (size_t)y * w]) {
						hit = 1;
						break;
					}
			if (hit) o[i + (size_t)j * w] = 1;
		}
	return o;
}

/* Mask::bounding_box (include/profit/image.h:451-483) */
static void or_mask_bbox(const uint8_t *m, unsigned w, unsigned h, unsigned *lx, unsigned *ly, unsigned *ux,
                         unsigned *uy)
{
	unsigned lbx = w, lby = h, ubx = 0, uby = 0;
	int any = 0;
	for (unsigned j = 0; j < h; j++)
		for (unsigned i = 0; i < w; i++) {
			if (!m[i + (size_t)j * w]) continue;
			if (i < lbx) lbx = i;
			if (j < lby) lby = j;
			if (i + 1 > ubx) ubx = i + 1;
			if (j + 1 > uby) uby = j + 1;
			any = 1;
		}
	if (!any) lbx = lby = ubx = uby = 0;
	*lx = lbx;
	*ly = lby;
	*ux = ubx;
	*uy = uby;
}

/* ======================================================================== */
/* Model::evaluate (src/model.cpp:152-360)                                    */
/* ======================================================================== */

typedef struct {
	unsigned dw, dh;     /* drawing dims */
	unsigned padx, pady; /* psf padding */
	int conv_required, mask_needs_psf_padding, mask_needs_convolution, mask_needs_adjustment;
} or_analysis;

static int or_has_mask(const pk_model *m) { return m->mask && m->mask_width > 0 && m->mask_height > 0; }
static int or_has_psf(const pk_model *m) { return m->psf && m->psf_width > 0 && m->psf_height > 0; }

/* analyze_expansion_requirements (src/model.cpp:152-184) */
static void or_expansion(const pk_model *m, or_analysis *a)
{
	unsigned fs = m->finesampling;
	a->mask_needs_convolution = 0;
	a->mask_needs_psf_padding = 0;
	int need_pad = a->conv_required;
	if (or_has_mask(m) && !m->adjust_mask) {
		need_pad = m->mask_width > m->width * fs && m->mask_height > m->height * fs;
	} else if (or_has_mask(m) && a->conv_required) {
		unsigned lx, ly, ux, uy;
		or_mask_bbox(m->mask, m->mask_width, m->mask_height, &lx, &ly, &ux, &uy);
		lx *= fs;
		ly *= fs;
		ux *= fs;
		uy *= fs;
		unsigned upx = m->mask_width * fs - ux, upy = m->mask_height * fs - uy;
		unsigned nx = m->psf_width / 2, ny = m->psf_height / 2;
		need_pad = !(lx >= nx && ly >= ny) || !(upx >= nx && upy >= ny);
		a->mask_needs_convolution = 1;
		a->mask_needs_psf_padding = need_pad;
	}
	if (need_pad) {
		a->padx = m->psf_width / 2;
		a->pady = m->psf_height / 2;
	} else {
		a->padx = a->pady = 0;
	}
	a->mask_needs_adjustment =
		or_has_mask(m) && (a->mask_needs_convolution || a->mask_needs_psf_padding || fs > 1);
	a->dw = m->width * fs + 2 * a->padx;
	a->dh = m->height * fs + 2 * a->pady;
}

/* analyze_inputs (src/model.cpp:186-219) */
static int or_analyze(const pk_model *m, or_analysis *a)
{
	if (!(m->width > 0 && m->height > 0)) return fail(PK_EINVAL, "Model's requested dimensions are 0");
	if (m->scale_x <= 0) return fail(PK_EINVAL, "Model's scale_x cannot be negative or zero");
	if (m->scale_y <= 0) return fail(PK_EINVAL, "Model's scale_y cannot be negative or zero");
	if (m->finesampling == 0) return fail(PK_EINVAL, "finesampling must be > 0");
	if (or_has_mask(m) && m->adjust_mask && (m->mask_width != m->width || m->mask_height != m->height))
		return fail(PK_EINVAL, "Mask dimensions != model dimensions");
	a->conv_required = 0;
	for (unsigned c = 0; c < m->n_components; c++)
		if (m->components[c].convolve) a->conv_required = 1;
	if (a->conv_required && !or_has_psf(m))
		return fail(PK_EINVAL, "No psf provided but profile(s) requested convolution");
	for (unsigned c = 0; c < m->n_components; c++) {
		int st = or_validate(&m->components[c], or_has_psf(m));
		if (st) return st;
	}
	or_expansion(m, a);
	return PK_OK;
}

static uint64_t or_last_refined, or_last_subcells;
uint64_t or_stat_refined(void) { return or_last_refined; }
uint64_t or_stat_subcells(void) { return or_last_subcells; }

int or_evaluate(const pk_model *m, double *out, size_t cap, unsigned *ow, unsigned *oh, int *offx, int *offy)
{
	or_analysis a;
	int st = or_analyze(m, &a);
	if (st) return st;
	or_last_refined = or_last_subcells = 0;
	unsigned fs = m->finesampling;
	if (m->dry_run) {
		size_t n = (size_t)a.dw * a.dh;
		if (n > cap) return fail(PK_ERANGE, "output buffer too small");
		memset(out, 0, n * sizeof(double));
		*ow = a.dw;
		*oh = a.dh;
		*offx = *offy = 0;
		return PK_OK;
	}
	/* normalised psf (Model::set_psf -> Image::normalize, src/image.cpp:208-221) */
	double *psf = NULL;
	if (or_has_psf(m)) {
		size_t pn = (size_t)m->psf_width * m->psf_height;
		psf = (double *)malloc(pn * sizeof(double));
		double sum = 0;
		for (size_t t = 0; t < pn; t++) sum += m->psf[t];
		for (size_t t = 0; t < pn; t++) psf[t] = sum > 0 ? m->psf[t] / sum : m->psf[t];
	}

	/* mask adjustment (src/model.cpp:230-246, 265-285) */
	uint8_t *amask = NULL;
	const uint8_t *use_mask = NULL;
	int adjusted = 0;
	if (or_has_mask(m)) {
		if (m->adjust_mask && a.mask_needs_adjustment) {
			unsigned w = m->mask_width, h = m->mask_height;
			uint8_t *cur = (uint8_t *)malloc((size_t)w * h);
			for (size_t t = 0; t < (size_t)w * h; t++) cur[t] = m->mask[t] ? 1 : 0;
			if (fs > 1) {
				uint8_t *n = or_mask_upsample(cur, w, h, fs);
				free(cur);
				cur = n;
				w *= fs;
				h *= fs;
			}
			if (a.mask_needs_psf_padding) {
				uint8_t *n = or_mask_extend(cur, w, h, a.dw, a.dh, a.padx, a.pady);
				free(cur);
				cur = n;
				w = a.dw;
				h = a.dh;
			}
			if (a.mask_needs_convolution) {
				uint8_t *n = or_mask_expand(cur, w, h, m->psf_width / 2, m->psf_height / 2);
				free(cur);
				cur = n;
			}
			amask = cur;
			use_mask = amask;
			adjusted = 1;
		} else {
			if (!m->adjust_mask && (m->mask_width != a.dw || m->mask_height != a.dh)) {
				char buf[200];
				snprintf(buf, sizeof(buf), "Mask dimensions != drawing dimensions: [%u, %u] != [%u, %u]",
				         m->mask_width, m->mask_height, a.dw, a.dh);
				free(psf);
				return fail(PK_EINVAL, buf);
			}
			use_mask = m->mask;
		}
	}

	/* produce_image (src/model.cpp:322-360) */
	size_t dn = (size_t)a.dw * a.dh;
	double *model_img = (double *)calloc(dn, sizeof(double));
	double *conv_img = a.conv_required ? (double *)calloc(dn, sizeof(double)) : NULL;
	double sx = m->scale_x / fs, sy = m->scale_y / fs;
	for (unsigned c = 0; c < m->n_components; c++) {
		const pk_component *pc = &m->components[c];
		double *target = pc->convolve ? conv_img : model_img;
		if (pc->kind == PK_PROFILE_NULL) continue;
		if (pc->kind == PK_PROFILE_SKY) {
			double bg = pc->p[PK_SKY_BG] / (fs * fs);
			for (size_t t = 0; t < dn; t++)
				if (!use_mask || use_mask[t]) target[t] += bg;
			continue;
		}
		if (pc->kind == PK_PROFILE_PSF) {
			or_psf_evaluate(pc, psf, m->psf_width, m->psf_height, m->psf_scale_x, m->psf_scale_y, target, a.dw,
			                a.dh, sx, sy, a.padx, a.pady, m->magzero);
			continue;
		}
		or_prof p;
		or_setup(&p, pc, m, sx, sy, a.padx, a.pady, fs);
		or_radial_evaluate(&p, target, a.dw, a.dh, use_mask, sx, sy, &or_last_refined, &or_last_subcells);
	}
	unsigned iw = a.dw, ih = a.dh;
	int ox_ = 0, oy_ = 0;
	if (a.conv_required) {
		double *cv;
		unsigned cw, ch;
		st = or_convolve_alloc(m->convolver, conv_img, a.dw, a.dh, psf, m->psf_width, m->psf_height, use_mask,
		                       m->crop, &cv, &cw, &ch, &ox_, &oy_);
		free(conv_img);
		if (st) {
			free(model_img);
			free(psf);
			free(amask);
			return st;
		}
		if (cw != a.dw || ch != a.dh) {
			double *e = (double *)calloc((size_t)cw * ch, sizeof(double));
			for (unsigned j = 0; j < a.dh; j++)
				for (unsigned i = 0; i < a.dw; i++)
					e[(i + ox_) + (size_t)(j + oy_) * cw] = model_img[i + (size_t)j * a.dw];
			free(model_img);
			model_img = e;
		}
		for (size_t t = 0; t < (size_t)cw * ch; t++) model_img[t] += cv[t];
		free(cv);
		iw = cw;
		ih = ch;
	}
	free(psf);

	/* crop (src/model.cpp:289-306) */
	if (a.padx > 0 && a.pady > 0) {
		int cox = (int)a.padx, coy = (int)a.pady;
		int cdw = (int)a.dw - 2 * (int)a.padx, cdh = (int)a.dh - 2 * (int)a.pady;
		if (!m->crop) {
			int alx, aly, ahx, ahy, ilx, ily, ihx, ihy;
			int ct = a.conv_required ? m->convolver : PK_CONV_BRUTE;
			or_convolver_padding(ct, a.dw, a.dh, m->psf_width, m->psf_height, &alx, &aly, &ahx, &ahy);
			or_convolver_padding(ct, a.dw - 2 * a.padx, a.dh - 2 * a.pady, m->psf_width, m->psf_height, &ilx,
			                     &ily, &ihx, &ihy);
			int odx = alx - ilx, ody = aly - ily, ddx = ahx - ihx, ddy = ahy - ihy;
			cdw = (int)iw - 2 * (int)a.padx - (odx + ddx);
			cdh = (int)ih - 2 * (int)a.pady - (ody + ddy);
			cox += odx;
			coy += ody;
			ox_ -= odx;
			oy_ -= ody;
		}
		double *cimg = (double *)malloc((size_t)cdw * cdh * sizeof(double));
		for (int j = 0; j < cdh; j++)
			for (int i = 0; i < cdw; i++)
				cimg[i + (size_t)j * cdw] = model_img[(i + cox) + (size_t)(j + coy) * iw];
		free(model_img);
		model_img = cimg;
		iw = cdw;
		ih = cdh;
	}
	/* downsample SUM (src/model.cpp:308-311, src/image.cpp:262-276) */
	if (fs > 1 && !m->return_finesampled) {
		unsigned nw = (iw + fs - 1) / fs, nh = (ih + fs - 1) / fs;
		double *d = (double *)calloc((size_t)nw * nh, sizeof(double));
		for (unsigned r = 0; r < ih; r++)
			for (unsigned c = 0; c < iw; c++) d[c / fs + (size_t)(r / fs) * nw] += model_img[c + (size_t)r * iw];
		free(model_img);
		model_img = d;
		iw = nw;
		ih = nh;
		ox_ /= (int)fs;
		oy_ /= (int)fs;
	}
	/* final masking with the user's mask (src/model.cpp:314-316); with a
	 * finesampled result the mask is applied upsampled */
	if (adjusted) {
		unsigned f = (iw == m->mask_width) ? 1 : fs;
		for (unsigned j = 0; j < ih; j++)
			for (unsigned i = 0; i < iw; i++) {
				unsigned mi = i / f, mj = j / f;
				if (mi < m->mask_width && mj < m->mask_height && !m->mask[mi + (size_t)mj * m->mask_width])
					model_img[i + (size_t)j * iw] = 0;
			}
	}
	free(amask);
	size_t n = (size_t)iw * ih;
	if (n > cap) {
		free(model_img);
		return fail(PK_ERANGE, "output buffer too small");
	}
	memcpy(out, model_img, n * sizeof(double));
	free(model_img);
	*ow = iw;
	*oh = ih;
	*offx = ox_;
	*offy = oy_;
	return PK_OK;
}

int or_output_dims(const pk_model *m, unsigned *w, unsigned *h)
{
	or_analysis a;
	int st = or_analyze(m, &a);
	if (st) return st;
	*w = a.dw;
	*h = a.dh;
	return PK_OK;
}

/* Per-component constants after set-up, for tests of the set-up stage:
 * out[0..11] = rscale, ie, acc, rscale_switch, rscale_max, resolution,
 * bn, flux_scale(scale 1), cos_ang, sin_ang, rescale_factor, lumtot-free ie*1 */
int or_component_setup(const pk_model *m, unsigned c, double *out)
{
	or_prof p;
	if (c >= m->n_components) return fail(PK_EINVAL, "component index out of range");
	or_setup(&p, &m->components[c], m, m->scale_x, m->scale_y, 0, 0, 1);
	out[0] = p.rscale;
	out[1] = p.ie;
	out[2] = p.acc;
	out[3] = p.rscale_switch;
	out[4] = p.rscale_max;
	out[5] = p.resolution;
	out[6] = p.bn;
	out[7] = p.flux_scale;
	out[8] = p.cos_ang;
	out[9] = p.sin_ang;
	out[10] = p.rescale_factor;
	return PK_OK;
}

/* evaluate_at of one component after set-up (no padding) */
double or_component_eval_at(const pk_model *m, unsigned c, double x, double y)
{
	or_prof p;
	or_setup(&p, &m->components[c], m, m->scale_x, m->scale_y, 0, 0, 1);
	return or_evaluate_at(&p, x, y);
}
