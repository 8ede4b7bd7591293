// Radial-profile constants and evaluation, shared by host set-up and device
// kernels.
//
// PkComp holds everything libprofit's RadialProfile computes before its pixel
// loop (RadialProfile::initial_calculations, src/radial.cpp:137-199, and the
// profile overrides), in FP64, plus FP32 copies for the FP32 bulk path.
// Sersic values in the FP32 path are computed as exp(-bn * rf): the factor
// exp(bn) is folded into `unfold` (FP64), which keeps the FP32 exponent
// argument below ~14 for every pixel brighter than 1e-6 of the peak.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pk_math.h"
#include "profit_hip.h"

namespace pk {

enum SersicVariant { SV_GENERAL = 0, SV_HALF, SV_ONE, SV_TWO, SV_THREE, SV_FOUR, SV_EIGHT, SV_SIXTEEN };

// What the FP32 pixel pass reads of a component, in one 96-byte block (two
// scalar loads): the FP32 classification constants, the FP32 profile
// constants and the FP64 centre.
//
// The classification compares the FP32 r^2 with guard bands [lo, hi] around
// r2_switch and r2_max (set by fill_pix_hot): r^2 > hi, r^2 < lo decide, a
// value within a band (or NaN) goes to the FP64 test.
enum PixFlags : int32_t {
	PIX_FP32 = 1,     // FP32 profile form
	PIX_CONVOLVE = 2  // added to the image that is convolved
};

struct alignas(16) PixHot {
	double xcen, ycen;
	double vscale;                 // unfold * flux_scale (sky: the pixel value)
	float cos32, sin32, iq32;      // cos_ang, sin_ang, inv_axrat
	float lo_sw, hi_sw;            // band of r2_switch: above hi direct, below lo refined
	float lo_mx, hi_mx;            // band of r2_max: above hi zero
	float f[8];                    // PkComp::f
	int32_t kind, boxy, flags;     // flags: PixFlags
};

struct PkComp {
	int32_t kind;       // PK_PROFILE_*
	int32_t convolve;
	int32_t rough;
	int32_t adjust;
	int32_t res, max_rec;        // sub-pixel resolution / max recursions after adjustment
	int32_t center_res, center_max_rec;  // Sersic centre-pixel override (src/sersic.cpp:331-347), 0 = none
	int32_t sv, boxy;            // Sersic evaluation variant (src/sersic.cpp:225-255)
	int32_t fp32;                // evaluated in FP32 in the FP32 mode
	int32_t pad_;
	double xcen, ycen;           // _xcen/_ycen: centre in the drawing frame (src/radial.cpp:254-255)
	double cos_ang, sin_ang, axrat, box;
	double rscale, rscale_switch, rscale_max, acc;
	double flux_scale;           // pixel area * Ie (* rescale factor), src/radial.cpp:219-222
	double unfold;               // FP32 values are multiplied by this to reach libprofit's units
	double win_x, win_y;         // Model image pixel scale (Sersic centre window, src/sersic.cpp:340-341)
	double q[8];                 // profile constants (FP64)
	float f[8];                  // profile constants (FP32)
	// values after adjustment, reported back (libprofit writes them into the profile)
	double bn, lumtot, ie;
	double inv_axrat;            // 1 / axrat (FP32-mode refinement)
	double total_flux;           // 10^(-0.4 (mag - magzero)), bound of any pixel value
	// gradient constants (set-up with SetupIn::grad): d bn / d nser and the
	// derivatives of log(flux_scale) with respect to re, nser, axrat, box
	double dbn, glnf_re, glnf_n, glnf_q, glnf_b;
	// squared radii of the pixel classification: (rscale_max rscale)^2 and
	// (rscale_switch rscale)^2
	double r2_max, r2_switch;
	PixHot hot;
};

// libprofit's test sqrt(x^2 + y^2) / rscale > limit (src/radial.cpp:316-321),
// decided from r2 = x^2 + y^2 against r2_limit = (limit rscale)^2; only within
// a relative 1e-10 of the limit (far wider than the rounding of either form)
// is the reference expression itself evaluated, so the decision is the same.
PK_HD bool radius_exceeds(double r2, double r2_limit, double rscale, double limit)
{
	if (r2 > r2_limit * (1 + 1e-10)) return true;
	if (r2 < r2_limit * (1 - 1e-10)) return false;
	return sqrt(r2) / rscale > limit;
}

// What one FP32 sub-cell evaluation of the refinement reads: one 128-byte
// record (a single cache line) per component.
struct alignas(128) CompLite {
	double cos_ang, sin_ang, inv_axrat;
	double wflux;    // flux_scale * fixed-point scale of the job (weight of an FP64 value)
	double wq;       // wflux * unfold (weight of an FP32 value)
	double acc;
	double hsx, hsy; // half the pixel scale of the job
	double q4, q5;   // king / ferrer FP64 constants (PkComp::q[4], q[5])
	double b2;       // box + 2
	float f[8];
	int8_t kind, boxy, fp32, convolve;
	int32_t pad_;
};

// A CompLite whose box flag is a compile-time constant: the FP32 profile
// functions read c.boxy, so their round / boxy branches fold away.
template <bool B>
struct CompLiteB : CompLite {
	static constexpr bool boxy = B;
	PK_HD explicit CompLiteB(const CompLite &c) : CompLite(c) {}
};

PK_HD bool almost_equals(double x, double y) { return fabs(x - y) < 1e-10; }

// ---------------------------------------------------------------------------
// FP64 evaluation, operation for operation as libprofit's CPU code.
// ---------------------------------------------------------------------------

PK_HD double boxy_r(const PkComp &c, double x, double y)
{
	// include/profit/radial.h:93-102
	if (c.box == 0) return sqrt(x * x + y * y);
	double b2 = c.box + 2.;
	return pow(pow(fabs(x), b2) + pow(fabs(y), b2), 1. / b2);
}

PK_HD double sersic_at(const PkComp &c, double x, double y)
{
	// src/sersic.cpp:79-205
	const double re = c.q[0], bn = c.q[2], invexp = c.q[3];
	double base, rf;
	if (c.boxy) {
		double B = c.box + 2;
		base = pow(fabs(x / re), B) + pow(fabs(y / re), B);
		switch (c.sv) {
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
		base = (x * x + y * y) / (re * re);
		switch (c.sv) {
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
	return exp(-bn * (rf - 1));
}

PK_HD double broken_exponential(double r, double h1, double h2, double rb, double a)
{
	// src/brokenexponential.cpp:39-76
	double base = r - rb;
	double expo = 1 / h1 - 1 / h2;
	if (a * base < 40) base = log(1 + exp(a * base)) / a;
	return exp(-r / h1 + expo * base);
}

PK_HD double king_r(const PkComp &c, double r)
{
	// src/king.cpp:51-59
	const double rc = c.q[0], rt = c.q[1], a = c.q[2];
	if (r < rt) return pow(1 / pow(1 + pow(r / rc, 2), 1 / a) - 1 / pow(1 + pow(rt / rc, 2), 1 / a), a);
	return 0;
}

PK_HD double coresersic_r(const PkComp &c, double r)
{
	// src/coresersic.cpp:51-60
	const double re = c.q[0], rb = c.q[1], nser = c.q[2], a = c.q[3], b = c.q[4], bn = c.q[5];
	return pow(1 + pow(r / rb, -a), b / a) * exp(-bn * pow((pow(r, a) + pow(rb, a)) / pow(re, a), 1 / (nser * a)));
}

PK_HD double evaluate_at(const PkComp &c, double x, double y)
{
	switch (c.kind) {
	case PK_PROFILE_SERSIC: return sersic_at(c, x, y);
	case PK_PROFILE_MOFFAT: {
		// src/moffat.cpp:52-58
		double r = boxy_r(c, x, y);
		double rf = r / c.rscale;
		return pow(1 + rf * rf, -c.q[1]);
	}
	case PK_PROFILE_FERRER: {
		// src/ferrer.cpp:50-61
		double r = boxy_r(c, x, y);
		double rf = r / c.rscale;
		if (rf < 1) return pow(1 - pow(rf, 2 - c.q[2]), c.q[1]);
		return 0;
	}
	case PK_PROFILE_KING: return king_r(c, boxy_r(c, x, y));
	case PK_PROFILE_BROKENEXP: return broken_exponential(boxy_r(c, x, y), c.q[0], c.q[1], c.q[2], c.q[3]);
	case PK_PROFILE_CORESERSIC: return coresersic_r(c, boxy_r(c, x, y));
	default: return 0;
	}
}

// ---------------------------------------------------------------------------
// FP32 evaluation (all radial kinds).  Values are in units of 1/unfold of the
// FP64 ones.  Built on the hardware transcendental instructions (V_EXP_F32,
// V_LOG_F32, V_RCP_F32, V_SQRT_F32: about 1 ulp; results below 2^-126 flush
// to 0); log1p, expm1 and exp are composed from them so that no accuracy is
// lost near 0.  A value that is not finite is left to the FP64 evaluation by
// the callers.
// ---------------------------------------------------------------------------

__device__ inline float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ inline float flog2(float x) { return __builtin_amdgcn_logf(x); }
__device__ inline float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

constexpr float F_LN2 = 0.693147180559945309f;

// e^x; the rounding error of x log2(e) is carried to first order
__device__ inline float fexp(float x)
{
	constexpr float L2E_HI = 1.44269502162933349609375f;  // log2(e) = L2E_HI + L2E_LO
	constexpr float L2E_LO = 1.925963033500011079e-8f;
	const float h = x * L2E_HI;
	const float l = fmaf(x, L2E_HI, -h) + x * L2E_LO;
	return fexp2(h) * fmaf(l, F_LN2, 1.f);
}

// log(1 + x), x > -1: log(u) x / (u - 1) with u = 1 + x rounded (W. Kahan)
__device__ inline float flog1p(float x)
{
	const float u = 1.f + x;
	const float d = u - 1.f;
	return d == 0.f ? x : flog2(u) * F_LN2 * (x * frcp(d));
}

// e^x - 1: (u - 1) x / log(u) with u = e^x rounded (W. Kahan)
__device__ inline float fexpm1(float x)
{
	const float u = fexp(x);
	const float um1 = u - 1.f;
	if (um1 == 0.f) return x;
	if (um1 == -1.f) return -1.f;
	return um1 * x * frcp(flog2(u) * F_LN2);
}

PK_HD double kq4(const PkComp &c) { return c.q[4]; }
PK_HD double kq5(const PkComp &c) { return c.q[5]; }
PK_HD double kb2(const PkComp &c) { return c.box + 2.; }
PK_HD double kq4(const CompLite &c) { return c.q4; }
PK_HD double kq5(const CompLite &c) { return c.q5; }
PK_HD double kb2(const CompLite &c) { return c.b2; }

// Boxy radius (|x|^b2 + |y|^b2)^(1/b2) (include/profit/radial.h:93-102),
// b2 = c.f[7], written as m (1 + (s/m)^b2)^(1/b2) with m = max(|x|, |y|),
// s = min(|x|, |y|): the powers stay within [0, 2], so that the FP32 result
// is good to a few ulps at any radius.  Returns m and log2(1 + (s/m)^b2) / b2.
template <typename CR>
__device__ inline float boxy_parts32(const CR &c, float x, float y, float &m)
{
	const float ax = fabsf(x), ay = fabsf(y);
	m = fmaxf(ax, ay);
	const float q = m > 0.f ? fminf(ax, ay) * frcp(m) : 0.f;
	const float b2 = c.f[7];
	return flog2(1.f + fexp2(b2 * flog2(q))) * frcp(b2);
}

template <typename CR>
__device__ inline float boxy_r32(const CR &c, float x, float y)
{
	float m;
	const float l = boxy_parts32(c, x, y, m);
	return m * fexp2(l);
}

// log2 of the radius (round or boxy)
template <typename CR>
__device__ inline float log2_r32(const CR &c, float x, float y)
{
	if (!c.boxy) return 0.5f * flog2(x * x + y * y);
	float m;
	const float l = boxy_parts32(c, x, y, m);
	return flog2(m) + l;
}

// Squared boxy radius (|x|^b2 + |y|^b2)^(2/b2) in FP64.  Kept out of line:
// its three FP64 pow calls, inlined at every call site, would raise the
// register count (and code size) of every kernel that can reach it.
PK_HD __attribute__((noinline)) double boxy_radius2_64(double x, double y, double b2)
{
	const double r = pow(pow(fabs(x), b2) + pow(fabs(y), b2), 1. / b2);
	return r * r;
}

// FP64 squared (boxy) radius, for the FP32 king and ferrer evaluation whose
// value near the truncation radius depends on rt^2 - r^2.
template <typename CR>
PK_HD double radius2_64_kf(const CR &c, double x, double y)
{
	if (!c.boxy) return x * x + y * y;
	return boxy_radius2_64(x, y, kb2(c));
}

// radius2_64_kf for king and ferrer, 0 for the other kinds
template <typename CR>
PK_HD double radius2_64(const CR &c, double x, double y)
{
	if (c.kind != PK_PROFILE_KING && c.kind != PK_PROFILE_FERRER) return 0;
	return radius2_64_kf(c, x, y);
}

template <typename CR>
__device__ inline float king32(const CR &c, double r2)
{
	// f[0] = 1/rc^2, f[1] = 1/a, f[2] = a; q[4] = rt^2, q[5] = 1 / (rc^2 + rt^2).
	// (t(r) - t(rt))^a with t(r) = (1 + r^2/rc^2)^(-1/a), written as
	// t(r) (1 - (1 - d)^(1/a)), d = (rt^2 - r^2) / (rc^2 + rt^2), so that
	// no two nearly equal numbers are subtracted near rt.
	if (!(r2 < kq4(c))) return 0.f;
	const float lt = -c.f[1] * flog2(1.f + (float)r2 * c.f[0]);
	const float d = (float)((kq4(c) - r2) * kq5(c));
	const float diff = -fexpm1(flog1p(-d) * c.f[1]);
	if (!(diff > 0.f)) return c.f[2] == 0.f ? 1.f : 0.f;
	return fexp2(c.f[2] * (lt + flog2(diff)));
}

template <typename CR>
__device__ inline float ferrer32(const CR &c, double r2)
{
	// f[0] = a, f[1] = (2 - b) / 2; q[4] = rout^2, q[5] = 1 / rout^2.
	// 1 - (r/rout)^(2-b) = -expm1((2-b)/2 log1p(-d)), d = 1 - r^2/rout^2
	if (!(r2 < kq4(c))) return 0.f;
	const float d = (float)((kq4(c) - r2) * kq5(c));
	const float base = -fexpm1(c.f[1] * flog1p(-d));
	if (!(base > 0.f)) return c.f[0] == 0.f ? 1.f : 0.f;
	return fexp2(c.f[0] * flog2(base));
}

template <typename CR>
__device__ inline float sersic32(const CR &c, float x, float y)
{
	// f[0] = -log2(re), f[1] = 1/nser, f[2] = bn, f[7] = box + 2: exp(-bn (r/re)^(1/nser))
	const float rf = fexp2(c.f[1] * (log2_r32(c, x, y) + c.f[0]));
	return fexp(-c.f[2] * rf);
}

template <typename CR>
__device__ inline float moffat32(const CR &c, float x, float y)
{
	// f[0] = 1/rscale^2, f[1] = con
	float r2f;
	if (!c.boxy) {
		r2f = (x * x + y * y) * c.f[0];
	} else {
		const float r = boxy_r32(c, x, y);
		r2f = r * r * c.f[0];
	}
	return fexp2(-c.f[1] * flog2(1.f + r2f));
}

template <typename CR>
__device__ inline float brokenexp32(const CR &c, float x, float y)
{
	// f[0] = 1/h1, f[1] = 1/h1 - 1/h2, f[2] = rb, f[3] = a, f[4] = 1/a
	const float r = c.boxy ? boxy_r32(c, x, y) : fsqrt(x * x + y * y);
	float base = r - c.f[2];
	if (c.f[3] * base < 40.f) base = flog1p(fexp(c.f[3] * base)) * c.f[4];
	return fexp(-r * c.f[0] + c.f[1] * base);
}

template <typename CR>
__device__ inline float coresersic32(const CR &c, float x, float y)
{
	// f[0] = log2(1/rb), f[1] = a, f[2] = b/a, f[3] = bn, f[4] = rb^a, f[5] = 1/re^a, f[6] = 1/(nser a)
	const float lr = log2_r32(c, x, y);
	const float a = c.f[1];
	const float ra = fexp2(a * lr);
	const float core = fexp2(c.f[2] * flog2(1.f + fexp2(-a * (lr + c.f[0]))));
	const float sers = fexp(-c.f[3] * fexp2(c.f[6] * flog2((ra + c.f[4]) * c.f[5])));
	return core * sers;
}

// `r2` is radius2_64 at the same point.
template <typename CR>
__device__ inline float evaluate_at32(const CR &c, float x, float y, double r2)
{
	switch (c.kind) {
	case PK_PROFILE_KING: return king32(c, r2);
	case PK_PROFILE_FERRER: return ferrer32(c, r2);
	case PK_PROFILE_SERSIC: return sersic32(c, x, y);
	case PK_PROFILE_MOFFAT: return moffat32(c, x, y);
	case PK_PROFILE_BROKENEXP: return brokenexp32(c, x, y);
	case PK_PROFILE_CORESERSIC: return coresersic32(c, x, y);
	default: return 0.f;
	}
}

// Values of a profile of kind `kind` at profile coordinates (xa, ya) and, if
// PAIR, at (xb, yb); both evaluations sit in one branch, so that they
// interleave.
template <bool PAIR, typename CR>
__device__ inline void evaluate_pair32(const CR &c, int kind, double xa, double ya, double xb, double yb, float &va,
                                       float &vb)
{
	const float x1 = (float)xa, y1 = (float)ya, x2 = (float)xb, y2 = (float)yb;
	switch (kind) {
	case PK_PROFILE_KING:
		va = king32(c, radius2_64_kf(c, xa, ya));
		if (PAIR) vb = king32(c, radius2_64_kf(c, xb, yb));
		break;
	case PK_PROFILE_FERRER:
		va = ferrer32(c, radius2_64_kf(c, xa, ya));
		if (PAIR) vb = ferrer32(c, radius2_64_kf(c, xb, yb));
		break;
	case PK_PROFILE_SERSIC:
		va = sersic32(c, x1, y1);
		if (PAIR) vb = sersic32(c, x2, y2);
		break;
	case PK_PROFILE_MOFFAT:
		va = moffat32(c, x1, y1);
		if (PAIR) vb = moffat32(c, x2, y2);
		break;
	case PK_PROFILE_BROKENEXP:
		va = brokenexp32(c, x1, y1);
		if (PAIR) vb = brokenexp32(c, x2, y2);
		break;
	case PK_PROFILE_CORESERSIC:
		va = coresersic32(c, x1, y1);
		if (PAIR) vb = coresersic32(c, x2, y2);
		break;
	default:
		va = -1.f;  // no FP32 form
		vb = -1.f;
	}
}

// ---------------------------------------------------------------------------
// Gradients of the Sersic profile value for the chi^2 adjoint.
//
// With x', y' the profile coordinates (src/radial.cpp:47-53),
//   x' = dx cos + dy sin,  y' = (-dx sin + dy cos) / q,  dx = x - xcen, dy = y - ycen,
// R = (|x'|^B + |y'|^B)^(1/B) / re (B = box + 2) and rf = R^(1/n), the value is
// f = exp(-bn (rf - 1)) (src/sersic.cpp:79-205), and with L = log f,
//   dL/dθ = -bn' (rf - 1) [θ = n] + K d log R / dθ,  K = -bn rf / n.
// The functions return the value and dL/dθ for θ in the order of the
// GradTerm enum (phi = angle in radians; the caller converts to degrees).
// The derivatives of log(flux_scale) are the per-component glnf_* constants.
// ---------------------------------------------------------------------------
enum GradTerm { GT_XC = 0, GT_YC, GT_RE, GT_N, GT_PHI, GT_Q, GT_B, GT_COUNT };

// Raw adjoint accumulators of one component: GA_W sums the weights
// a = G w flux_scale f, GA_<θ> sums a dL/dθ.
enum GradAcc { GA_XC = 0, GA_YC, GA_W, GA_RE, GA_N, GA_PHI, GA_Q, GA_B, GA_SLOTS };

// FP32: f[0] = -log2 re, f[1] = 1/n, f[2] = bn, f[3] = dbn/dn, f[4] = 1/re,
// f[7] = B; value in units of 1/unfold = e^-bn.
template <typename CR>
__device__ inline float sersic_grad32(const CR &c, float x, float y, float g[GT_COUNT])
{
	const float ax = fabsf(x), ay = fabsf(y);
	const float m = fmaxf(ax, ay);
	const float qq = m > 0.f ? fminf(ax, ay) * frcp(m) : 0.f;
	const bool xmax = ax >= ay;
	float l2R, dlx, dly, dlb;  // log2 R; d log R / d x', d y', d B
	if (!c.boxy) {
		const float s2 = x * x + y * y;
		const float inv = s2 > 0.f ? frcp(s2) : 0.f;
		l2R = 0.5f * flog2(s2) + c.f[0];
		dlx = x * inv;
		dly = y * inv;
		// B = 2: -log(1 + q^2) / 4 + q^2 log q / (2 (1 + q^2))
		const float q2 = qq * qq;
		const float lq = qq > 0.f ? flog2(qq) * F_LN2 : 0.f;
		dlb = -0.25f * flog1p(q2) + 0.5f * q2 * lq * frcp(1.f + q2);
	} else {
		// R re = m (1 + t)^(1/B), t = q^B, m = max(|x'|, |y'|), q = min/max
		const float B = c.f[7];
		const float lq2 = qq > 0.f ? flog2(qq) : -INFINITY;
		const float t = qq > 0.f ? fexp2(B * lq2) : 0.f;
		const float l1t = flog1p(t);
		const float iB = frcp(B);
		l2R = (m > 0.f ? flog2(m) : -INFINITY) + l1t * (iB / F_LN2) + c.f[0];
		const float den = m > 0.f ? frcp(m * (1.f + t)) : 0.f;
		const float wmin = qq > 0.f ? t * frcp(qq) : 0.f;  // q^(B-1)
		dlx = copysignf((xmax ? 1.f : wmin) * den, x);
		dly = copysignf((xmax ? wmin : 1.f) * den, y);
		dlb = -l1t * iB * iB + (qq > 0.f ? t * (lq2 * F_LN2) * frcp(B * (1.f + t)) : 0.f);
	}
	const float rf = fexp2(c.f[1] * l2R);
	const float v = fexp(-c.f[2] * rf);
	const float K = -c.f[2] * rf * c.f[1];
	const float cs = (float)c.cos_ang, sn = (float)c.sin_ang, iq = (float)c.inv_axrat, q = 1.f / iq;
	const float Lx = K * dlx, Ly = K * dly;
	g[GT_XC] = -cs * Lx + sn * iq * Ly;
	g[GT_YC] = -sn * Lx - cs * iq * Ly;
	g[GT_RE] = -K * c.f[4];
	g[GT_N] = -c.f[3] * (rf - 1.f) - K * (l2R * F_LN2) * c.f[1];
	g[GT_PHI] = q * y * Lx - x * iq * Ly;
	g[GT_Q] = -y * iq * Ly;
	g[GT_B] = K * dlb;
	return v;
}

// FP64 counterpart (libprofit's units: value exp(-bn (rf - 1))).
PK_HD double sersic_grad64(const PkComp &c, double x, double y, double g[GT_COUNT])
{
	const double re = c.q[0], n = c.q[1], bn = c.q[2], B = c.box + 2;
	const double ax = fabs(x), ay = fabs(y);
	const double m = fmax(ax, ay);
	const double qq = m > 0 ? fmin(ax, ay) / m : 0;
	const bool xmax = ax >= ay;
	double lR, dlx, dly, dlb;
	if (!c.boxy) {
		const double s2 = x * x + y * y;
		const double inv = s2 > 0 ? 1 / s2 : 0;
		lR = 0.5 * log(s2) - log(re);
		dlx = x * inv;
		dly = y * inv;
		const double q2 = qq * qq;
		const double lq = qq > 0 ? log(qq) : 0;
		dlb = -0.25 * log1p(q2) + 0.5 * q2 * lq / (1 + q2);
	} else {
		const double lq = qq > 0 ? log(qq) : -INFINITY;
		const double t = qq > 0 ? exp(B * lq) : 0;
		const double l1t = log1p(t);
		lR = (m > 0 ? log(m) : -INFINITY) + l1t / B - log(re);
		const double den = m > 0 ? 1 / (m * (1 + t)) : 0;
		const double wmin = qq > 0 ? t / qq : 0;
		dlx = copysign((xmax ? 1. : wmin) * den, x);
		dly = copysign((xmax ? wmin : 1.) * den, y);
		dlb = -l1t / (B * B) + (qq > 0 ? t * lq / (B * (1 + t)) : 0.);
	}
	const double rf = exp(lR / n);
	const double v = exp(-bn * (rf - 1));
	const double K = -bn * rf / n;
	const double cs = c.cos_ang, sn = c.sin_ang, iq = c.inv_axrat, q = c.axrat;
	const double Lx = K * dlx, Ly = K * dly;
	g[GT_XC] = -cs * Lx + sn * iq * Ly;
	g[GT_YC] = -sn * Lx - cs * iq * Ly;
	g[GT_RE] = -K / re;
	g[GT_N] = -c.dbn * (rf - 1) - K * lR / n;
	g[GT_PHI] = q * y * Lx - x * iq * Ly;
	g[GT_Q] = -y * iq * Ly;
	g[GT_B] = K * dlb;
	return v;
}

// a = weight of one evaluation point, g = dL/dθ: adds into the raw accumulators.
template <typename TG>
PK_HD void grad_accumulate(double acc[GA_SLOTS], double a, const TG g[GT_COUNT])
{
	acc[GA_W] += a;
	acc[GA_XC] += a * (double)g[GT_XC];
	acc[GA_YC] += a * (double)g[GT_YC];
	acc[GA_RE] += a * (double)g[GT_RE];
	acc[GA_N] += a * (double)g[GT_N];
	acc[GA_PHI] += a * (double)g[GT_PHI];
	acc[GA_Q] += a * (double)g[GT_Q];
	acc[GA_B] += a * (double)g[GT_B];
}

// ---------------------------------------------------------------------------
// Set-up: RadialProfile::initial_calculations + profile overrides, FP64.
// `sx`, `sy` are the (finesampled) pixel scales passed to Profile::evaluate,
// (pad_x, pad_y) the PSF padding offset (src/model.cpp:336-342).
// ---------------------------------------------------------------------------

struct SetupIn {
	pk_component c;
	double magzero, sx, sy, img_sx, img_sy, pad_x, pad_y;
	int32_t fp32_mode;
	uint32_t fs;               // finesampling factor (sky bg is divided by fs^2, src/sky.cpp:42-45)
	double fx_scale;           // fixed-point scale of the FP32-mode refinement of the model
	int32_t grad;              // also compute the gradient constants
	int32_t integrator;        // luminosity integrals: PK_INTEGRATOR_QUADPACK or PK_INTEGRATOR_EXACT
};

PK_HD double sersic_fluxfrac(double re, double nser, double bn, double fraction)
{
	// src/sersic.cpp:257-260
	double ratio = qgamma(fraction, 2 * nser) / bn;
	return re * pow(ratio, nser);
}

PK_HD double moffat_fluxfrac(double rscale, double con, double fraction)
{
	// src/moffat.cpp:72-75
	return rscale * sqrt(pow(1 - fraction, 1 / (1 - con)) - 1);
}

// qgamma(p[k], a) for k = 0, 1, 2: one per lane where lanes run side by side,
// else one after the other (the same values either way).
template <typename LN>
PK_HD void qgamma3(double a, const double p[3], double q[3], const LN &ln)
{
	if (ln.count() >= 3) {
		const int l = ln.lane();
		const double v = qgamma(p[l == 0 ? 0 : l == 1 ? 1 : 2], a);
		for (int k = 0; k < 3; k++) q[k] = ln.pick(v, k);
	} else {
		for (int k = 0; k < 3; k++) q[k] = qgamma(p[k], a);
	}
}

template <typename LN = SerialLanes>
PK_HD void setup_component_core(const SetupIn &in, PkComp &o, const LN &ln = LN())
{
	const pk_component &c = in.c;
	o = PkComp{};
	o.kind = c.kind;
	o.convolve = c.convolve;
	o.rough = c.rough;
	o.adjust = c.adjust;
	o.axrat = c.axrat;
	o.box = c.box;
	o.acc = c.acc;
	o.rscale_switch = c.rscale_switch;
	o.rscale_max = c.rscale_max;
	o.res = (int32_t)(c.resolution > 4096u ? 4096u : c.resolution);
	// deeper levels than 64 split cells below the resolution of a double
	o.max_rec = (int32_t)(c.max_recursions > 64u ? 64u : c.max_recursions);
	o.win_x = in.img_sx;
	o.win_y = in.img_sy;
	o.unfold = 1;
	if (c.kind == PK_PROFILE_SKY) {
		o.q[0] = c.p[PK_SKY_BG] / (double)(in.fs * in.fs);
		o.q[1] = 1 / (double)(in.fs * in.fs);  // d value / d bg
		return;
	}
	if (c.kind == PK_PROFILE_NULL || c.kind == PK_PROFILE_PSF) return;

	const double *p = c.p;
	double rscale = 0, lum = 0, bn = 0;
	double sq[3] = {0, 0, 0};  // Sersic quantiles: bn and those of rscale_switch, rscale_max
	o.boxy = c.box != 0;
	switch (c.kind) {
	case PK_PROFILE_SERSIC: {
		double re = p[PK_SERSIC_RE], n = p[PK_SERSIC_NSER];
		// bn (src/sersic.cpp:297) and the quantiles of the flux fractions of
		// the adjusted rscale_switch and rscale_max (src/sersic.cpp:257-260)
		const double fr[3] = {0.5, 1. - n * n / 2e3, 0.9999};
		qgamma3(2 * n, fr, sq, ln);
		bn = sq[0];
		rscale = re;
		double inv = n;
		if (o.boxy) inv *= (c.box + 2);
		o.sv = almost_equals(inv, 0.5) ? SV_HALF : almost_equals(inv, 1) ? SV_ONE : almost_equals(inv, 2) ? SV_TWO :
		       almost_equals(inv, 3) ? SV_THREE : almost_equals(inv, 4) ? SV_FOUR : almost_equals(inv, 8) ? SV_EIGHT :
		       almost_equals(inv, 16) ? SV_SIXTEEN : SV_GENERAL;
		o.q[0] = re;
		o.q[1] = n;
		o.q[2] = bn;
		o.q[3] = inv;
		lum = pow(re, 2) * 2 * PK_PI * n * gammafn(2 * n) * exp(bn) / pow(bn, 2 * n);  // src/sersic.cpp:283-289
		break;
	}
	case PK_PROFILE_MOFFAT: {
		double fwhm = p[PK_MOFFAT_FWHM], con = p[PK_MOFFAT_CON];
		rscale = fwhm / (2 * sqrt(pow(2, (1 / con)) - 1));  // src/moffat.cpp:81-83
		o.q[0] = fwhm;
		o.q[1] = con;
		lum = pow(rscale, 2) * PK_PI / (con - 1);
		break;
	}
	case PK_PROFILE_FERRER: {
		double rout = p[PK_FERRER_ROUT], a = p[PK_FERRER_A], b = p[PK_FERRER_B];
		rscale = rout;
		o.q[0] = rout;
		o.q[1] = a;
		o.q[2] = b;
		lum = pow(rout, 2) * PK_PI * (a * beta(a, 1 + 2 / (2 - b)));  // src/ferrer.cpp:85-101
		break;
	}
	case PK_PROFILE_KING: {
		o.q[0] = p[PK_KING_RC];
		o.q[1] = p[PK_KING_RT];
		o.q[2] = p[PK_KING_A];
		rscale = o.q[1];
		const PkComp &k = o;
		lum = 2 * PK_PI * integrate_adaptive([&](double r) { return r * king_r(k, r); }, 0.0, o.q[1], in.integrator, ln);
		break;
	}
	case PK_PROFILE_BROKENEXP: {
		o.q[0] = p[PK_BROKENEXP_H1];
		o.q[1] = p[PK_BROKENEXP_H2];
		o.q[2] = p[PK_BROKENEXP_RB];
		o.q[3] = p[PK_BROKENEXP_A];
		rscale = o.q[0] * 4;
		const double h1 = o.q[0], h2 = o.q[1], rb = o.q[2], a = o.q[3];
		lum = 2 * PK_PI * integrate_to_inf([&](double r) { return r * broken_exponential(r, h1, h2, rb, a); }, 0.0,
		                                   in.integrator, ln);
		break;
	}
	case PK_PROFILE_CORESERSIC: {
		o.q[0] = p[PK_CORESERSIC_RE];
		o.q[1] = p[PK_CORESERSIC_RB];
		o.q[2] = p[PK_CORESERSIC_NSER];
		o.q[3] = p[PK_CORESERSIC_A];
		o.q[4] = p[PK_CORESERSIC_B];
		bn = qgamma(0.5, 2 * o.q[2]);
		o.q[5] = bn;
		rscale = o.q[0];
		const PkComp &k = o;
		lum = 2 * PK_PI * integrate_to_inf([&](double r) { return r * coresersic_r(k, r); }, 0.0, in.integrator, ln);
		break;
	}
	default: break;
	}
	o.rscale = rscale;
	o.bn = bn;

	// src/radial.cpp:150-153
	double b2 = c.box + 2;
	double r_box = PK_PI * b2 / (2 * beta(1 / b2, 1 / b2));
	double lumtot = lum * c.axrat / r_box;
	double ie = pow(10, -0.4 * (c.mag - in.magzero)) / lumtot;
	o.lumtot = lumtot;
	o.ie = ie;

	if (c.adjust) {
		// src/radial.cpp:159-186 with the profile overrides
		double rsw = 0;
		switch (c.kind) {
		case PK_PROFILE_SERSIC:
			rsw = ceil(o.q[0] * pow(sq[1] / bn, o.q[1]));  // sersic_fluxfrac(re, n, bn, 1 - n^2 / 2000)
			rsw = fmax(fmin(rsw, 20.), 2.);
			rsw = rsw / o.q[0];
			break;
		case PK_PROFILE_MOFFAT:
			rsw = fmax(moffat_fluxfrac(rscale, o.q[1], 0.9999), 1.823 * o.q[0]);
			rsw = fmax(fmin(rsw, 20.), 2.) / rscale;
			break;
		case PK_PROFILE_FERRER:
		case PK_PROFILE_KING: rsw = 0.5; break;
		case PK_PROFILE_BROKENEXP:
		case PK_PROFILE_CORESERSIC: rsw = 1; break;
		}
		o.rscale_switch = rsw;
		unsigned res = (unsigned)ceil(160 / (rsw * rscale));
		res += res % 2;
		res = res > 16u ? 16u : res;
		res = res < 4u ? 4u : res;
		o.res = (int32_t)res;
		if (o.rscale_max == 0) {
			double rmax = 0;
			switch (c.kind) {
			case PK_PROFILE_SERSIC:
				rmax = ceil(fmax(o.q[0] * pow(sq[2] / bn, o.q[1]), 2.) / o.q[0]);  // sersic_fluxfrac(.., 0.9999)
				break;
			case PK_PROFILE_MOFFAT: rmax = ceil(fmax(moffat_fluxfrac(rscale, o.q[1], 0.9999), 2.) / rscale); break;
			case PK_PROFILE_FERRER: rmax = 1; break;
			case PK_PROFILE_KING: rmax = 1.5; break;
			case PK_PROFILE_BROKENEXP:
			case PK_PROFILE_CORESERSIC: rmax = 100; break;
			}
			o.rscale_max = rmax;
		}
		double acc = o.acc;
		if (c.kind == PK_PROFILE_SERSIC) {
			acc /= sqrt(o.q[1]);  // src/sersic.cpp:278-281
			acc = fmax(0.1, acc) / c.axrat;
		} else if (c.kind == PK_PROFILE_MOFFAT) {
			acc = 0.1 / c.axrat;  // src/moffat.cpp:97-99
		}
		o.acc = acc;
	}

	o.r2_max = (o.rscale_max * rscale) * (o.rscale_max * rscale);
	o.r2_switch = (o.rscale_switch * rscale) * (o.rscale_switch * rscale);

	double angrad = fmod(c.ang + 90, 360.) * PK_PI / 180.;
	o.inv_axrat = 1 / c.axrat;
	o.total_flux = pow(10, -0.4 * (c.mag - in.magzero));
	o.cos_ang = cos(angrad);
	o.sin_ang = sin(angrad);
	o.xcen = c.xcen + in.pad_x * in.sx;
	o.ycen = c.ycen + in.pad_y * in.sy;

	double flux_scale = in.sx * in.sy * ie;
	if (c.kind == PK_PROFILE_SERSIC && c.rescale_flux) {
		// src/sersic.cpp:303-325; libprofit leaves the factor unset without
		// adjust, here it is 1 then.
		double rescale = 1;
		if (c.adjust) {
			double flux_r = bn * pow(o.rscale_max / o.q[0], 1 / o.q[1]);
			rescale = 1 / pgamma(flux_r, 2 * o.q[1]);
		}
		flux_scale *= rescale;
	}
	o.flux_scale = flux_scale;

	if (c.kind == PK_PROFILE_SERSIC && c.adjust && o.q[1] > 1) {
		o.center_res = 8;
		o.center_max_rec = 10;
	}

	if (in.grad && c.kind == PK_PROFILE_SERSIC) {
		// flux_scale = sx sy 10^(-0.4 (mag - magzero)) r_box(B) / (lum axrat), B = box + 2,
		// lum = re^2 2 pi n Gamma(2n) e^bn / bn^(2n)  (src/sersic.cpp:283-289, src/radial.cpp:150-153)
		const double n = o.q[1], re = o.q[0], B = c.box + 2;
		o.dbn = sersic_dbn_dn(n);
		o.glnf_re = -2 / re;
		o.glnf_n = -(1 / n + 2 * digamma(2 * n) + o.dbn - 2 * log(bn) - 2 * n * o.dbn / bn);
		o.glnf_q = -1 / c.axrat;
		// d log r_box / dB, r_box = pi B / (2 Beta(1/B, 1/B))
		o.glnf_b = 1 / B + 2 * (digamma(1 / B) - digamma(2 / B)) / (B * B);
	}

	// FP32 constants
	o.fp32 = 0;
	if (in.fp32_mode) {
		switch (c.kind) {
		case PK_PROFILE_SERSIC:
			o.fp32 = 1;
			o.f[0] = (float)(-log2(o.q[0]));
			o.f[1] = (float)(1 / o.q[1]);
			o.f[2] = (float)bn;
			o.f[3] = (float)o.dbn;
			o.f[4] = (float)(1 / o.q[0]);
			o.f[7] = (float)(c.box + 2);
			o.unfold = exp(bn);
			break;
		case PK_PROFILE_MOFFAT:
			o.fp32 = 1;
			o.f[0] = (float)(1 / (rscale * rscale));
			o.f[1] = (float)o.q[1];
			o.f[7] = (float)(c.box + 2);
			break;
		case PK_PROFILE_BROKENEXP:
			o.fp32 = 1;
			o.f[0] = (float)(1 / o.q[0]);
			o.f[1] = (float)(1 / o.q[0] - 1 / o.q[1]);
			o.f[2] = (float)o.q[2];
			o.f[3] = (float)o.q[3];
			o.f[4] = (float)(1 / o.q[3]);
			o.f[7] = (float)(c.box + 2);
			break;
		case PK_PROFILE_CORESERSIC:
			o.fp32 = 1;
			o.f[0] = (float)(-log2(o.q[1]));
			o.f[1] = (float)o.q[3];
			o.f[2] = (float)(o.q[4] / o.q[3]);
			o.f[3] = (float)bn;
			o.f[4] = (float)pow(o.q[1], o.q[3]);
			o.f[5] = (float)(1 / pow(o.q[0], o.q[3]));
			o.f[6] = (float)(1 / (o.q[2] * o.q[3]));
			o.f[7] = (float)(c.box + 2);
			break;
		case PK_PROFILE_KING:
			o.fp32 = 1;
			o.f[0] = (float)(1 / (o.q[0] * o.q[0]));
			o.f[1] = (float)(1 / o.q[2]);
			o.f[2] = (float)o.q[2];
			o.q[4] = o.q[1] * o.q[1];
			o.q[5] = 1 / (o.q[0] * o.q[0] + o.q[1] * o.q[1]);
			break;
		case PK_PROFILE_FERRER:
			o.fp32 = 1;
			o.f[0] = (float)o.q[1];
			o.f[1] = (float)((2 - o.q[2]) / 2);
			o.q[4] = o.q[0] * o.q[0];
			o.q[5] = 1 / o.q[4];
			break;
		default: break;
		}
	}
}

PK_HD void fill_pix_hot(PkComp &o)
{
	PixHot &h = o.hot;
	h = PixHot{};
	h.kind = o.kind;
	h.boxy = o.boxy;
	h.xcen = o.xcen;
	h.ycen = o.ycen;
	h.vscale = o.kind == PK_PROFILE_SKY ? o.q[0] : o.unfold * o.flux_scale;
	for (int k = 0; k < 8; k++) h.f[k] = o.f[k];
	h.flags = (o.convolve ? PIX_CONVOLVE : 0) | (o.fp32 ? PIX_FP32 : 0);
	h.cos32 = (float)o.cos_ang;
	h.sin32 = (float)o.sin_ang;
	h.iq32 = (float)o.inv_axrat;
	// Guard bands. With x', y' formed in FP32 from the FP32-rounded offsets
	// to the centre (u = 2^-24 per operand and operation), the FP32 r^2
	// differs from the exact one by at most 31 u M r^2, M = max(q^-2, q^2)
	// (q = axrat).  A band of relative half-width c = 3e-5 M (16 times that,
	// and far wider than the FP64 rounding of libprofit's own test) leaves
	// every pixel outside it on the side of the limit its exact radius is on.
	const double inf = (double)INFINITY;
	const double iq2 = o.inv_axrat * o.inv_axrat;
	const double cband = 3e-5 * fmax(iq2, 1 / iq2);
	auto fp32_range = [](double l) { return l == 0 || (l >= 1e-20 && l <= 1e20); };
	const bool has_max = o.rscale_max > 0;
	const bool cls32 = o.fp32 && fp32_range(o.r2_switch) && (!has_max || fp32_range(o.r2_max)) &&
	                   o.inv_axrat >= 1e-4 && o.inv_axrat <= 1e4;
	if (!cls32) {
		// every pixel to the FP64 test
		h.lo_sw = -INFINITY;
		h.hi_sw = INFINITY;
		h.lo_mx = h.hi_mx = INFINITY;
		return;
	}
	if (o.rough) {
		h.lo_sw = h.hi_sw = -INFINITY;  // never refined
	} else {
		h.lo_sw = (float)(o.r2_switch * (1 - cband) - 1e-30);
		h.hi_sw = (float)(o.r2_switch * (1 + cband) + 1e-30);
	}
	if (has_max) {
		h.lo_mx = (float)(o.r2_max * (1 - cband) - 1e-30);
		h.hi_mx = (float)(o.r2_max * (1 + cband) + 1e-30);
	} else {
		h.lo_mx = h.hi_mx = (float)inf;
	}
}

template <typename LN = SerialLanes>
PK_HD void setup_component(const SetupIn &in, PkComp &o, const LN &ln = LN())
{
	setup_component_core(in, o, ln);
	fill_pix_hot(o);
}

PK_HD CompLite make_lite(const PkComp &c, const SetupIn &in)
{
	CompLite l{};
	l.cos_ang = c.cos_ang;
	l.sin_ang = c.sin_ang;
	l.inv_axrat = c.inv_axrat;
	l.wflux = c.flux_scale * in.fx_scale;
	l.wq = l.wflux * c.unfold;
	l.acc = c.acc;
	l.hsx = 0.5 * in.sx;
	l.hsy = 0.5 * in.sy;
	l.q4 = c.q[4];
	l.q5 = c.q[5];
	l.b2 = c.box + 2;
	for (int k = 0; k < 8; k++) l.f[k] = c.f[k];
	l.kind = (int8_t)c.kind;
	l.boxy = (int8_t)c.boxy;
	l.fp32 = (int8_t)c.fp32;
	l.convolve = (int8_t)(c.convolve ? 1 : 0);
	return l;
}

}  // namespace pk
