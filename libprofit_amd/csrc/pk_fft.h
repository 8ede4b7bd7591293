// FFT convolution (the FFT convolver, src/convolve.cpp:234-312), FP64.
//
// libprofit zero-pads source and kernel to 2 max(src, krn) and multiplies
// their 1-D r2c FFTs over the flattened buffer, which equals their 2-D linear
// convolution L.  Here L is computed directly by 2-D FFTs of size P x Q
// (powers of two with P >= sw + kw - 1, Q >= sh + kh - 1, so the circular
// convolution does not wrap) and then sampled in the convolver's geometry:
//   fft_pack_kernel    Z = src + i c krn on the P x Q grid (both real, at the origin;
//                      c a power of two that gives both the same peak)
//   fft_pass_kernel    in-place radix-4 (+ one radix-2) FFT along one axis of
//                      several rows or columns per workgroup, staged in LDS.
//                      Forward passes are decimation in frequency, whose
//                      output sits in digit-reversed order; inverse passes
//                      undo the same stages in reverse (conjugate twiddles), so
//                      no permutation pass is needed anywhere.
//   fft_mul_kernel     S K from Z: S(k) = (Z(k) + Z*(-k)) / 2, K(k) = (Z(k) - Z*(-k)) / 2i,
//                      so S K = (Z(k)^2 - Z*(-k)^2) / 4i, with the 1/(PQ) of the
//                      inverse; k and -k are visited by one thread
//   fft_sample_kernel  conv(X, Y) = Re L(X + csx + kw-1-kw/2, Y + csy + kh-1-kh/2),
//                      the same-size convolution of conv_kernel at the shifted
//                      position (0 outside L), plus the rows libprofit's
//                      circular buffer of 2 max(src, krn) rows wraps onto it
//                      (images much taller than wide), then the mask
// FFT rounding error is relative to the peak (~log2(PQ) eps max|L|): ~1e-15 of
// the peak in FP64.
#pragma once

#include <hip/hip_runtime.h>

#include "pk_kernels.h"

namespace pk {

constexpr int FFT_MAX_LOG2 = 12;           // longest axis: 4096
constexpr int FFT_TWIDDLES = 1 << FFT_MAX_LOG2;
constexpr int FFT_LDS_COMPLEX = 4096;      // complex doubles staged per workgroup (64 KB)

struct FftPlan {
	// stage radices of one axis, first (largest blocks) first
	int nstages;
	int radix[FFT_MAX_LOG2];
	int n, log2n;
};

PK_HD FftPlan fft_plan(int log2n)
{
	FftPlan p{};
	p.log2n = log2n;
	p.n = 1 << log2n;
	int left = log2n;
	if (left & 1) {
		p.radix[p.nstages++] = 2;
		left -= 1;
	}
	while (left > 0) {
		p.radix[p.nstages++] = 4;
		left -= 2;
	}
	return p;
}

// frequency index held at position `pos` after the forward passes: the digit
// of the first stage is the most significant of the position and the least
// significant of the frequency
__device__ inline uint32_t fft_freq_of_pos(const FftPlan &p, uint32_t pos)
{
	uint32_t k = 0, mult = 1, weight = (uint32_t)p.n;
	for (int s = 0; s < p.nstages; s++) {
		const uint32_t r = (uint32_t)p.radix[s];
		weight /= r;
		k += ((pos / weight) % r) * mult;
		mult *= r;
	}
	return k;
}

__device__ inline uint32_t fft_pos_of_freq(const FftPlan &p, uint32_t k)
{
	uint32_t pos = 0, weight = (uint32_t)p.n;
	for (int s = 0; s < p.nstages; s++) {
		const uint32_t r = (uint32_t)p.radix[s];
		weight /= r;
		pos += (k % r) * weight;
		k /= r;
	}
	return pos;
}

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiplication by -i (forward) or +i (inverse)
__device__ inline double2 cmul_mi(double2 a) { return make_double2(a.y, -a.x); }
__device__ inline double2 cmul_pi(double2 a) { return make_double2(-a.y, a.x); }

// max |src| as the bits of a non-negative double (their integer order is the
// order of the values); `mx` must be 0 before the launch
template <typename TI>
__global__ __launch_bounds__(256) void fft_maxabs_kernel(const TI *__restrict__ src, size_t n,
                                                         unsigned long long *__restrict__ mx)
{
	double m = 0;
	for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
		const double v = fabs((double)src[t]);
		if (v > m) m = v;  // NaN never wins
	}
	for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
	if (__lane_id() == 0 && m > 0) atomicMax(mx, (unsigned long long)__double_as_longlong(m));
}

__global__ void fft_zero_kernel(unsigned long long *mx) { *mx = 0; }

// Power of two that brings the kernel's peak to the source's: src and krn are
// transformed together as src + i c krn, and each of the two recovered
// transforms carries rounding error relative to the larger of the pair, so
// both are kept at the same magnitude (c is exact to apply and to undo).
__device__ inline double fft_krn_scale(const unsigned long long *mx, double krn_max)
{
	const double sm = __longlong_as_double((long long)*mx);
	if (!(sm > 0) || !(krn_max > 0) || !isfinite(sm)) return 1.0;
	return ldexp(1.0, ilogb(sm) - ilogb(krn_max));
}

template <typename TI>
__global__ __launch_bounds__(256) void fft_pack_kernel(const TI *__restrict__ src, uint32_t sw, uint32_t sh,
                                                       const double *__restrict__ krn, uint32_t kw, uint32_t kh,
                                                       double krn_max, const unsigned long long *__restrict__ mx,
                                                       double2 *__restrict__ z, uint32_t P, uint32_t Q)
{
	const size_t n = (size_t)P * Q;
	const double c = fft_krn_scale(mx, krn_max);
	for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
		const uint32_t x = (uint32_t)(t % P), y = (uint32_t)(t / P);
		double re = 0, im = 0;
		if (x < sw && y < sh) re = (double)src[x + (size_t)y * sw];
		if (x < kw && y < kh) im = krn[x + (size_t)y * kw] * c;
		z[t] = make_double2(re, im);
	}
}

// One axis of a 2-D FFT: workgroup g transforms `per_group` lines of length
// plan.n, line t at z + t * line_stride with elements elem_stride apart.
// `tw[t] = exp(-2 pi i t / FFT_TWIDDLES)`.  INVERSE undoes the forward stages.
template <bool INVERSE>
__global__ __launch_bounds__(256) void fft_pass_kernel(double2 *__restrict__ z, FftPlan plan, uint32_t n_lines,
                                                       uint32_t per_group, size_t line_stride, size_t elem_stride,
                                                       const double2 *__restrict__ tw)
{
	__shared__ double2 buf[FFT_LDS_COMPLEX];
	const uint32_t n = (uint32_t)plan.n;
	const uint32_t l0 = blockIdx.x * per_group;
	const uint32_t nl = min(per_group, n_lines - l0);
	const uint32_t total = nl * n;
	// lines of contiguous elements are read element-fastest, strided lines
	// (columns) line-fastest, so that both read whole cache lines
	const bool rows = elem_stride == 1;
	for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
		const uint32_t t = rows ? i / n : i % nl, e = rows ? i % n : i / nl;
		buf[t * n + e] = z[(size_t)(l0 + t) * line_stride + (size_t)e * elem_stride];
	}
	__syncthreads();
	for (int si = 0; si < plan.nstages; si++) {
		const int s = INVERSE ? plan.nstages - 1 - si : si;
		// block length of stage s: n / (product of the earlier radices)
		uint32_t L = n;
		for (int k = 0; k < s; k++) L /= (uint32_t)plan.radix[k];
		const uint32_t r = (uint32_t)plan.radix[s];
		const uint32_t q = L / r;                 // butterfly span
		const uint32_t tw_step = FFT_TWIDDLES / L;  // W_L^m = tw[m * tw_step]
		const uint32_t nb = total / r;
		for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
			const uint32_t t = b / (n / r), u = b % (n / r);
			const uint32_t blk = u / q, j = u % q;
			double2 *x = buf + t * n + blk * L + j;
			if (r == 2) {
				const double2 w = tw[j * tw_step];
				if (!INVERSE) {
					const double2 a0 = x[0], a1 = x[q];
					x[0] = cadd(a0, a1);
					x[q] = cmul(csub(a0, a1), w);
				} else {
					const double2 a0 = x[0], a1 = cmul(x[q], cconj(w));
					x[0] = cadd(a0, a1);
					x[q] = csub(a0, a1);
				}
			} else {
				const double2 w1 = tw[j * tw_step], w2 = tw[2 * j * tw_step], w3 = tw[3 * j * tw_step];
				if (!INVERSE) {
					const double2 a0 = x[0], a1 = x[q], a2 = x[2 * q], a3 = x[3 * q];
					const double2 s02 = cadd(a0, a2), d02 = csub(a0, a2), s13 = cadd(a1, a3),
					              d13 = cmul_mi(csub(a1, a3));
					x[0] = cadd(s02, s13);                  // a0 + a1 + a2 + a3
					x[q] = cmul(cadd(d02, d13), w1);        // a0 - i a1 - a2 + i a3
					x[2 * q] = cmul(csub(s02, s13), w2);    // a0 - a1 + a2 - a3
					x[3 * q] = cmul(csub(d02, d13), w3);    // a0 + i a1 - a2 - i a3
				} else {
					const double2 b0 = x[0], b1 = cmul(x[q], cconj(w1)), b2 = cmul(x[2 * q], cconj(w2)),
					              b3 = cmul(x[3 * q], cconj(w3));
					const double2 s02 = cadd(b0, b2), d02 = csub(b0, b2), s13 = cadd(b1, b3),
					              d13 = cmul_pi(csub(b1, b3));
					x[0] = cadd(s02, s13);
					x[q] = cadd(d02, d13);
					x[2 * q] = csub(s02, s13);
					x[3 * q] = csub(d02, d13);
				}
			}
		}
		__syncthreads();
	}
	for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
		const uint32_t t = rows ? i / n : i % nl, e = rows ? i % n : i / nl;
		z[(size_t)(l0 + t) * line_stride + (size_t)e * elem_stride] = buf[t * n + e];
	}
}

// Z -> S K / (P Q c) in place, in the digit-reversed order of both axes.
__global__ __launch_bounds__(256) void fft_mul_kernel(double2 *__restrict__ z, FftPlan px, FftPlan py, double krn_max,
                                                      const unsigned long long *__restrict__ mx)
{
	const uint32_t P = (uint32_t)px.n, Q = (uint32_t)py.n;
	const size_t n = (size_t)P * Q;
	const double scale = 1.0 / ((double)P * (double)Q) / fft_krn_scale(mx, krn_max);
	for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
		const uint32_t p1 = (uint32_t)(t % P), p2 = (uint32_t)(t / P);
		const uint32_t k1 = fft_freq_of_pos(px, p1), k2 = fft_freq_of_pos(py, p2);
		const uint32_t m1 = (P - k1) & (P - 1), m2 = (Q - k2) & (Q - 1);
		const size_t u = fft_pos_of_freq(px, m1) + (size_t)fft_pos_of_freq(py, m2) * P;
		if (u < t) continue;  // the pair is done by the thread of its first position
		const double2 a = z[t], b = z[u];
		// S K at k: (a^2 - conj(b)^2) / 4i;  at -k: (b^2 - conj(a)^2) / 4i;  x / 4i = -i x / 4
		const double2 ca = cconj(a), cb = cconj(b);
		const double2 pk = cmul_mi(csub(cmul(a, a), cmul(cb, cb)));
		const double2 pm = cmul_mi(csub(cmul(b, b), cmul(ca, ca)));
		const double f = 0.25 * scale;
		z[t] = make_double2(pk.x * f, pk.y * f);
		if (u != t) z[u] = make_double2(pm.x * f, pm.y * f);
	}
}

// conv(X, Y) for the job's conv geometry from the linear convolution in z.
template <typename TI>
__global__ __launch_bounds__(256) void fft_sample_kernel(const PkJob *__restrict__ job, const double2 *__restrict__ z,
                                                         uint32_t P, const uint8_t *__restrict__ masks,
                                                         TI *__restrict__ dst)
{
	const PkJob &jb = *job;
	const size_t n = (size_t)jb.cw * jb.ch;
	const int W = (int)jb.dw, H = (int)jb.dh, kw = (int)jb.kw, kh = (int)jb.kh;
	const int lw = W + kw - 1, lh = H + kh - 1;
	for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
		const int X = (int)(t % jb.cw), Y = (int)(t / jb.cw);
		const int u = X + jb.csx + (kw - 1 - kw / 2), v = Y + jb.csy + (kh - 1 - kh / 2);
		// the FFT convolver's buffer is circular in y with wrap_y rows: rows
		// v + m wrap_y of L land on row v as well
		double val = 0;
		if (u >= 0 && u < lw)
			for (int vv = v; vv < lh; vv += jb.wrap_y ? (int)jb.wrap_y : lh) {
				if (vv >= 0) val += z[(size_t)u + (size_t)vv * P].x;
			}
		if (jb.mask_off != NONE) {
			const int mi = X - jb.cmx, mj = Y - jb.cmy;
			if (mi < 0 || mj < 0 || mi >= W || mj >= H || !masks[jb.mask_off + mi + (size_t)mj * W]) val = 0;
		}
		dst[jb.conv_off + t] = (TI)val;
	}
}

}  // namespace pk
