// gfx950 kernels of the image-generation path (included once, by pk_api.hip).
//
//  coarse_kernel   one lane per drawing pixel: classifies every component as
//                  zero / direct / needs-refinement (src/radial.cpp:291-344),
//                  evaluates the direct ones, appends refinement work with one
//                  atomic per wavefront (ballot + mbcnt compaction).
//  refine_kernel   one wavefront per refinement entry: the recursive
//                  subsample_pixel (src/radial.cpp:55-135) as an iterative LIFO
//                  of cells kept in LDS; sub-cells of one or more cells are
//                  spread over the 64 lanes, cells that fail the accuracy test
//                  are compacted back onto the stack with a ballot.
//  psf_kernel      PsfProfile rebinning (src/psf.cpp:58-145).
//  conv_kernel     direct "same" convolution (src/convolve.cpp:101-231), one
//                  64x16 output tile per 256-lane workgroup, source tile in LDS,
//                  reversed PSF read through scalar loads.
//  compose_kernel  crop of the PSF padding, SUM downsampling and the final mask
//                  (src/model.cpp:289-316).
//
// All FP64 code is compiled with -ffp-contract=off so that products and sums
// round as in libprofit's CPU build; FP32 hot loops use explicit fmaf.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

#include "pk_kernels.h"

namespace pk {

// ---------------------------------------------------------------------------
// image -> profile coordinates (src/radial.cpp:47-53)
// ---------------------------------------------------------------------------
__device__ inline void to_prof(const PkComp &c, double x, double y, double &xp, double &yp)
{
	x -= c.xcen;
	y -= c.ycen;
	xp = x * c.cos_ang + y * c.sin_ang;
	yp = -x * c.sin_ang + y * c.cos_ang;
	yp /= c.axrat;
}

// Job of workgroup `block` in the tiling `which` (0 coarse, 1 convolution,
// 2 composition, 3 convolution strips): from the batch's per-tile table when
// given (one scalar load), else by binary search over the jobs' tile offsets.
__device__ inline uint32_t find_job(const PkJob *jobs, uint32_t njobs, uint32_t block, int which,
                                    const uint32_t *__restrict__ tile_job = nullptr)
{
	if (tile_job) return tile_job[block];
	int lo = 0, hi = (int)njobs - 1;
	while (lo < hi) {
		int mid = (lo + hi + 1) >> 1;
		uint32_t start = which == 0   ? jobs[mid].tile_off
		                 : which == 1 ? jobs[mid].ctile_off
		                 : which == 2 ? jobs[mid].otile_off
		                              : jobs[mid].stile_off;
		if (start <= block) lo = mid;
		else hi = mid - 1;
	}
	return (uint32_t)lo;
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

template <int N>
struct IntTag {
	static constexpr int value = N;
};

// f(IntTag<0>{}), ..., f(IntTag<N - 1>{}) in turn, each index a compile-time
// constant (register arrays indexed by it stay in registers whatever the
// loop unroller decides)
template <class F, int... S>
__device__ __forceinline__ void static_for_seq(F &&f, std::integer_sequence<int, S...>)
{
	(f(IntTag<S>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f)
{
	static_for_seq(f, std::make_integer_sequence<int, N>{});
}

// ---------------------------------------------------------------------------
// pixel passes (coarse, residual, coarse adjoint, composition): a 256-lane
// workgroup covers a 64 x PX_ROWS tile, each lane one column of PX_RPT rows,
// so that the per-component constants are loaded once for PX_RPT pixels
// ---------------------------------------------------------------------------
constexpr int PX_ROWS = 16;  // tile rows
constexpr int PX_RPT = 4;    // rows per lane

// ---------------------------------------------------------------------------
// coarse pass
// ---------------------------------------------------------------------------
enum { CLS_ZERO = 0, CLS_DIRECT = 1, CLS_REFINE = 2, CLS_FP64 = -1 };

// FP32 classification of one pixel against one component (FP32 mode), from
// its FP32 profile coordinates and the guard bands of the PixHot block:
// outside the bands the decision is libprofit's (src/radial.cpp:316-321),
// within one (or for NaN) CLS_FP64 asks for the FP64 test.
__device__ inline int classify32(const PixHot &h, float xf, float yf)
{
	const float r2 = fmaf(xf, xf, yf * yf);
	int cls = r2 > h.hi_sw ? CLS_DIRECT : r2 < h.lo_sw ? CLS_REFINE : CLS_FP64;
	if (!(r2 < h.lo_mx)) cls = CLS_FP64;
	if (r2 > h.hi_mx) cls = CLS_ZERO;
	return cls;
}

// libprofit's classification in FP64 (src/radial.cpp:316-321)
__device__ inline int classify64(const PkComp &c, double xp, double yp)
{
	const double r2 = xp * xp + yp * yp;
	if (c.rscale_max > 0 && radius_exceeds(r2, c.r2_max, c.rscale, c.rscale_max)) return CLS_ZERO;
	if (c.rough || radius_exceeds(r2, c.r2_switch, c.rscale, c.rscale_switch)) return CLS_DIRECT;
	return CLS_REFINE;
}

// Stores the pixels of one lane (PX_RPT rows) and appends those that need
// refinement: one atomic for the wavefront's rows, row r's pixels after those
// of rows < r.
template <typename TI>
__device__ inline void coarse_store(const PkJob &jb, uint32_t jn, uint32_t i, uint32_t j0, const bool inside[PX_RPT],
                                    const double acc_nc[PX_RPT], const double acc_c[PX_RPT],
                                    const uint32_t cmask[PX_RPT], TI *__restrict__ img_nc, TI *__restrict__ img_c,
                                    RefineEntry *__restrict__ entries, PkCounters *__restrict__ counters,
                                    unsigned long long *__restrict__ acc)
{
	uint64_t ballot[PX_RPT];
	uint32_t n_need = 0;
#pragma unroll
	for (int r = 0; r < PX_RPT; r++) {
		const uint32_t pix = (j0 + r) * jb.dw + i;
		if (inside[r]) {
			if (jb.has_nc && !jb.nc_sky) img_nc[jb.img_off + pix] = (TI)acc_nc[r];
			if (jb.has_conv) img_c[jb.img_off + pix] = (TI)acc_c[r];
		}
		ballot[r] = __ballot(cmask[r] != 0);
		n_need += __popcll(ballot[r]);
	}
	if (n_need) {
		uint32_t base = 0;
		if (lane_id() == 0) base = atomicAdd(&counters->n_entries, n_need);
		base = __shfl(base, 0);
#pragma unroll
		for (int r = 0; r < PX_RPT; r++) {
			if (cmask[r]) {
				const uint32_t e = base + __popcll(ballot[r] & ((1ull << lane_id()) - 1ull));
				entries[e] = RefineEntry{jn, (j0 + r) * jb.dw + i, cmask[r]};
				if (acc) {
					acc[2 * (size_t)e] = 0;
					acc[2 * (size_t)e + 1] = 0;
				}
			}
			base += __popcll(ballot[r]);
		}
	}
}

// Evaluation count of the wavefront (COUNT instantiations), stored in its own slot.
__device__ inline void coarse_store_count(uint32_t nev, uint32_t *__restrict__ wave_evals)
{
	for (int off = 32; off > 0; off >>= 1) nev += __shfl_xor(nev, off);
	if (lane_id() == 0) wave_evals[blockIdx.x * 4 + (threadIdx.x >> 6)] = nev;
}

// FP64 mode: libprofit's pixel loop (src/radial.cpp:291-344) in FP64.
// COUNT: classification only -- no values, writes or appends; each
// wavefront stores how many evaluate_at calls its pixels need (the work the
// product instantiation, COUNT = false, does).
template <typename TI, bool COUNT = false>
__global__ __launch_bounds__(256) void coarse_kernel(const PkJob *__restrict__ jobs, uint32_t njobs,
                                                     const PkComp *__restrict__ comps,
                                                     const uint8_t *__restrict__ masks, TI *__restrict__ img_nc,
                                                     TI *__restrict__ img_c, RefineEntry *__restrict__ entries,
                                                     PkCounters *__restrict__ counters,
                                                     unsigned long long *__restrict__ acc,
                                                     const uint32_t *__restrict__ tile_job,
                                                     uint32_t *__restrict__ wave_evals)
{
	const uint32_t jn = find_job(jobs, njobs, blockIdx.x, 0, tile_job);
	const PkJob &jb = jobs[jn];
	const uint32_t t = blockIdx.x - jb.tile_off;
	const uint32_t i = (t % jb.tiles_x) * 64 + (threadIdx.x & 63);
	const uint32_t j0 = (t / jb.tiles_x) * PX_ROWS + (threadIdx.x >> 6) * PX_RPT;
	bool inside[PX_RPT], live[PX_RPT];
	double acc_nc[PX_RPT], acc_c[PX_RPT];
	uint32_t cmask[PX_RPT];
#pragma unroll
	for (int r = 0; r < PX_RPT; r++) {
		inside[r] = i < jb.dw && j0 + r < jb.dh;
		live[r] = inside[r] && !(jb.mask_off != NONE && !masks[jb.mask_off + (j0 + r) * jb.dw + i]);
		acc_nc[r] = acc_c[r] = 0;
		cmask[r] = 0;
	}
	const double hx = jb.sx / 2., hy = jb.sy / 2.;
	const double x = hx + i * jb.sx;
	uint32_t nev = 0;  // COUNT: evaluate_at calls of this lane
	for (uint32_t c = 0; c < jb.ncomp; c++) {
		const PkComp &cp = comps[jb.comp_off + c];
#pragma unroll
		for (int r = 0; r < PX_RPT; r++) {
			if (!live[r]) continue;
			const double y = hy + (j0 + r) * jb.sy;
			double v;
			if (cp.kind == PK_PROFILE_SKY) {
				v = cp.q[0];
			} else {
				double xp, yp;
				to_prof(cp, x, y, xp, yp);
				const int cls = classify64(cp, xp, yp);
				if (cls == CLS_ZERO) continue;
				if (cls == CLS_REFINE) {
					cmask[r] |= 1u << c;
					continue;
				}
				if (COUNT) {
					nev++;
					continue;
				}
				v = evaluate_at(cp, xp, yp) * cp.flux_scale;
			}
			if (cp.convolve) acc_c[r] += v;
			else acc_nc[r] += v;
		}
	}
	if (COUNT) {
		coarse_store_count(nev, wave_evals);
		return;
	}
	coarse_store<TI>(jb, jn, i, j0, inside, acc_nc, acc_c, cmask, img_nc, img_c, entries, counters, acc);
}

// FP32 profile value of kind K at (x, y) (the FP32 forms of the radial profiles
// that take FP32 coordinates; king and ferrer take an FP64 r^2).
template <int K, typename CR>
__device__ inline float value32(const CR &c, float x, float y)
{
	switch (K) {
	case PK_PROFILE_SERSIC: return sersic32(c, x, y);
	case PK_PROFILE_MOFFAT: return moffat32(c, x, y);
	case PK_PROFILE_BROKENEXP: return brokenexp32(c, x, y);
	case PK_PROFILE_CORESERSIC: return coresersic32(c, x, y);
	default: return -1.f;
	}
}

// FP32 mode. Per component: the PixHot block in scalar registers, the FP32
// classification of the lane's PX_RPT pixels (FP64 test only inside the
// guard bands), then the direct values.  Branches are wavefront-uniform
// wherever possible (a row is evaluated in all lanes when any of its lanes
// needs the value, the others discard theirs): divergent branches would hold
// one execution mask per row and condition in scalar registers.
template <typename TI, bool COUNT = false>
__global__ __launch_bounds__(256) void coarse32_kernel(const PkJob *__restrict__ jobs, uint32_t njobs,
                                                       const PkComp *__restrict__ comps,
                                                       const uint8_t *__restrict__ masks, TI *__restrict__ img_nc,
                                                       TI *__restrict__ img_c, RefineEntry *__restrict__ entries,
                                                       PkCounters *__restrict__ counters,
                                                       unsigned long long *__restrict__ acc,
                                                       const uint32_t *__restrict__ tile_job,
                                                       uint32_t *__restrict__ wave_evals)
{
	const uint32_t jn = find_job(jobs, njobs, blockIdx.x, 0, tile_job);
	const PkJob &jb = jobs[jn];
	const uint32_t t = blockIdx.x - jb.tile_off;
	const uint32_t i = (t % jb.tiles_x) * 64 + (threadIdx.x & 63);
	const uint32_t j0 = (t / jb.tiles_x) * PX_ROWS + (threadIdx.x >> 6) * PX_RPT;
	bool inside[PX_RPT];
	uint32_t live = 0;  // bit r: row r is drawn and not masked
	double acc_nc[PX_RPT], acc_c[PX_RPT];
	uint32_t cmask[PX_RPT];
#pragma unroll
	for (int r = 0; r < PX_RPT; r++) {
		inside[r] = i < jb.dw && j0 + r < jb.dh;
		if (inside[r] && !(jb.mask_off != NONE && !masks[jb.mask_off + (j0 + r) * jb.dw + i])) live |= 1u << r;
		acc_nc[r] = acc_c[r] = 0;
		cmask[r] = 0;
	}
	const double hx = jb.sx / 2., hy = jb.sy / 2.;
	const double x = hx + i * jb.sx;
	uint32_t nev = 0;  // COUNT: evaluate_at calls of this lane
	for (uint32_t c = 0; c < jb.ncomp; c++) {
		const PkComp &cp = comps[jb.comp_off + c];
		const PixHot h = cp.hot;
		const bool conv = (h.flags & PIX_CONVOLVE) != 0;
		if (h.kind == PK_PROFILE_SKY) {
#pragma unroll
			for (int r = 0; r < PX_RPT; r++) {
				const double v = (live >> r) & 1u ? h.vscale : 0.0;
				if (conv) acc_c[r] += v;
				else acc_nc[r] += v;
			}
			continue;
		}
		// classification: profile coordinates in FP32 from the FP64 offsets
		// to the centre (src/radial.cpp:47-53)
		const float dxf = (float)(x - h.xcen);
		const float dxc = dxf * h.cos32, ndxs = -dxf * h.sin32;
		int cls[PX_RPT];
		float xf[PX_RPT], yf[PX_RPT];
		bool any64 = false;
#pragma unroll
		for (int r = 0; r < PX_RPT; r++) {
			const float dyf = (float)((hy + (j0 + r) * jb.sy) - h.ycen);
			xf[r] = fmaf(dyf, h.sin32, dxc);
			yf[r] = fmaf(dyf, h.cos32, ndxs) * h.iq32;
			const int k = classify32(h, xf[r], yf[r]);
			cls[r] = (live >> r) & 1u ? k : (int)CLS_ZERO;
			any64 = any64 || cls[r] == CLS_FP64;
		}
		if (__ballot(any64)) {
			// within a guard band, or no FP32 classification: libprofit's test
#pragma unroll
			for (int r = 0; r < PX_RPT; r++) {
				if (cls[r] != CLS_FP64) continue;
				double xp, yp;
				to_prof(cp, x, hy + (j0 + r) * jb.sy, xp, yp);
				cls[r] = classify64(cp, xp, yp);
				xf[r] = (float)xp;
				yf[r] = (float)yp;
			}
		}
		uint32_t drow = 0;  // bit r: some lane of the wavefront evaluates row r directly
#pragma unroll
		for (int r = 0; r < PX_RPT; r++) {
			if (cls[r] == CLS_REFINE) cmask[r] |= 1u << c;
			if (COUNT) nev += cls[r] == CLS_DIRECT ? 1u : 0u;
			if (__ballot(cls[r] == CLS_DIRECT)) drow |= 1u << r;
		}
		if (COUNT || !drow) continue;
		// direct values: FP32 forms; kind and rows are wavefront-uniform branches
		float v32[PX_RPT];
#pragma unroll
		for (int r = 0; r < PX_RPT; r++) v32[r] = -1.f;
		auto rows32 = [&](auto kind_tag) {
			constexpr int K = decltype(kind_tag)::value;
#pragma unroll
			for (int r = 0; r < PX_RPT; r++)
				if ((drow >> r) & 1u) v32[r] = value32<K>(h, xf[r], yf[r]);
		};
		if (h.flags & PIX_FP32) {
			switch (h.kind) {
			case PK_PROFILE_SERSIC: rows32(IntTag<PK_PROFILE_SERSIC>{}); break;
			case PK_PROFILE_MOFFAT: rows32(IntTag<PK_PROFILE_MOFFAT>{}); break;
			case PK_PROFILE_BROKENEXP: rows32(IntTag<PK_PROFILE_BROKENEXP>{}); break;
			case PK_PROFILE_CORESERSIC: rows32(IntTag<PK_PROFILE_CORESERSIC>{}); break;
			case PK_PROFILE_KING:
			case PK_PROFILE_FERRER:
				// r^2 in FP64: their values vanish at the truncation radius
				// through a difference of nearly equal terms
#pragma unroll
				for (int r = 0; r < PX_RPT; r++) {
					if (!((drow >> r) & 1u)) continue;
					const double dx = x - h.xcen, dy = (hy + (j0 + r) * jb.sy) - h.ycen;
					const double xp = dx * cp.cos_ang + dy * cp.sin_ang;
					const double yp = (dy * cp.cos_ang - dx * cp.sin_ang) * cp.inv_axrat;
					const double r2 = radius2_64_kf(cp, xp, yp);
					v32[r] = h.kind == PK_PROFILE_KING ? king32(cp, r2) : ferrer32(cp, r2);
				}
				break;
			default: break;
			}
		}
		double v[PX_RPT];
		bool any_bad = false;
#pragma unroll
		for (int r = 0; r < PX_RPT; r++) {
			const bool direct = cls[r] == CLS_DIRECT;
			const bool ok = v32[r] >= 0.f && v32[r] <= 3e38f;
			v[r] = direct && ok ? (double)v32[r] * h.vscale : 0.0;
			any_bad = any_bad || (direct && !ok);
		}
		if (__ballot(any_bad)) {
			// no FP32 form, or its value out of range: FP64
#pragma unroll
			for (int r = 0; r < PX_RPT; r++) {
				if (cls[r] != CLS_DIRECT || (v32[r] >= 0.f && v32[r] <= 3e38f)) continue;
				double xp, yp;
				to_prof(cp, x, hy + (j0 + r) * jb.sy, xp, yp);
				v[r] = evaluate_at(cp, xp, yp) * cp.flux_scale;
			}
		}
#pragma unroll
		for (int r = 0; r < PX_RPT; r++) {
			if (conv) acc_c[r] += v[r];
			else acc_nc[r] += v[r];
		}
	}
	if (COUNT) {
		coarse_store_count(nev, wave_evals);
		return;
	}
	coarse_store<TI>(jb, jn, i, j0, inside, acc_nc, acc_c, cmask, img_nc, img_c, entries, counters, acc);
}

// ---------------------------------------------------------------------------
// refinement
// ---------------------------------------------------------------------------
constexpr int RQ_CAP = 768;
constexpr uint32_t REF_CHUNK = 16;

struct RefStack {
	double x0[RQ_CAP], x1[RQ_CAP], y0[RQ_CAP], y1[RQ_CAP];
	uint8_t lvl[RQ_CAP];
};

__device__ inline double wave_sum(double v)
{
	// fixed butterfly order: deterministic across runs
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
	return v;
}

// Integrates one component over one pixel; returns the pixel value in
// libprofit units (before flux_scale), identical on every lane.
template <bool FP32, bool GRAD = false>
__device__ double refine_pixel(const PkComp &cp, RefStack &st, double x0, double x1, double y0, double y1,
                               int res, int max_rec, uint32_t &overflow, unsigned long long &nev,
                               double gw = 0, double *gacc = nullptr)
{
	const int lane = (int)lane_id();
	if (res <= 0) return NAN;  // libprofit divides 0 by 0 (src/radial.cpp:134)
	const int r2 = res * res;
	const int tasks_per_pop = r2 >= 64 ? 1 : 64 / r2;
	const int rounds = (r2 + 63) / 64;
	const double r2d = (double)r2;
	const double acc_thr = cp.acc;
	const bool lane_fp32 = FP32 && cp.fp32;
	double total = 0;
	if (lane == 0) {
		st.x0[0] = x0;
		st.x1[0] = x1;
		st.y0[0] = y0;
		st.y1[0] = y1;
		st.lvl[0] = 0;
	}
	__syncthreads();
	int top = 1;
	while (top > 0) {
		const int ntask = top < tasks_per_pop ? top : tasks_per_pop;
		const int base = top - ntask;
		// every lane loads the cell of its task before anything is pushed
		const int my_task = r2 >= 64 ? 0 : lane / r2;
		const bool task_ok = my_task < ntask;
		const int slot = base + (task_ok ? my_task : 0);
		const double cx0 = st.x0[slot], cx1 = st.x1[slot], cy0 = st.y0[slot], cy1 = st.y1[slot];
		const int level = st.lvl[slot];
		__syncthreads();
		top = base;
		const double xbin = (cx1 - cx0) / res;
		const double ybin = (cy1 - cy0) / res;
		const double hxb = xbin / 2.;
		const double hyb = ybin / 2.;
		const bool recurse = res > 1 && level < max_rec;
		double w = 1.0;
		for (int k = 0; k <= level; k++) w /= r2d;
		for (int rd = 0; rd < rounds; rd++) {
			const int s = r2 >= 64 ? lane + 64 * rd : lane % r2;
			const bool active = task_ok && s < r2;
			bool fail = false;
			double x = 0, y = 0;
			if (active) {
				const int si = s / res, sj = s % res;
				// sub-cell centres accumulated exactly as src/radial.cpp:83-104
				x = cx0;
				for (int k = 0; k < si; k++) {
					x += hxb;
					x += hxb;
				}
				x += hxb;
				y = cy0;
				for (int k = 0; k < sj; k++) {
					y += hyb;
					y += hyb;
				}
				y += hyb;
				double xp, yp;
				to_prof(cp, x, y, xp, yp);
				double val;
				if (recurse) {
					const double dyp = (-xbin * cp.sin_ang + ybin * cp.cos_ang) / cp.axrat;
					const double txp = fabs(xp), typ = fabs(yp) + fabs(dyp);
					bool decided = false;
					if (lane_fp32) {
						const float v32 = evaluate_at32(cp, (float)xp, (float)yp, radius2_64(cp, xp, yp));
						const float t32 = evaluate_at32(cp, (float)txp, (float)typ, radius2_64(cp, txp, typ));
						const float m = fabsf(t32 / v32 - 1.f);
						const double guard = 1e-4 * (1.0 + acc_thr);
						if (v32 > 0.f && isfinite(m) && fabs((double)m - acc_thr) > guard) {
							fail = (double)m > acc_thr;
							val = (double)v32 * cp.unfold;
							decided = true;
						}
						nev += 2;
					}
					if (!decided) {
						val = evaluate_at(cp, xp, yp);
						const double tv = evaluate_at(cp, txp, typ);
						fail = fabs(tv / val - 1.0) > acc_thr;
						nev += 2;
					}
				} else {
					if (lane_fp32)
						val = (double)evaluate_at32(cp, (float)xp, (float)yp, radius2_64(cp, xp, yp)) * cp.unfold;
					else val = evaluate_at(cp, xp, yp);
					nev += 1;
				}
				if (!fail) total += val * w;
				if (GRAD && !fail) {
					// adjoint: the same accepted sub-cells, weight G * flux_scale * w * f
					double g[GT_COUNT];
					const double vg = sersic_grad64(cp, xp, yp, g);
					grad_accumulate(gacc, gw * w * vg, g);
				}
			}
			// compaction of failing sub-cells onto the LIFO
			const uint64_t ballot = __ballot(fail);
			if (ballot) {
				const int n = __popcll(ballot);
				if (top + n > RQ_CAP) {
					overflow = 1;
				} else {
					if (fail) {
						const int pos = top + __popcll(ballot & ((1ull << lane) - 1ull));
						st.x0[pos] = x - hxb;
						st.x1[pos] = x + hxb;
						st.y0[pos] = y - hyb;
						st.y1[pos] = y + hyb;
						st.lvl[pos] = (uint8_t)(level + 1);
					}
					top += n;
				}
			}
		}
		__syncthreads();
	}
	return wave_sum(total);
}

// GRAD: instead of adding the pixel values, accumulate the chi^2 adjoint of
// the same sub-cells (G_nc / G_c: adjoint images of the non-convolved /
// convolved drawing images) into gacc[component][GA_SLOTS].
template <typename TI, bool FP32, bool GRAD = false>
__global__ __launch_bounds__(64) void refine_kernel(const PkJob *__restrict__ jobs, const PkComp *__restrict__ comps,
                                                    const RefineEntry *__restrict__ entries,
                                                    PkCounters *__restrict__ counters, TI *__restrict__ img_nc,
                                                    TI *__restrict__ img_c, const TI *__restrict__ G_nc = nullptr,
                                                    const TI *__restrict__ G_c = nullptr,
                                                    double *__restrict__ gacc = nullptr)
{
	__shared__ RefStack st;
	const uint32_t n_entries = counters->n_entries;
	uint32_t overflow = 0;
	unsigned long long nev = 0;
	uint32_t chunk_next = 0, chunk_end = 0;
	const uint32_t chunk = max(1u, min(REF_CHUNK, n_entries / (8u * gridDim.x)));
	for (;;) {
		if (chunk_next >= chunk_end) {
			// claim a chunk of entries with one atomic on the queue head: up to
			// REF_CHUNK, fewer when there are few entries per wavefront
			uint32_t base = 0;
			if (lane_id() == 0) base = atomicAdd(&counters->next_entry, chunk);
			base = __shfl(base, 0);
			if (base >= n_entries) break;
			chunk_next = base;
			chunk_end = min(base + chunk, n_entries);
		}
		const uint32_t e = chunk_next++;
		const RefineEntry en = entries[e];
		const PkJob &jb = jobs[en.job];
		const uint32_t i = en.pix % jb.dw, j = en.pix / jb.dw;
		const double hx = jb.sx / 2., hy = jb.sy / 2.;
		const double x = hx + i * jb.sx, y = hy + j * jb.sy;
		double add_nc = 0, add_c = 0;
		uint32_t cm = en.cmask;
		while (cm) {
			const int c = __ffs(cm) - 1;
			cm &= cm - 1;
			const PkComp &cp = comps[jb.comp_off + c];
			int res = cp.res, mr = cp.max_rec;
			if (cp.center_res) {
				// SersicProfile::subsampling_params (src/sersic.cpp:331-347)
				const bool center = fabs(x - cp.xcen) < cp.win_x && fabs(y - cp.ycen) < cp.win_y;
				if (center) {
					res = cp.center_res;
					mr = cp.center_max_rec;
				}
			}
			if (GRAD) {
				const size_t p = jb.img_off + en.pix;
				const double gw = (double)(cp.convolve ? G_c[p] : G_nc[p]) * cp.flux_scale;
				double ga[GA_SLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
				(void)refine_pixel<FP32, true>(cp, st, x - hx, x + hx, y - hy, y + hy, res, mr, overflow, nev, gw, ga);
#pragma unroll
				for (int k = 0; k < GA_SLOTS; k++) {
					const double v = wave_sum(ga[k]);
					if (lane_id() == 0 && v != 0) atomicAdd(&gacc[(size_t)(jb.comp_off + c) * GA_SLOTS + k], v);
				}
				continue;
			}
			const double v = refine_pixel<FP32>(cp, st, x - hx, x + hx, y - hy, y + hy, res, mr, overflow, nev);
			if (cp.convolve) add_c += cp.flux_scale * v;
			else add_nc += cp.flux_scale * v;
		}
		if (!GRAD && lane_id() == 0) {
			if (jb.has_nc) img_nc[jb.img_off + en.pix] += (TI)add_nc;
			if (jb.has_conv) img_c[jb.img_off + en.pix] += (TI)add_c;
		}
	}
	// statistics and overflow flag, one atomic per wavefront
	for (int off = 32; off > 0; off >>= 1) nev += __shfl_xor(nev, off);
	const uint64_t ov = __ballot(overflow != 0);
	if (lane_id() == 0) {
		atomicAdd(&counters->n_subcells, nev);
		if (ov) atomicOr(&counters->overflow, 1u);
	}
}

// ---------------------------------------------------------------------------
// PSF profile (src/psf.cpp:58-145); one lane per image pixel of the item's box
// ---------------------------------------------------------------------------
__device__ inline uint32_t bind_u(double value, uint32_t max)
{
	int iv = (int)floor(value);
	if (iv < 0) return 0;
	uint32_t u = (uint32_t)iv;
	return u < max ? u : max;
}

template <typename TI>
__global__ void psf_kernel(const PkPsfItem *__restrict__ items, const PkJob *__restrict__ jobs,
                           const double *__restrict__ psfs, TI *__restrict__ img_nc, TI *__restrict__ img_c)
{
	const PkPsfItem it = items[blockIdx.y];
	const PkJob &jb = jobs[it.job];
	const uint32_t bw = it.x1 - it.x0 + 1, bh = it.y1 - it.y0 + 1;
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	if (it.x1 < it.x0 || it.y1 < it.y0 || t >= bw * bh) return;
	const uint32_t px = it.x0 + t % bw, py = it.y0 + t / bw;
	const double sx = jb.sx, sy = jb.sy, psx = it.psx, psy = it.psy;
	const double x = px * sx, y = py * sy;
	const int qx0 = (int)bind_u(floor((x - it.origin_x) / psx), it.pw - 1);
	const int qx1 = (int)bind_u(floor((x - it.origin_x + sx) / psx), it.ph - 1);
	const int qy0 = (int)bind_u(floor((y - it.origin_y) / psy), it.pw - 1);
	const int qy1 = (int)bind_u(floor((y - it.origin_y + sy) / psy), it.ph - 1);
	const double *psf = psfs + it.psf_off;
	double val = 0;
	for (int qy = qy0; qy <= qy1; qy++) {
		const double ppy = qy * psy + it.origin_y;
		for (int qx = qx0; qx <= qx1; qx++) {
			const double ppx = qx * psx + it.origin_x;
			const double ix = fmin(x + sx, ppx + psx) - fmax(x, ppx);
			const double iy = fmin(y + sy, ppy + psy) - fmax(y, ppy);
			const uint32_t q = (uint32_t)qx + (uint32_t)qy * it.pw;
			if (q < it.pw * it.ph) val += psf[q] * (ix * iy) / (psx * psy);
		}
	}
	TI *img = it.to_conv ? img_c : img_nc;
	img[jb.img_off + px + (uint64_t)py * jb.dw] += (TI)(val * it.scale);
}

// ---------------------------------------------------------------------------
// direct convolution
// ---------------------------------------------------------------------------
constexpr int CT_W = 64;   // tile width (one wavefront)
constexpr int CT_R = 8;    // output rows per lane
constexpr int CT_H = 32;   // tile height (4 wavefronts x CT_R rows)
constexpr int CT_KMIN = 5, CT_KMAX = 49;  // square odd PSFs with a compile-time loop nest

// Per-output sum of per-PSF-column partial sums: FP32 columns are combined in
// FP64, so a column of <= kh FMAs is the only FP32 rounding chain.
template <typename T> struct ColSum;
template <> struct ColSum<float> {
	double s = 0;
	__device__ void add(float v) { s += (double)v; }
	__device__ float get() const { return (float)s; }
};
template <> struct ColSum<double> {
	double s = 0;
	__device__ void add(double v) { s += v; }
	__device__ double get() const { return s; }
};

__device__ inline double fma_t(double a, double b, double c) { return fma(a, b, c); }
__device__ inline float fma_t(float a, float b, float c) { return fmaf(a, b, c); }

// One CT_W x CT_H tile of out(i,j) = sum_{l in [l0, l0+nl), k} src(i+k-hw, j+l-hh) krn(kw-1-k, kh-1-l).
// `kt` holds the PSF reversed and transposed: kt[k * kh + l] = krn(kw-1-k, kh-1-l),
// so the values of one PSF column are contiguous and uniform across the
// wavefront (scalar loads).  NL > 0 fixes nl (and l0 = 0) at compile time.
template <typename T>
struct ConvOut {
	T *__restrict__ dst;           // convolved image buffer
	T *__restrict__ out;           // returned images (jobs with fuse_out), or nullptr
	const T *__restrict__ img_nc;  // unconvolved drawing images added to them
	T nc_const;                    // the job's unconvolved value when it has nc_sky
};

// Unconvolved value of a job with nc_sky: its unconvolved skies summed in
// component order in FP64 and rounded, as the pixel pass would store it.
template <typename T>
__device__ inline T nc_sky_value(const PkJob &jb, const PkComp *__restrict__ comps)
{
	double v = 0;
	if (jb.nc_sky && comps)
		for (uint32_t k = jb.comp_off; k < jb.comp_off + jb.ncomp; k++) {
			const PixHot &h = comps[k].hot;
			if (h.kind == PK_PROFILE_SKY && !(h.flags & PIX_CONVOLVE)) v += h.vscale;
		}
	return (T)v;
}

// Unconvolved part of returned pixel (X, Y) of a job with fuse_out (0 when
// the job has none there, or the pixel lies outside the convolved frame).
template <typename T>
__device__ inline T conv_unconvolved(const PkJob &jb, const ConvOut<T> &o, int X, int Y)
{
	if (!(o.out && jb.fuse_out && jb.has_nc) || X >= (int)jb.cw || Y >= (int)jb.ch) return (T)0;
	const int mx = X - jb.mox, my = Y - jb.moy;
	if (mx < 0 || my < 0 || mx >= (int)jb.dw || my >= (int)jb.dh) return (T)0;
	if (jb.nc_sky) return o.nc_const;
	return o.img_nc[jb.img_off + mx + (size_t)my * jb.dw];
}

// Stores convolved pixel (X, Y) of job jb: into the convolved image, or -- a
// job with fuse_out -- as returned pixel (X, Y) with its unconvolved part nc
// added exactly as the composition does (sum in FP64, src/model.cpp:289-306).
template <typename T>
__device__ inline void conv_store(const PkJob &jb, const ConvOut<T> &o, int X, int Y, T v, T nc)
{
	if (o.out && jb.fuse_out) o.out[jb.out_off + X + (size_t)Y * jb.ow] = (T)((double)v + (double)nc);
	else o.dst[jb.conv_off + X + (size_t)Y * jb.cw] = v;
}

template <typename T, int NL>
__device__ inline void conv_tile(const PkJob &jb, const T *__restrict__ src, const T *__restrict__ kt, T *tile,
                                 int bx, int by, int l0, int nl_rt, bool accumulate, const uint8_t *__restrict__ masks,
                                 const ConvOut<T> &o)
{
	const int nl = NL > 0 ? NL : nl_rt;
	const int W = (int)jb.dw, H = (int)jb.dh, kw = (int)jb.kw, kh = (int)jb.kh;
	const int hw = kw / 2, hh = kh / 2;
	const int ox = bx + jb.csx, oy = by + jb.csy;  // tile origin in the same-size frame
	const int TW = CT_W + kw - 1, TH = CT_H + nl - 1;
	// source tile + halo, four loads in flight per thread; the row of element e
	// from a float reciprocal, corrected by one step (no integer division)
	const int n_el = TW * TH;
	const float inv_tw = 1.f / (float)TW;
	for (int e0 = (int)threadIdx.x; e0 < n_el; e0 += 4 * 256) {
		T v[4];
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int e = e0 + u * 256;
			v[u] = 0;
			if (e < n_el) {
				int ty = (int)(((float)e + 0.5f) * inv_tw);
				int tx = e - ty * TW;
				if (tx < 0) {
					ty--;
					tx += TW;
				} else if (tx >= TW) {
					ty++;
					tx -= TW;
				}
				const int si = ox + tx - hw, sj = oy + ty + l0 - hh;
				if (si >= 0 && si < W && sj >= 0 && sj < H) v[u] = src[si + (size_t)sj * W];
			}
		}
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int e = e0 + u * 256;
			if (e < n_el) tile[e] = v[u];
		}
	}
	__syncthreads();
	const int lx = threadIdx.x & 63, ly = (threadIdx.x >> 6) * CT_R;
	// a wavefront whose rows all lie below the image has no output (nothing
	// after this point waits for the workgroup)
	if (by + ly >= (int)jb.ch) return;
	// unconvolved parts of the lane's returned pixels, loaded before the sums
	// so that their latency hides behind them
	T ncv[CT_R];
#pragma unroll
	for (int r = 0; r < CT_R; r++) ncv[r] = conv_unconvolved(jb, o, bx + lx, by + ly + r);
	ColSum<T> acc[CT_R];
	const T *trow = tile + (size_t)ly * TW + lx;
	for (int k = 0; k < kw; k++) {
		const T *kc = kt + (size_t)k * kh + l0;
		T col[CT_R];
#pragma unroll
		for (int r = 0; r < CT_R; r++) col[r] = 0;
		if (NL > 0) {
#pragma unroll
			for (int yy = 0; yy < NL + CT_R - 1; yy++) {
				const T sv = trow[(size_t)yy * TW + k];
#pragma unroll
				for (int r = 0; r < CT_R; r++) {
					const int m = yy - r;
					if (m >= 0 && m < NL) col[r] = fma_t(sv, kc[m], col[r]);
				}
			}
		} else {
			for (int yy = 0; yy < nl + CT_R - 1; yy++) {
				const T sv = trow[(size_t)yy * TW + k];
#pragma unroll
				for (int r = 0; r < CT_R; r++) {
					const int m = yy - r;
					if (m >= 0 && m < nl) col[r] = fma_t(sv, kc[m], col[r]);
				}
			}
		}
#pragma unroll
		for (int r = 0; r < CT_R; r++) acc[r].add(col[r]);
	}
#pragma unroll
	for (int r = 0; r < CT_R; r++) {
		const int X = bx + lx, Y = by + ly + r;
		if (X < (int)jb.cw && Y < (int)jb.ch) {
			T v = acc[r].get();
			if (accumulate) v += o.dst[jb.conv_off + X + (size_t)Y * jb.cw];
			if (jb.mask_off != NONE) {
				const int mi = X - jb.cmx, mj = Y - jb.cmy;
				if (mi < 0 || mj < 0 || mi >= W || mj >= H || !masks[jb.mask_off + mi + (size_t)mj * W]) v = 0;
			}
			conv_store(jb, o, X, Y, v, ncv[r]);
		}
	}
}

template <typename T, int K>
__device__ inline bool conv_dispatch(int kw, const PkJob &jb, const T *src, const T *kt, T *tile, int bx, int by,
                                     const uint8_t *masks, const ConvOut<T> &o)
{
	if (kw == K) {
		conv_tile<T, K>(jb, src, kt, tile, bx, by, 0, K, false, masks, o);
		return true;
	}
	if constexpr (K + 2 <= CT_KMAX) return conv_dispatch<T, K + 2>(kw, jb, src, kt, tile, bx, by, masks, o);
	return false;
}

// Direct "same" convolution of every job (src/convolve.cpp:101-231).  With
// l0/nl the launch covers PSF rows [l0, l0+nl) only and, if l0 > 0, adds to
// dst (large PSFs are worked through in row chunks that fit in LDS).
// The launch bounds' second argument (one workgroup per CU at least) is
// deliberate: with it hipcc allocates 88 VGPRs for the FP32 kernel (five
// wavefronts per SIMD) instead of 153, and the kernel ran 4 % faster.
template <typename T>
__global__ __launch_bounds__(256, 1) void conv_kernel(const PkJob *__restrict__ jobs, uint32_t njobs,
                                                   const T *__restrict__ src_all, const T *__restrict__ krn_all,
                                                   const uint8_t *__restrict__ masks, T *__restrict__ dst_all,
                                                   int l0, int nl, const uint32_t *__restrict__ tile_job,
                                                   T *__restrict__ out_all, const T *__restrict__ img_nc,
                                                   uint32_t tile_base, const PkComp *__restrict__ comps)
{
	extern __shared__ unsigned char smem_raw[];
	T *tile = reinterpret_cast<T *>(smem_raw);
	const uint32_t gt = tile_base + blockIdx.x;  // convolution tile of this workgroup
	const uint32_t jn = find_job(jobs, njobs, gt, 1, tile_job);
	const PkJob &jb = jobs[jn];
	const uint32_t t = gt - jb.ctile_off;
	const int bx = (int)(t % jb.ctiles_x) * CT_W, by = (int)(t / jb.ctiles_x) * CT_H;
	const T *src = src_all + jb.img_off;
	const T *kt = krn_all + jb.psf_off;
	const ConvOut<T> o{dst_all, out_all, img_nc, nc_sky_value<T>(jb, comps)};
	const int kw = (int)jb.kw, kh = (int)jb.kh;
	if (l0 == 0 && nl >= kh && kw == kh && (kw & 1) && kw >= CT_KMIN && kw <= CT_KMAX &&
	    conv_dispatch<T, CT_KMIN>(kw, jb, src, kt, tile, bx, by, masks, o))
		return;
	const int n = nl < kh - l0 ? nl : kh - l0;
	if (n <= 0) return;
	conv_tile<T, 0>(jb, src, kt, tile, bx, by, l0, n, l0 > 0, masks, o);
}

// ---------------------------------------------------------------------------
// direct convolution on the matrix cores (FP32 mode)
//
// For one PSF row l, a 16 x 16 block of outputs is a matrix product
//   out[y][x] += sum_j A_l[y][j] B_l[j][x],
//   A_l[y][j] = tile(y + l, x0 + j)   (a window of the source tile in LDS),
//   B_l[j][x] = K'(j - x, l)          (the PSF row as a band, 0 off the band),
// with j over the 16 + kw - 1 source columns the block reads: the direct
// convolution recast as a tiled contraction for v_mfma_f32_16x16x4_f32
// (exact FP32 products).  The band costs (16 + kw - 1) / kw times the useful
// multiply-adds.  The FP32 accumulators of each PSF row are added in FP64, so
// a row of <= (16 + kw - 1) products is the only FP32 rounding chain, as in
// conv_kernel.  Same tiles, geometry and mask as conv_kernel.
// Lane layout of the 16x16x4 FP32 product: A[i][k] from lane i + 16 k,
// B[k][j] from lane j + 16 k, D[i][j] in lane j + 16 (i / 4), register i % 4.
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));

// LDS layout of the matrix-core form.  Source tile: (CT_W + kw - 1) columns
// of source plus CM_PAD zero columns (a block reads up to 3 columns past its
// band), odd row stride so that the 16 rows one operand read touches fall in
// distinct banks.  PSF band: row l holds K'(d, l) at d + 16 for
// d in [-16, kw + 20), zero outside [0, kw), so that B[j][x] = K'(j - x, l) is
// one read without a bounds test.
constexpr int CM_PAD = 4;
// PSF rows whose products are summed in FP32 before the FP64 sum: per output,
// at most 3 kw non-zero products of one sign (the band's zeros add exactly),
// worst-case rounding 3 kw 2^-24 relative, below 1e-5 for kw <= 55
constexpr int CM_ROWS32 = 3;
__host__ __device__ inline int conv_mfma_stride(int kw) { return (CT_W + kw - 1 + CM_PAD) | 1; }
__host__ __device__ inline int conv_mfma_band(int kw) { return kw + 36; }

__host__ __device__ inline size_t conv_mfma_lds_bytes(int kw, int kh)
{
	return sizeof(float) *
	       ((size_t)conv_mfma_stride(kw) * (CT_H + kh - 1) + (size_t)conv_mfma_band(kw) * (size_t)kh);
}

// Products of all PSF rows for this wavefront's two 16 x 16 output blocks
// (rows ry0.., columns cx0.. and cx0 + 16..; TWO = false: the first only, the
// second lies right of the image).  NJ > 0: 4 * NJ source columns per row
// known at compile time (fully unrolled, operands read before the products);
// NJ = 0: run-time count.
template <int NJ, bool TWO>
__device__ inline void conv_mfma_rows(const float *tile, const float *band, int TS, int BS, int kh, int nj_rt,
                                      int lane, int ry0, int cx0, double out[2][4])
{
	const int lo = lane & 15, kq = lane >> 4;
	for (int l0 = 0; l0 < kh; l0 += CM_ROWS32) {
	floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
	const int l1 = l0 + CM_ROWS32 < kh ? l0 + CM_ROWS32 : kh;
	for (int l = l0; l < l1; l++) {
		const float *arow = tile + (size_t)(ry0 + lo + l) * TS + cx0 + kq;  // A[lo][j], j = 4 jc + kq
		const float *brow = band + (size_t)l * BS + 16 + kq - lo;           // B[j][lo] = K'(j - lo, l)
		if (NJ > 0) {
			float a0[NJ > 0 ? NJ : 1], a1[NJ > 0 ? NJ : 1], b[NJ > 0 ? NJ : 1];
#pragma unroll
			for (int jc = 0; jc < NJ; jc++) {
				a0[jc] = arow[4 * jc];
				if (TWO) a1[jc] = arow[16 + 4 * jc];
				b[jc] = brow[4 * jc];
			}
#pragma unroll
			for (int jc = 0; jc < NJ; jc++) {
				acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[jc], b[jc], acc0, 0, 0, 0);
				if (TWO) acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[jc], b[jc], acc1, 0, 0, 0);
			}
		} else {
			for (int jc = 0; jc < nj_rt; jc++) {
				const float bv = brow[4 * jc];
				acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(arow[4 * jc], bv, acc0, 0, 0, 0);
				if (TWO) acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(arow[16 + 4 * jc], bv, acc1, 0, 0, 0);
			}
		}
	}
#pragma unroll
	for (int r = 0; r < 4; r++) {
		out[0][r] += (double)acc0[r];
		out[1][r] += (double)acc1[r];
	}
	}
}

// NJ = ceil((16 + kw - 1) / 4) for the odd PSF widths CT_KMIN..CT_KMAX
template <int K, bool TWO>
__device__ inline bool conv_mfma_dispatch(int kw, const float *tile, const float *band, int TS, int BS, int kh,
                                          int lane, int ry0, int cx0, double out[2][4])
{
	if (kw == K) {
		conv_mfma_rows<(16 + K - 1 + 3) / 4, TWO>(tile, band, TS, BS, kh, 0, lane, ry0, cx0, out);
		return true;
	}
	if constexpr (K + 2 <= CT_KMAX)
		return conv_mfma_dispatch<K + 2, TWO>(kw, tile, band, TS, BS, kh, lane, ry0, cx0, out);
	return false;
}

#ifndef PK_STRIP_KERNELS_ONLY  // (pk_strips.hip: the strip-form kernels alone)
__global__ __launch_bounds__(256, 1) void conv_mfma_kernel(const PkJob *__restrict__ jobs, uint32_t njobs,
                                                           const float *__restrict__ src_all,
                                                           const float *__restrict__ krn_all,
                                                           const uint8_t *__restrict__ masks,
                                                           float *__restrict__ dst_all,
                                                           const uint32_t *__restrict__ tile_job,
                                                           float *__restrict__ out_all,
                                                           const float *__restrict__ img_nc, uint32_t tile_base,
                                                           const PkComp *__restrict__ comps)
{
	extern __shared__ unsigned char smem_raw[];
	float *tile = reinterpret_cast<float *>(smem_raw);
	const uint32_t gt = tile_base + blockIdx.x;  // convolution tile of this workgroup
	const uint32_t jn = find_job(jobs, njobs, gt, 1, tile_job);
	const PkJob &jb = jobs[jn];
	const uint32_t t = gt - jb.ctile_off;
	const int bx = (int)(t % jb.ctiles_x) * CT_W, by = (int)(t / jb.ctiles_x) * CT_H;
	const float *src = src_all + jb.img_off;
	const float *krn = krn_all + jb.psf_off;
	const ConvOut<float> o{dst_all, out_all, img_nc, nc_sky_value<float>(jb, comps)};
	const int W = (int)jb.dw, H = (int)jb.dh, kw = (int)jb.kw, kh = (int)jb.kh;
	const int hw = kw / 2, hh = kh / 2;
	const int ox = bx + jb.csx, oy = by + jb.csy;
	const int TW = CT_W + kw - 1, TH = CT_H + kh - 1, TS = conv_mfma_stride(kw), BS = conv_mfma_band(kw);
	float *band = tile + (size_t)TS * TH;
	// source tile (zero outside the image and in the pad columns), four loads
	// in flight per thread; row of element e by float reciprocal + one step
	{
		const int n_el = TS * TH;
		const float inv_ts = 1.f / (float)TS;
		for (int e0 = (int)threadIdx.x; e0 < n_el; e0 += 4 * 256) {
			float v[4];
#pragma unroll
			for (int u = 0; u < 4; u++) {
				const int e = e0 + u * 256;
				v[u] = 0.f;
				if (e < n_el) {
					int ty = (int)(((float)e + 0.5f) * inv_ts);
					int tx = e - ty * TS;
					if (tx < 0) {
						ty--;
						tx += TS;
					} else if (tx >= TS) {
						ty++;
						tx -= TS;
					}
					const int si = ox + tx - hw, sj = oy + ty - hh;
					if (tx < TW && si >= 0 && si < W && sj >= 0 && sj < H) v[u] = src[si + (size_t)sj * W];
				}
			}
#pragma unroll
			for (int u = 0; u < 4; u++) {
				const int e = e0 + u * 256;
				if (e < n_el) tile[e] = v[u];
			}
		}
	}
	// PSF band: band[l * BS + c] = K'(c - 16, l); kt[k * kh + l] = K'(k, l)
	for (int e = (int)threadIdx.x; e < kh * BS; e += 256) {
		const int l = e / BS, d = e - l * BS - 16;
		band[e] = (d >= 0 && d < kw) ? krn[(size_t)d * kh + l] : 0.f;
	}
	__syncthreads();

	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const int ry0 = (w >> 1) * 16, cx0 = (w & 1) * 32;  // this wave: rows ry0.., columns cx0.. (two 16 x 16 blocks)
	// no output rows or columns: nothing below waits for the workgroup
	if (by + ry0 >= (int)jb.ch || bx + cx0 >= (int)jb.cw) return;
	const int lo = lane & 15, kq = lane >> 4;
	double out[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
	const bool odd = (kw & 1) && kw >= CT_KMIN && kw <= CT_KMAX;
	if (bx + cx0 + 16 < (int)jb.cw) {
		if (!(odd && conv_mfma_dispatch<CT_KMIN, true>(kw, tile, band, TS, BS, kh, lane, ry0, cx0, out)))
			conv_mfma_rows<0, true>(tile, band, TS, BS, kh, (16 + kw - 1 + 3) / 4, lane, ry0, cx0, out);
	} else {
		if (!(odd && conv_mfma_dispatch<CT_KMIN, false>(kw, tile, band, TS, BS, kh, lane, ry0, cx0, out)))
			conv_mfma_rows<0, false>(tile, band, TS, BS, kh, (16 + kw - 1 + 3) / 4, lane, ry0, cx0, out);
	}
#pragma unroll
	for (int s2 = 0; s2 < 2; s2++)
#pragma unroll
		for (int r = 0; r < 4; r++) {
			const int X = bx + cx0 + 16 * s2 + lo, Y = by + ry0 + 4 * kq + r;
			if (X < (int)jb.cw && Y < (int)jb.ch) {
				float v = (float)out[s2][r];
				if (jb.mask_off != NONE) {
					const int mi = X - jb.cmx, mj = Y - jb.cmy;
					if (mi < 0 || mj < 0 || mi >= W || mj >= H || !masks[jb.mask_off + mi + (size_t)mj * W]) v = 0.f;
				}
				conv_store(jb, o, X, Y, v, conv_unconvolved(jb, o, X, Y));
			}
		}
}
#endif

// ---------------------------------------------------------------------------
// Matrix-core convolution by strips.  A wavefront owns a strip of 16 output
// columns and up to ST_H rows and streams the source rows through it.  For
// source row t of the strip (row 0: the strip's first output row shifted up by
// the PSF's half height),
//   D_t[l][x] = sum_d K'(d, l) S(t, x + d),   l < 16 C, d < 4 NJ,
// is one chain of NJ v_mfma_f32_16x16x4f32 per 16 PSF rows (A: the PSF, in
// registers for the whole strip; B: the source row, from LDS), and D_t[l][x]
// is the part of output row t - l that source row t gives.  Of the products,
// only the padding of kw to 4 NJ and of kh to 16 C is not a PSF tap, against
// 16 + kw - 1 (rounded up to 4) contracted columns for kw taps in the banded
// form; the cost is the source rows above and below the strip's rows, which
// a block of 16 PSF rows skips where it adds to none of them.
//
// Diagonal sums.  Lane (x, q) of D holds the rows l = 16 c + 4 q + r, r < 4.
// It adds them into a ring of R = 16 C running sums, the sum of output row y
// in slot (y + 4 q) mod R, i.e. slot (t - 16 c - r) mod R at step t: a
// compile-time index in a step loop unrolled R times.  The lane's last term of
// row y comes at step y + 4 q + 16 (C - 1) + 3 and the slot's next row starts
// at step y + R + 4 q, so at step t the lane takes slot (t + 1) mod R out:
// lane group q hands over row t + 1 - R - 4 q to a 32-row LDS stage.  Row y
// has all four hand-overs after step y + R + 11; rows are summed over the
// groups and stored 16 at a time (at fixed steps of the unrolled loop, as R
// is a multiple of 16), with the unconvolved parts of the returned pixels
// loaded one batch ahead.
// Tail rows.  A PSF of kh = 16 C + TM rows, 0 < TM <= ST_TAIL, takes C
// blocks of matrix-core steps and its last TM rows on the vector ALU: lane
// (x, q) sums K'(4 j + q, 16 C + i) S(t, x + 4 j + q) over j (the B operands it
// holds) and adds the sum to its slot of output row t - 16 C - i, one of four
// compile-time slots by lane group, so the ring is that of C + 1 blocks
// (R = 16 (C + 1)) and the hand-over is unchanged.  Without it the last block
// would be 16 rows of products for TM PSF rows.
// Rounding: per output a chain of NJ matrix-core steps (4 NJ additions), at
// most 4 C ring additions and 3 group additions, so for terms of one sign the
// FP32 error stays below (4 NJ + 4 C + 3) 2^-24 = 3.8e-6 of the sum (kw, kh <= 48);
// a tail row's NJ - 1 additions and one ring addition are fewer.
// ---------------------------------------------------------------------------
constexpr int ST_TAIL = 3;   // most tail rows (PSF rows past the last full block of 16 taken on the ALU)
constexpr int ST_H = 512;    // most output rows of a strip tile (4 strips of 16 columns); an image is split
                             // into equal tiles of at most ST_H rows (PkJob::st_rows)
constexpr int ST_KMAX = 48;  // largest PSF side of the strip form (NJ <= 12, C <= 3)
constexpr int ST_PF = 4;     // source rows loaded ahead

struct StripLds {
	float row[2][64];         // source rows t and t + 1, columns x0 - kw/2 + 0..63
	float stage[32][4][16];   // hand-overs of the lane groups, by output row mod 32
};

template <int NJ, int C, int TM>
__device__ inline void conv_strip(const PkJob &jb, const float *__restrict__ src, const float *__restrict__ krn,
                                  const uint8_t *__restrict__ masks, const ConvOut<float> &o, int x0, int y0,
                                  int nrows, StripLds &L, int lane)
{
	constexpr int R = 16 * (C + (TM > 0));
	static_assert(R % ST_PF == 0 && R % 16 == 0, "ring length");
	static_assert(TM >= 0 && TM <= ST_TAIL && 16 * C + TM <= ST_KMAX, "tail rows");
	const int W = (int)jb.dw, H = (int)jb.dh, kw = (int)jb.kw, kh = (int)jb.kh;
	const int lo = lane & 15, kq = lane >> 4;
	// A operand: lane (l, q) of step j of row block c holds K'(4 j + q, 16 c + l)
	float kr[C][NJ];
#pragma unroll
	for (int c = 0; c < C; c++)
#pragma unroll
		for (int j = 0; j < NJ; j++) {
			const int d = 4 * j + kq, l = 16 * c + lo;
			kr[c][j] = (d < kw && l < kh) ? krn[(size_t)d * kh + l] : 0.f;
		}
	// tail rows: lane (x, q) holds K'(4 j + q, 16 C + i)
	float ktl[TM > 0 ? TM : 1][NJ];
#pragma unroll
	for (int i = 0; i < TM; i++)
#pragma unroll
		for (int j = 0; j < NJ; j++) {
			const int d = 4 * j + kq, l = 16 * C + i;
			ktl[i][j] = (d < kw && l < kh) ? krn[(size_t)d * kh + l] : 0.f;
		}
	// source row t: image row sy0 + t, columns sx0 + lane, zero outside the
	// image; loads from a clamped address (no branch), zeroed when written to LDS
	const int sx0 = x0 + jb.csx - kw / 2, sy0 = y0 + jb.csy - kh / 2;
	const int sc = sx0 + lane;
	const bool col_in = sc >= 0 && sc < W;
	const float *colp = src + (col_in ? sc : 0);
	const int nsrc = nrows + kh - 1;    // steps with products
	const int nsteps = nrows + R + 11;  // and the hand-over of the last row
	auto load_row = [&](int t) -> float {
		const int sr = min(max(sy0 + t, 0), H - 1);
		return colp[(size_t)sr * W];
	};
	auto row_in = [&](int t) -> bool { return t < nsrc && sy0 + t >= 0 && sy0 + t < H; };
	// row r is kept in pf[(r - 1) % ST_PF] until it is written to LDS at step r - 1
	float pf[ST_PF];
#pragma unroll
	for (int u = 0; u < ST_PF; u++) pf[u] = load_row(u + 1);
	{
		const float v0 = load_row(0);
		L.row[0][lane] = (col_in && row_in(0)) ? v0 : 0.f;
	}
	// B operands of the current row, read one step ahead
	const float *rb0 = &L.row[0][lo + kq];
	float b[NJ];
#pragma unroll
	for (int j = 0; j < NJ; j++) b[j] = rb0[4 * j];
	// output: rows are stored 16 at a time, lane (x, q) rows yb + q + 4 i,
	// each the sum of the four groups' hand-overs, as conv_store and
	// conv_unconvolved do it, from values taken before the first store (a
	// store could alias the job record and reload its fields); the
	// unconvolved parts of the next 16 rows are loaded one batch ahead
	const int X = x0 + lo;
	const bool col_out = X < (int)jb.cw;
	const bool fuse = o.out && jb.fuse_out;
	float *const outp = fuse ? o.out + jb.out_off + X : o.dst + jb.conv_off + X;
	const size_t ostride = fuse ? jb.ow : jb.cw;
	const bool with_nc = fuse && jb.has_nc && X - jb.mox >= 0 && X - jb.mox < (int)jb.dw;
	const float *const ncb = with_nc ? o.img_nc + jb.img_off + (X - jb.mox) : o.img_nc;
	const int nc_y0 = y0 - jb.moy, nc_h = (int)jb.dh;
	const size_t nc_stride = jb.dw;
	const bool nc_const = jb.nc_sky != 0;
	const bool masked = jb.mask_off != NONE;
	float ncp[4];
	auto load_nc = [&](int yb) {
#pragma unroll
		for (int i = 0; i < 4; i++) {
			const int yy = yb + kq + 4 * i, my = nc_y0 + yy;
			const bool in = with_nc && yy < nrows && my >= 0 && my < nc_h;
			ncp[i] = !in ? 0.f : nc_const ? o.nc_const : ncb[(size_t)my * nc_stride];
		}
	};
	auto flush = [&](int yb) {
		float v[4];
#pragma unroll
		for (int i = 0; i < 4; i++) {
			const float *g = &L.stage[(yb + kq + 4 * i) & 31][0][lo];
			v[i] = ((g[0] + g[16]) + g[32]) + g[48];
		}
		float nc[4];
#pragma unroll
		for (int i = 0; i < 4; i++) nc[i] = ncp[i];
		load_nc(yb + 16);
#pragma unroll
		for (int i = 0; i < 4; i++) {
			const int yy = yb + kq + 4 * i, Y = y0 + yy;
			if (yy < nrows && col_out) {
				float w = v[i];
				if (masked) {
					const int mi = X - jb.cmx, mj = Y - jb.cmy;
					if (mi < 0 || mj < 0 || mi >= W || mj >= H || !masks[jb.mask_off + mi + (size_t)mj * W]) w = 0.f;
				}
				outp[(size_t)Y * ostride] = fuse ? (float)((double)w + (double)nc[i]) : w;
			}
		}
	};
	load_nc(0);
	float ring[R];
#pragma unroll
	for (int u = 0; u < R; u++) ring[u] = 0.f;
	// one step: row t + 1 to LDS, its operands read, the products of row t,
	// hand-over.  EDGE (the steps near the strip's ends): row block c only
	// for t in [16 c, nrows + 16 c + 15), where it adds to the strip's rows,
	// and t < nsrc; else all row blocks, their products interleaved.  Each
	// interior step is a block of its own (the test of t against `guard`, a
	// copy of nsrc the compiler cannot see through, so it cannot drop the
	// test): without that the compiler interleaves the products of many
	// steps and runs out of registers.
	int guard = nsrc;
	asm volatile("" : "+s"(guard));
	// the tail rows' part of source row t: row t - 16 C - i is in slot
	// (t - 16 C - i + 4 q) mod R of lane group q (in the interior steps also
	// where it adds to no row of the strip: such a row is never stored)
	auto tail = [&](auto tag) __attribute__((always_inline)) {
		constexpr int s = decltype(tag)::value;
#pragma unroll
		for (int i = 0; i < TM; i++) {
			float p = ktl[i][0] * b[0];
#pragma unroll
			for (int j = 1; j < NJ; j++) p = fmaf(ktl[i][j], b[j], p);
#pragma unroll
			for (int q = 0; q < 4; q++) ring[((s - 16 * C - i + 4 * q) % R + R) % R] += kq == q ? p : 0.f;
		}
	};
	auto step = [&](auto tag, auto edge, int tb) __attribute__((always_inline)) {
		constexpr int s = decltype(tag)::value;
		constexpr bool EDGE = decltype(edge)::value;
		const int t = tb + s;
		L.row[(s + 1) & 1][lane] = (col_in && row_in(t + 1)) ? pf[s % ST_PF] : 0.f;
		pf[s % ST_PF] = load_row(t + 1 + ST_PF);
		float bn[NJ];
		const float *rb = &L.row[(s + 1) & 1][lo + kq];
#pragma unroll
		for (int j = 0; j < NJ; j++) bn[j] = rb[4 * j];
		if constexpr (EDGE) {
#pragma unroll
			for (int c = 0; c < C; c++) {
				if (t < nsrc && t >= 16 * c && t < nrows + 16 * c + 15) {
					floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
					for (int j = 0; j < NJ; j++)
						acc = __builtin_amdgcn_mfma_f32_16x16x4f32(kr[c][j], b[j], acc, 0, 0, 0);
#pragma unroll
					for (int r = 0; r < 4; r++) ring[((s - 16 * c - r) % R + R) % R] += acc[r];
				}
			}
			if (TM > 0 && t < nsrc && t >= 16 * C && t < nrows + 16 * C + TM - 1) tail(tag);
		} else if (t < guard) {
			floatx4 acc[C];
#pragma unroll
			for (int c = 0; c < C; c++) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
			for (int j = 0; j < NJ; j++)
#pragma unroll
				for (int c = 0; c < C; c++) acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(kr[c][j], b[j], acc[c], 0, 0, 0);
#pragma unroll
			for (int c = 0; c < C; c++)
#pragma unroll
				for (int r = 0; r < 4; r++) ring[((s - 16 * c - r) % R + R) % R] += acc[c][r];
			tail(tag);
		}
#pragma unroll
		for (int j = 0; j < NJ; j++) b[j] = bn[j];
		// hand-over: group q gives row t + 1 - R - 4 q
		constexpr int sl = (s + 1) % R;
		L.stage[(t + 1 - R - 4 * kq) & 31][kq][lo] = ring[sl];
		ring[sl] = 0.f;
		// rows up to t + 1 - R - 12 have all four hand-overs; the last of a
		// batch of 16 at s = 10 (mod 16), as tb and R are multiples of 16
		if constexpr (s % 16 == 10) {
			const int y3 = t + 1 - R - 12;
			if (y3 >= 0) flush(y3 - 15);
		}
	};
	// interior steps: all row blocks add to the strip's rows, t in
	// [16 (C - 1), nrows + 15) and < nsrc; edge steps before and after (three
	// loops: one loop holding both forms of the step runs out of registers)
	const int t_full = min(nsrc, nrows + 15);
	const int tb_full = C == 1 ? 0 : R;  // first multiple of R >= 16 (C - 1)
	int tb = 0;
	for (; tb < tb_full && tb < nsteps; tb += R)
		static_for<R>([&](auto tag) __attribute__((always_inline)) { step(tag, IntTag<1>{}, tb); });
	for (; tb + R <= t_full; tb += R)
		static_for<R>([&](auto tag) __attribute__((always_inline)) { step(tag, IntTag<0>{}, tb); });
	for (; tb < nsteps; tb += R)
		static_for<R>([&](auto tag) __attribute__((always_inline)) { step(tag, IntTag<1>{}, tb); });
	// the last, partial batch of rows
	const int y_last = tb - R - 12;  // last row with all hand-overs
	if ((y_last & 15) != 15) flush(y_last & ~15);
}

// One workgroup: 4 strips (64 columns) of one strip tile; the wavefronts work
// independently (their own LDS, no barrier).
template <int NJ, int C, int TM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(C + (TM > 0) < 3 ? 4 : NJ < 12 ? 3 : 2))) void
conv_strip_kernel(
    const PkJob *__restrict__ jobs, uint32_t njobs, const float *__restrict__ src_all,
    const float *__restrict__ krn_all, const uint8_t *__restrict__ masks, float *__restrict__ dst_all,
    const uint32_t *__restrict__ tile_job, float *__restrict__ out_all, const float *__restrict__ img_nc,
    uint32_t tile_base, const PkComp *__restrict__ comps)
{
	__shared__ StripLds lds[4];
	const uint32_t gt = tile_base + blockIdx.x;
	const PkJob &jb = jobs[find_job(jobs, njobs, gt, 3, tile_job)];
	const uint32_t t = gt - jb.stile_off;
	const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
	const int x0 = (int)(t % jb.ctiles_x) * CT_W + 16 * w, y0 = (int)(t / jb.ctiles_x) * (int)jb.st_rows;
	if (x0 >= (int)jb.cw || y0 >= (int)jb.ch) return;
	const int nrows = min((int)jb.st_rows, (int)jb.ch - y0);
	const ConvOut<float> o{dst_all, out_all, img_nc, nc_sky_value<float>(jb, comps)};
	conv_strip<NJ, C, TM>(jb, src_all + jb.img_off, krn_all + jb.psf_off, masks, o, x0, y0, nrows, lds[w], lane);
}

#ifndef PK_STRIP_KERNELS_ONLY
// ---------------------------------------------------------------------------
// composition: returned(i,j) = sum over the fs x fs block of
//   conv(X,Y) + model(X - mox, Y - moy), (X,Y) = crop origin + fine position
// ---------------------------------------------------------------------------
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void compose_kernel(const PkJob *__restrict__ jobs, uint32_t njobs,
                                                      const TI *__restrict__ img_nc,
                                                      const TI *__restrict__ conv_all,
                                                      const uint8_t *__restrict__ omasks, TO *__restrict__ out_all,
                                                      const uint32_t *__restrict__ tile_job)
{
	const uint32_t jn = find_job(jobs, njobs, blockIdx.x, 2, tile_job);
	const PkJob &jb = jobs[jn];
	const uint32_t t = blockIdx.x - jb.otile_off;
	const uint32_t i = (t % jb.otiles_x) * 64 + (threadIdx.x & 63);
	const uint32_t j0 = (t / jb.otiles_x) * PX_ROWS + (threadIdx.x >> 6) * PX_RPT;
	if (i >= jb.ow) return;
	if (jb.fs == 1 && jb.omask_off == NONE) {
		// no downsampling, no final mask (the batch case): all loads of the
		// lane's rows are issued before the first sum
		TI pc[PX_RPT], pn[PX_RPT];
		const int X = jb.cox + (int)i;
		const int mx = X - jb.mox;
		const bool xin = X >= 0 && X < (int)jb.cw && (int)i < (int)jb.crw;
		const bool mxin = mx >= 0 && mx < (int)jb.dw;
#pragma unroll
		for (int r = 0; r < PX_RPT; r++) {
			const uint32_t j = j0 + r;
			const int Y = jb.coy + (int)j, my = Y - jb.moy;
			const bool in = xin && j < jb.oh && (int)j < (int)jb.crh && Y >= 0 && Y < (int)jb.ch;
			pc[r] = (jb.has_conv && in) ? conv_all[jb.conv_off + X + (size_t)Y * jb.cw] : (TI)0;
			pn[r] = (jb.has_nc && in && mxin && my >= 0 && my < (int)jb.dh) ? img_nc[jb.img_off + mx + (size_t)my * jb.dw]
			                                                               : (TI)0;
		}
#pragma unroll
		for (int r = 0; r < PX_RPT; r++) {
			const uint32_t j = j0 + r;
			if (j < jb.oh) out_all[jb.out_off + i + (size_t)j * jb.ow] = (TO)((double)pc[r] + (double)pn[r]);
		}
		return;
	}
#pragma unroll
	for (int r = 0; r < PX_RPT; r++) {
		const uint32_t j = j0 + r;
		if (j >= jb.oh) break;
		double v = 0;
		for (uint32_t b = 0; b < jb.fs; b++) {
			const uint32_t fy = j * jb.fs + b;
			if (fy >= jb.crh) break;
			for (uint32_t a = 0; a < jb.fs; a++) {
				const uint32_t fx = i * jb.fs + a;
				if (fx >= jb.crw) break;
				const int X = jb.cox + (int)fx, Y = jb.coy + (int)fy;
				if (X < 0 || Y < 0 || X >= (int)jb.cw || Y >= (int)jb.ch) continue;
				double p = 0;
				if (jb.has_conv) p += (double)conv_all[jb.conv_off + X + (size_t)Y * jb.cw];
				if (jb.has_nc) {
					const int mx = X - jb.mox, my = Y - jb.moy;
					if (mx >= 0 && my >= 0 && mx < (int)jb.dw && my < (int)jb.dh)
						p += (double)img_nc[jb.img_off + mx + (size_t)my * jb.dw];
				}
				v += p;
			}
		}
		if (jb.omask_off != NONE) {
			const uint32_t mi = i / jb.omask_f, mj = j / jb.omask_f;
			if (mi < jb.omask_w && mj < jb.omask_h && !omasks[jb.omask_off + mi + (size_t)mj * jb.omask_w]) v = 0;
		}
		out_all[jb.out_off + i + (size_t)j * jb.ow] = (TO)v;
	}
}

// ---------------------------------------------------------------------------
// chi^2 adjoint
// ---------------------------------------------------------------------------

// Sums v[0..NV) over a 256-lane workgroup; thread 0 gets the totals.
template <int NV>
__device__ inline void block_sum_256(double v[NV])
{
	__shared__ double red[4][NV];
#pragma unroll
	for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
	const int w = threadIdx.x >> 6;
	if ((threadIdx.x & 63) == 0)
#pragma unroll
		for (int k = 0; k < NV; k++) red[w][k] = v[k];
	__syncthreads();
	if (threadIdx.x == 0)
#pragma unroll
		for (int k = 0; k < NV; k++) v[k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
	__syncthreads();
}

// Residuals of the returned images and their adjoint on the drawing frame:
//   chi^2 = sum_p (data_p - model_p)^2 ivar_p,   G = d chi^2 / d drawing pixel.
// The returned image is the crop of the drawing frame, SUM-downsampled by fs
// (src/model.cpp:289-311), so every drawing pixel of the crop gets
// -2 (data - model) ivar of its returned pixel and pixels outside get 0.
template <typename TI>
__global__ __launch_bounds__(256) void resid_kernel(const PkJob *__restrict__ jobs, uint32_t njobs,
                                                    const TI *__restrict__ out_all, const TI *__restrict__ data,
                                                    const TI *__restrict__ ivar, TI *__restrict__ G,
                                                    double *__restrict__ chi2, const uint32_t *__restrict__ tile_job,
                                                    const uint8_t *__restrict__ masks)
{
	const uint32_t jn = find_job(jobs, njobs, blockIdx.x, 0, tile_job);
	const PkJob &jb = jobs[jn];
	const uint32_t t = blockIdx.x - jb.tile_off;
	const uint32_t i = (t % jb.tiles_x) * 64 + (threadIdx.x & 63);
	const uint32_t j0 = (t / jb.tiles_x) * PX_ROWS + (threadIdx.x >> 6) * PX_RPT;
	double c2[1] = {0};
#pragma unroll
	for (int r = 0; r < PX_RPT; r++) {
		const uint32_t j = j0 + r;
		if (!(i < jb.dw && j < jb.dh)) continue;
		double g = 0;
		// drawing pixel i is pixel i + mox of the convolved frame, whose crop starts at cox
		const int fx = (int)i + jb.mox - jb.cox, fy = (int)j + jb.moy - jb.coy;
		if (fx >= 0 && fy >= 0 && fx < (int)jb.crw && fy < (int)jb.crh) {
			const uint32_t oi = (uint32_t)fx / jb.fs, oj = (uint32_t)fy / jb.fs;
			const size_t p = jb.out_off + oi + (size_t)oj * jb.ow;
			const double rr = (double)data[p] - (double)out_all[p];
			const double iv = (double)ivar[p];
			g = -2 * rr * iv;
			if ((uint32_t)fx % jb.fs == 0 && (uint32_t)fy % jb.fs == 0) c2[0] += rr * rr * iv;
			// a returned pixel the final mask zeroes, or a drawing pixel the
			// calculation mask leaves out (neither its profiles nor its
			// convolved value reach the image), carries no gradient
			if (jb.omask_off != NONE) {
				const uint32_t mi = oi / jb.omask_f, mj = oj / jb.omask_f;
				if (mi < jb.omask_w && mj < jb.omask_h && !masks[jb.omask_off + mi + (size_t)mj * jb.omask_w]) g = 0;
			}
			if (jb.mask_off != NONE && !masks[jb.mask_off + i + (size_t)j * jb.dw]) g = 0;
		}
		G[jb.img_off + i + (size_t)j * jb.dw] = (TI)g;
	}
	block_sum_256<1>(c2);
	if (threadIdx.x == 0 && c2[0] != 0) atomicAdd(&chi2[jn], c2[0]);
}

// Adjoint weights of the refinement entries (FP32 mode): gent[2 e + c] =
// G_c or G_nc at entry e's pixel, divided by its job's fixed-point scale
// (the refinement's values carry that scale).
template <typename T>
__global__ __launch_bounds__(256) void grad_entry_weights_kernel(const PkJob *__restrict__ jobs,
                                                                 const RefineEntry *__restrict__ entries,
                                                                 const PkCounters *__restrict__ cnt,
                                                                 const T *__restrict__ G_nc,
                                                                 const T *__restrict__ G_c,
                                                                 double *__restrict__ gent)
{
	const uint32_t n = cnt->n_entries;
	for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
		const RefineEntry en = entries[e];
		const PkJob &jb = jobs[en.job];
		const size_t p = jb.img_off + en.pix;
		const double inv = 1.0 / jb.fx_scale;
		gent[2 * (size_t)e] = (double)G_nc[p] * inv;
		gent[2 * (size_t)e + 1] = jb.has_conv ? (double)G_c[p] * inv : 0.0;
	}
}

// Adjoint of the coarse pass: every pixel a component evaluates directly
// (src/radial.cpp:316-324) contributes G * flux_scale * f * dL/dθ; the
// pixels refined by the forward pass are left to the refinement adjoint.
template <typename TI, bool FP32>
__global__ __launch_bounds__(256) void coarse_grad_kernel(const PkJob *__restrict__ jobs, uint32_t njobs,
                                                          const PkComp *__restrict__ comps,
                                                          const uint8_t *__restrict__ masks,
                                                          const TI *__restrict__ G_nc, const TI *__restrict__ G_c,
                                                          double *__restrict__ gacc,
                                                          const uint32_t *__restrict__ tile_job)
{
	const uint32_t jn = find_job(jobs, njobs, blockIdx.x, 0, tile_job);
	const PkJob &jb = jobs[jn];
	const uint32_t t = blockIdx.x - jb.tile_off;
	const uint32_t i = (t % jb.tiles_x) * 64 + (threadIdx.x & 63);
	const uint32_t j0 = (t / jb.tiles_x) * PX_ROWS + (threadIdx.x >> 6) * PX_RPT;
	bool live[PX_RPT];
	double gn[PX_RPT], gc[PX_RPT];
#pragma unroll
	for (int r = 0; r < PX_RPT; r++) {
		const uint32_t j = j0 + r;
		live[r] = i < jb.dw && j < jb.dh;
		gn[r] = gc[r] = 0;
		if (live[r]) {
			const size_t pix = jb.img_off + j * jb.dw + i;
			if (jb.mask_off != NONE && !masks[jb.mask_off + j * jb.dw + i]) live[r] = false;
			gn[r] = (double)G_nc[pix];
			if (jb.has_conv) gc[r] = (double)G_c[pix];
		}
	}
	const double x = jb.sx / 2. + i * jb.sx;
	for (uint32_t c = 0; c < jb.ncomp; c++) {
		const PkComp &cp = comps[jb.comp_off + c];
		double acc[GA_SLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
		for (int r = 0; r < PX_RPT; r++) {
			if (!live[r]) continue;
			const double G = cp.convolve ? gc[r] : gn[r];
			if (cp.kind == PK_PROFILE_SKY) {
				acc[GA_W] += G;
			} else if (cp.kind == PK_PROFILE_SERSIC && G != 0) {
				const double y = jb.sy / 2. + (j0 + r) * jb.sy;
				double xp, yp;
				to_prof(cp, x, y, xp, yp);
				const double r2 = xp * xp + yp * yp;
				const bool zero = cp.rscale_max > 0 && radius_exceeds(r2, cp.r2_max, cp.rscale, cp.rscale_max);
				if (!zero && (cp.rough || radius_exceeds(r2, cp.r2_switch, cp.rscale, cp.rscale_switch))) {
					if (FP32) {
						float g[GT_COUNT];
						const float v = sersic_grad32(cp, (float)xp, (float)yp, g);
						grad_accumulate(acc, G * cp.flux_scale * cp.unfold * (double)v, g);
					} else {
						double g[GT_COUNT];
						const double v = sersic_grad64(cp, xp, yp, g);
						grad_accumulate(acc, G * cp.flux_scale * v, g);
					}
				}
			}
		}
		if (cp.kind == PK_PROFILE_SKY) {
			// the sky's one slot
			block_sum_256<1>(&acc[GA_W]);
			if (threadIdx.x == 0 && acc[GA_W] != 0)
				atomicAdd(&gacc[(size_t)(jb.comp_off + c) * GA_SLOTS + GA_W], acc[GA_W]);
			continue;
		}
		block_sum_256<GA_SLOTS>(acc);
		if (threadIdx.x == 0)
#pragma unroll
			for (int k = 0; k < GA_SLOTS; k++)
				if (acc[k] != 0) atomicAdd(&gacc[(size_t)(jb.comp_off + c) * GA_SLOTS + k], acc[k]);
	}
}

// Raw accumulators -> gradient slots of include/profit_hip.h (PK_GRAD_*).
__global__ void grad_finalize_kernel(const PkComp *__restrict__ comps, const double *__restrict__ gacc,
                                     double *__restrict__ gout, uint32_t n)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	const PkComp &c = comps[t];
	const double *a = gacc + (size_t)t * GA_SLOTS;
	double *o = gout + (size_t)t * PK_GRAD_SLOTS;
	for (int k = 0; k < PK_GRAD_SLOTS; k++) o[k] = 0;
	if (c.kind == PK_PROFILE_SKY) {
		o[PK_GRAD_SKY_BG] = a[GA_W] * c.q[1];
	} else if (c.kind == PK_PROFILE_SERSIC) {
		const double w = a[GA_W];
		o[PK_GRAD_XCEN] = a[GA_XC];
		o[PK_GRAD_YCEN] = a[GA_YC];
		o[PK_GRAD_MAG] = -0.4 * 2.302585092994045684 * w;  // d log(10^(-0.4 mag)) / d mag
		o[PK_GRAD_RE] = a[GA_RE] + c.glnf_re * w;
		o[PK_GRAD_NSER] = a[GA_N] + c.glnf_n * w;
		o[PK_GRAD_ANG] = a[GA_PHI] * (PK_PI / 180.);  // the angle is in degrees (src/radial.cpp:191-196)
		o[PK_GRAD_AXRAT] = a[GA_Q] + c.glnf_q * w;
		o[PK_GRAD_BOX] = a[GA_B] + c.glnf_b * w;
	}
}

// Clears the adjoint accumulators and the refinement-adjoint counters (a
// kernel rather than memset nodes, so that the sequence can be captured in a
// graph); the refinement entries of the forward pass are reused.
__global__ void reset_grad_kernel(PkCounters *__restrict__ g, const PkCounters *__restrict__ fwd,
                                  double *__restrict__ gacc, uint32_t nacc, double *__restrict__ chi2, uint32_t nj)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	for (uint32_t k = t; k < nacc; k += gridDim.x * blockDim.x) gacc[k] = 0;
	for (uint32_t k = t; k < nj; k += gridDim.x * blockDim.x) chi2[k] = 0;
	if (t == 0) {
		g->n_entries = fwd->n_entries;
		g->next_entry = 0;
		g->overflow = 0;
		g->n_bad = 0;
		g->n_subcells = 0;
		for (int p = 0; p < R32_PASSES; p++) {
			g->claim[p] = 0;
			g->n_cells[p] = 0;
			g->n_pend[p] = 0;
		}
	}
}

}  // namespace pk

#include "pk_refine32.h"

namespace pk {

// counters of one evaluation are cleared by a kernel (not a memset node) so
// that the launch sequence can be captured into a graph and replayed
__global__ void reset_counters_kernel(PkCounters *c)
{
	if (threadIdx.x == 0) {
		c->n_entries = 0;
		c->next_entry = 0;
		c->overflow = 0;
		c->n_bad = 0;
		c->n_subcells = 0;
		for (int p = 0; p < R32_PASSES; p++) {
			c->claim[p] = 0;
			c->n_cells[p] = 0;
			c->n_pend[p] = 0;
		}
	}
}

// Refined pixels per component (bits of the entries' component masks):
// LDS histogram per workgroup, one atomic per component and workgroup
__global__ __launch_bounds__(256) void integrations_kernel(const RefineEntry *__restrict__ entries,
                                                            const PkCounters *__restrict__ cnt,
                                                            unsigned long long *__restrict__ hist)
{
	__shared__ unsigned int h[32];
	if (threadIdx.x < 32) h[threadIdx.x] = 0;
	__syncthreads();
	const uint32_t n = cnt->n_entries;
	for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
		uint32_t m = entries[e].cmask;
		while (m) {
			const int c = __ffs(m) - 1;
			m &= m - 1;
			atomicAdd(&h[c], 1u);
		}
	}
	__syncthreads();
	if (threadIdx.x < 32 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// Sum of the per-wavefront counts of a COUNT pass into *out (cleared
// before): grid-stride, one atomic per wavefront
__global__ __launch_bounds__(256) void sum_wave_evals_kernel(const uint32_t *__restrict__ w, uint64_t n,
                                                              unsigned long long *__restrict__ out)
{
	unsigned long long v = 0;
	for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
		v += w[k];
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
	if (lane_id() == 0 && v) atomicAdd(out, v);
}

// ---------------------------------------------------------------------------
// set-up of all components on the device (batched path)
// ---------------------------------------------------------------------------
// The lanes of a wavefront as the quadrature's Lanes policy (pk_math.h): the
// butterfly leaves rounding-level differences between lanes, so the total
// is taken from lane 0 for all.
struct WaveLanes {
	QagWork *ws;  // the wavefront's QUADPACK interval list (LDS)
	__device__ QagWork *work() const { return ws; }
	__device__ int lane() const { return (int)__lane_id(); }
	__device__ int count() const { return 64; }
	__device__ double pick(double v, int l) const { return __shfl(v, l); }
	__device__ double sum(double v) const
	{
		for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
		return __shfl(v, 0);
	}
};

// One wavefront per component: every lane runs the (sequential) set-up, the
// luminosity quadratures spread their nodes over the lanes (the QUADPACK
// interval list in LDS, written alike by every lane), lane 0 writes.
__global__ __launch_bounds__(64) void setup_kernel(const SetupIn *__restrict__ in, PkComp *__restrict__ out,
                                                   CompLite *__restrict__ lite, uint32_t n)
{
	const uint32_t t = blockIdx.x;
	if (t >= n) return;
	__shared__ QagWork qws;
	PkComp c;
	setup_component(in[t], c, WaveLanes{&qws});
	if (__lane_id() == 0) {
		out[t] = c;
		if (lite) lite[t] = make_lite(c, in[t]);
	}
}

#endif  // PK_STRIP_KERNELS_ONLY

}  // namespace pk
