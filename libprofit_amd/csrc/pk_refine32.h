// FP32-mode sub-pixel refinement (included by pk_device.h).
//
// Work items are cells (RCell): a pixel of one component integrated from
// level 0, or a failing sub-cell at a deeper level.  Each wavefront keeps a
// LIFO of cells in LDS.  A round takes cells from its top until their
// sub-cells (res^2 each) fill the 64 lanes -- a cell with more than 64
// sub-cells is spread over several rounds -- and every lane evaluates one
// sub-cell and its accuracy test (src/radial.cpp:86-109).  Failing sub-cells
// go back onto the LIFO, or onto a global spill queue worked off by a second
// launch when the LIFO is full.  Accepted values of a cell are summed as exact
// 64-bit fixed-point integers and added atomically to the pixel's
// accumulator: neither the packing of cells into rounds nor the scheduling of
// wavefronts can change the result.
//
// Latency: a round reads one 128-byte CompLite record per cell (one cache
// line); refills claim WQ_CHUNK items with one atomic and load them with one
// coalesced access, expanding the pixels of a chunk lane-parallel.
#pragma once

namespace pk {

// LIFO capacity: small in the main pass (occupancy), large in the pass over
// spilled cells, which has nowhere further to spill.  A cell of resolution
// res can push res^2 children, so depth-first processing needs at most
// about (res^2 - 1) * levels + res^2 cells: 1024 covers res 16 with 3 levels
// and the Sersic centre pixels (res 8, 10 levels).
constexpr int R32_CAP_MAIN = 192;
constexpr int R32_CAP_SPILL = 1024;
constexpr uint32_t WQ_CHUNK = 64;

constexpr int R32_TASKS = 16;  // cells per round at most (their records are staged in LDS)

template <int CAP>
struct R32Stack {
	CompLite lite[R32_TASKS];
	uint32_t gcomp[CAP], entry[CAP], meta[CAP];
	double dx0[CAP], dy0[CAP], f[CAP];
};

__device__ inline uint32_t wave_sum_u32(uint32_t v)
{
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
	return v;
}

__device__ inline uint32_t wave_incl_scan_u32(uint32_t v)
{
	const int lane = (int)lane_id();
	for (int off = 1; off < 64; off <<= 1) {
		const uint32_t t = __shfl_up(v, off);
		if (lane >= off) v += t;
	}
	return v;
}

// cell metadata: level | res << 8 | max recursions << 16 | convolved << 24
__device__ inline uint32_t cm_level(uint32_t m) { return m & 255u; }
__device__ inline uint32_t cm_res(uint32_t m) { return (m >> 8) & 255u; }
__device__ inline uint32_t cm_maxrec(uint32_t m) { return (m >> 16) & 255u; }
__device__ inline uint32_t cm_conv(uint32_t m) { return (m >> 24) & 1u; }

// FP64 value and accuracy decision of a sub-cell whose FP32 decision fell
// within the guard band of the threshold (kept out of line: it is rare, and
// inlined it would raise the register count of the whole kernel).
__device__ __attribute__((noinline)) double refine_fp64_decision(const PkComp &cp, double xp, double yp, double ddy,
                                                                 bool recurse, bool *fail)
{
	const double v = evaluate_at(cp, xp, yp);
	*fail = false;
	if (recurse) {
		const double tv = evaluate_at(cp, fabs(xp), fabs(yp) + fabs(ddy));
		*fail = fabs(tv / v - 1.0) > cp.acc;
	}
	return v;
}

template <int SOURCE>  // 0: pixels appended by the coarse pass, 1: the spill queue
__global__ __launch_bounds__(64) void refine32_kernel(const PkJob *__restrict__ jobs, const PkComp *__restrict__ comps,
                                                      const CompLite *__restrict__ lite,
                                                      const RefineEntry *__restrict__ entries,
                                                      PkCounters *__restrict__ cnt, RCell *__restrict__ spill,
                                                      uint32_t spill_cap, unsigned long long *__restrict__ acc)
{
	constexpr int R32_CAP = SOURCE == 0 ? R32_CAP_MAIN : R32_CAP_SPILL;
	__shared__ R32Stack<R32_CAP> st;
	const int lane = (int)lane_id();
	int top = 0;
	bool source_done = false;
	const uint32_t n_source = SOURCE == 0 ? cnt->n_entries : min(cnt->n_spill, spill_cap);
	uint32_t overflow = 0, bad = 0;
	unsigned long long nev = 0;
	uint32_t chunk_next = 0, chunk_end = 0;
	for (;;) {
		// ---- refill while the cells on the LIFO cannot fill the lanes ----
		for (;;) {
			uint32_t r2 = 0;
			if (lane < top) {
				const uint32_t res = cm_res(st.meta[top - 1 - lane]);
				r2 = res * res;
			}
			if (wave_sum_u32(r2) >= 64u || source_done) break;
			if (R32_CAP - top < (SOURCE == 0 ? 32 : 1)) break;
			if (chunk_next >= chunk_end) {
				uint32_t base = 0;
				if (lane == 0) base = atomicAdd(SOURCE == 0 ? &cnt->next_entry : &cnt->next_spill, WQ_CHUNK);
				base = __shfl(base, 0);
				if (base >= n_source) {
					source_done = true;
					break;
				}
				chunk_next = base;
				chunk_end = min(base + WQ_CHUNK, n_source);
			}
			const uint32_t avail = chunk_end - chunk_next;
			const uint32_t space = (uint32_t)(R32_CAP - top);
			if (SOURCE == 0) {
				// lane L expands pixel chunk_next + L into one cell per refined component
				const bool valid = (uint32_t)lane < avail;
				RefineEntry en{0, 0, 0};
				if (valid) en = entries[chunk_next + lane];
				const uint32_t nb = valid ? (uint32_t)__popc(en.cmask) : 0u;
				const uint32_t incl = wave_incl_scan_u32(nb);
				const uint64_t ok = __ballot(valid && incl <= space);
				const int take = __popcll(ok);  // >= 1: space >= 32 >= cells of one pixel
				if ((uint64_t)lane < (uint64_t)take) {
					const PkJob &jb = jobs[en.job];
					const uint32_t i = en.pix % jb.dw, j = en.pix / jb.dw;
					const double hx = jb.sx / 2., hy = jb.sy / 2.;
					const double x = hx + i * jb.sx, y = hy + j * jb.sy;  // pixel centre (src/radial.cpp:313-314)
					int pos = top + (int)(incl - nb);
					uint32_t m = en.cmask;
					while (m) {
						const int c = __ffs(m) - 1;
						m &= m - 1;
						const uint32_t gc = jb.comp_off + (uint32_t)c;
						const PkComp &cp = comps[gc];
						uint32_t res = (uint32_t)cp.res, mr = (uint32_t)cp.max_rec;
						if (cp.center_res && fabs(x - cp.xcen) < cp.win_x && fabs(y - cp.ycen) < cp.win_y) {
							// SersicProfile::subsampling_params (src/sersic.cpp:331-347)
							res = (uint32_t)cp.center_res;
							mr = (uint32_t)cp.center_max_rec;
						}
						st.gcomp[pos] = gc;
						st.entry[pos] = chunk_next + (uint32_t)lane;
						st.meta[pos] = (res << 8) | (mr << 16) | ((cp.convolve ? 1u : 0u) << 24);
						st.dx0[pos] = (x - hx) - cp.xcen;
						st.dy0[pos] = (y - hy) - cp.ycen;
						st.f[pos] = 1.0;
						pos++;
					}
				}
				top += (int)__shfl(incl, take - 1);
				chunk_next += (uint32_t)take;
			} else {
				const uint32_t n = min(min(avail, space), 64u);
				if ((uint32_t)lane < n) {
					const RCell c = spill[chunk_next + lane];
					const int pos = top + lane;
					st.gcomp[pos] = c.gcomp;
					st.entry[pos] = c.entry;
					st.meta[pos] = c.meta;
					st.dx0[pos] = c.dx0;
					st.dy0[pos] = c.dy0;
					st.f[pos] = c.f;
				}
				top += (int)n;
				chunk_next += n;
			}
			__syncthreads();
		}
		if (top == 0) break;

		// ---- cells on top of the LIFO -> lanes ----
		uint32_t r2 = 0;
		if (lane < top) {
			const uint32_t res = cm_res(st.meta[top - 1 - lane]);
			r2 = res * res;
		}
		const uint32_t incl = wave_incl_scan_u32(r2);
		const uint64_t fit = __ballot(lane < top && lane < R32_TASKS && incl <= 64u);
		const int ntask = fit ? __popcll(fit) : 1;
		const uint32_t lanes_needed = __shfl(incl, ntask - 1);
		const int rounds = (int)((lanes_needed + 63u) / 64u);
		// task of this lane (rounds > 1 only with a single task)
		int t_local = 0;
		for (int t = 0; t < ntask; t++) t_local += (__shfl(incl, t) <= (uint32_t)lane) ? 1 : 0;
		if (rounds > 1 || t_local >= ntask) t_local = rounds > 1 ? 0 : ntask - 1;
		// every lane takes part in the shuffles (a lane that is inactive in a
		// branch would be read as 0)
		const uint32_t prev = __shfl(incl, t_local > 0 ? t_local - 1 : 0);
		const uint32_t seg_end1 = __shfl(incl, t_local);
		const uint32_t before = t_local > 0 ? prev : 0u;  // first lane of the task (single round)
		const int slot = top - 1 - t_local;
		const uint32_t gcomp = st.gcomp[slot], entry = st.entry[slot], meta = st.meta[slot];
		const double dx0 = st.dx0[slot], dy0 = st.dy0[slot], f = st.f[slot];
		// lane t < ntask stages the component record of task t in LDS
		if (lane < ntask) st.lite[lane] = lite[st.gcomp[top - 1 - lane]];
		__syncthreads();
		top -= ntask;

		const CompLite &cl = st.lite[t_local];
		const int level = (int)cm_level(meta), res = (int)cm_res(meta), mr = (int)cm_maxrec(meta);
		const double fr = f / res;
		const double hx = 0.5 * fr * cl.sx, hy = 0.5 * fr * cl.sy;
		const double wflux = fr * fr * cl.wflux;
		const bool recurse = res > 1 && level < mr;
		const double acc_thr = cl.acc;
		const double ddy = (-2 * hx * cl.sin_ang + 2 * hy * cl.cos_ang) * cl.inv_axrat;
		const size_t target = 2 * (size_t)entry + cm_conv(meta);
		for (int rd = 0; rd < rounds; rd++) {
			const uint32_t g = (uint32_t)lane + 64u * (uint32_t)rd;
			const bool active = g < lanes_needed;
			const uint32_t s = g - before;
			bool fail = false;
			double dx = 0, dy = 0;
			unsigned long long fxv = 0;
			if (active) {
				const int si = (int)s / res, sj = (int)s % res;
				dx = dx0 + (2 * si + 1) * hx;
				dy = dy0 + (2 * sj + 1) * hy;
				const double xp = dx * cl.cos_ang + dy * cl.sin_ang;
				const double yp = (-dx * cl.sin_ang + dy * cl.cos_ang) * cl.inv_axrat;
				double val = 0;
				bool decided = false;
				if (cl.fp32) {
					const float v32 = evaluate_at32(cl, (float)xp, (float)yp, radius2_64(cl, xp, yp));
					nev += 1;
					if (recurse) {
						const double txp = fabs(xp), typ = fabs(yp) + fabs(ddy);
						const float t32 = evaluate_at32(cl, (float)txp, (float)typ, radius2_64(cl, txp, typ));
						nev += 1;
						const float m = fabsf(t32 / v32 - 1.f);
						if (v32 > 0.f && isfinite(m) && fabs((double)m - acc_thr) > 1e-4 * (1.0 + acc_thr)) {
							fail = (double)m > acc_thr;
							val = (double)v32 * cl.unfold;
							decided = true;
						}
					} else {
						val = (double)v32 * cl.unfold;
						decided = true;
					}
				}
				if (!decided) {
					val = refine_fp64_decision(comps[gcomp], xp, yp, ddy, recurse, &fail);
					nev += recurse ? 2 : 1;
				}
				if (!fail) {
					const double q = val * wflux;
					if (q >= 0 && q < 9.2e18) fxv = (unsigned long long)__double2ll_rn(q);
					else bad++;
				}
			}
			// exact per-cell sums: segmented scan over the cell's contiguous lanes,
			// the last lane adds the sum to the pixel accumulator
			const uint32_t sbeg = rounds > 1 ? 0u : before;
			const uint32_t send = rounds > 1 ? min(64u, lanes_needed - 64u * (uint32_t)rd) : seg_end1;
			unsigned long long v = fxv;
			for (int off = 1; off < 64; off <<= 1) {
				const unsigned long long t = __shfl_up(v, (unsigned int)off);
				if ((uint32_t)lane >= sbeg + (uint32_t)off) v += t;
			}
			if (active && (uint32_t)lane == send - 1 && v) atomicAdd(&acc[target], v);
			// failing sub-cells: onto the LIFO, or the spill queue when it is full
			const uint64_t ballot = __ballot(fail);
			if (ballot) {
				const int n = __popcll(ballot);
				const int rank = __popcll(ballot & ((1ull << lane) - 1ull));
				const double cdx0 = dx - hx, cdy0 = dy - hy;
				const uint32_t cmeta = (meta & ~255u) | (uint32_t)(level + 1);
				if (top + n <= R32_CAP) {
					if (fail) {
						const int pos = top + rank;
						st.gcomp[pos] = gcomp;
						st.entry[pos] = entry;
						st.meta[pos] = cmeta;
						st.dx0[pos] = cdx0;
						st.dy0[pos] = cdy0;
						st.f[pos] = fr;
					}
					top += n;
				} else if (SOURCE == 0) {
					const int leader = __ffsll((long long)ballot) - 1;
					uint32_t base = 0;
					if (lane == leader) base = atomicAdd(&cnt->n_spill, (uint32_t)n);
					base = __shfl(base, leader);
					if (fail) {
						if (base + rank < spill_cap) {
							RCell c;
							c.gcomp = gcomp;
							c.entry = entry;
							c.meta = cmeta;
							c.pad_ = 0;
							c.dx0 = cdx0;
							c.dy0 = cdy0;
							c.f = fr;
							spill[base + rank] = c;
						} else {
							overflow = 1;
						}
					}
				} else {
					overflow = 1;
				}
			}
			__syncthreads();
		}
	}
	for (int off = 32; off > 0; off >>= 1) nev += __shfl_xor(nev, off);
	const uint64_t ov = __ballot(overflow != 0);
	const uint32_t nbad = wave_sum_u32(bad);
	if (lane == 0) {
		atomicAdd(&cnt->n_subcells, nev);
		if (ov) atomicOr(&cnt->overflow, 1u);
		if (nbad) atomicAdd(&cnt->n_bad, nbad);
	}
}

// adds the refined contributions to the pixels (one writer per pixel)
template <typename TI>
__global__ __launch_bounds__(256) void finalize32_kernel(const PkJob *__restrict__ jobs,
                                                         const RefineEntry *__restrict__ entries,
                                                         const PkCounters *__restrict__ cnt,
                                                         const unsigned long long *__restrict__ acc,
                                                         TI *__restrict__ img_nc, TI *__restrict__ img_c)
{
	const uint32_t n = cnt->n_entries;
	for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
		const RefineEntry en = entries[e];
		const PkJob &jb = jobs[en.job];
		const double inv = 1.0 / jb.fx_scale;
		const unsigned long long a0 = acc[2 * (size_t)e], a1 = acc[2 * (size_t)e + 1];
		if (a0) img_nc[jb.img_off + en.pix] += (TI)((double)a0 * inv);
		if (a1) img_c[jb.img_off + en.pix] += (TI)((double)a1 * inv);
	}
}

}  // namespace pk
