// FP32-mode sub-pixel refinement (included by pk_device.h).
//
// Work items are cells (RCell): a pixel of one component integrated from
// level 0, or a failing sub-cell at a deeper level.  Each wavefront keeps a
// LIFO of cells in LDS.  A round takes cells from its top until their
// sub-cells (res^2 each) fill the 64 lanes -- a cell with more than 64
// sub-cells is spread over several rounds -- and every lane evaluates one
// sub-cell and its accuracy test (src/radial.cpp:86-109).  Failing sub-cells
// go back onto the LIFO.  Accepted values of a cell are summed as exact 64-bit
// fixed-point integers and added atomically to the pixel's accumulator:
// neither the packing of cells into rounds nor the scheduling of wavefronts
// can change the result.
//
// Load balance: the refinement runs as R32_PASSES launches.  Pass 0 takes
// pixels from the coarse pass's entry list, pass p > 0 the cells handed over
// by pass p - 1.  Work is claimed in chunks that shrink as the source runs
// dry.  Once a set share of a pass's wavefronts (R32Tune::handover_div; a
// quarter by default) have stopped, the others hand everything they still
// hold (their LIFO and unclaimed chunk) to the next pass and stop too, so a few deep pixels never keep a lone wavefront busy
// while the rest of the GPU idles.  The last pass finishes all that is left,
// with a deeper LIFO.  A LIFO that fills up before that also hands its new
// cells to the next pass.
//
// Latency: lane-crossing scans and whole-wavefront sums use DPP; a round
// holding several cells sums each cell's sub-cells with LDS integer atomics
// (one slot per cell) and adds the slots to the pixels.  The 128-byte CompLite
// record of a cell is staged in LDS and kept there while the following rounds
// work on the same component.
#pragma once

namespace pk {

// LIFO capacity: small in the handing-over passes (occupancy), large in the
// last pass, which has nowhere further to go.  Depth-first processing holds at
// most about res^2 * (max recursions) cells: 1024 covers res 16 with 3 levels
// and the Sersic centre pixels (res 8, 10 levels).
constexpr int R32_CAP_MAIN = 192;
constexpr int R32_CAP_LAST = 1024;
constexpr uint32_t R32_CELL_RUN = 4;       // log2 of the runs in which handed-over cells are claimed
constexpr int R32_TASKS = 16;              // cells per round at most (their records are staged in LDS)
// untuned settings: claims of up to 64 pixels / 16 cells, the hand-over test
// every 16 rounds, hand-over once a quarter of the wavefronts stopped
constexpr R32Tune R32_TUNE_DEFAULT = {4u, 64u, 16u, 16u, 0u};

template <int CAP>
struct R32Stack {
	CompLite lite[R32_TASKS];
	uint32_t lite_g[R32_TASKS];   // component whose record lite[t] holds
	unsigned long long csum[R32_TASKS];  // fixed-point sum of task t in the current round (0 between rounds)
	RCell cell[CAP];
};

// ---- wavefront primitives on DPP ----------------------------------------------
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118;
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

// v of the DPP source lane; 0 where the source lies outside the row, or the
// lane's row is outside ROW_MASK.  Call with all 64 lanes active.
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline uint32_t dpp0(uint32_t v)
{
	return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ inline unsigned long long dpp0_u64(unsigned long long v)
{
	const uint32_t lo = dpp0<CTRL, ROW_MASK>((uint32_t)v), hi = dpp0<CTRL, ROW_MASK>((uint32_t)(v >> 32));
	return ((unsigned long long)hi << 32) | lo;
}

// inclusive prefix sum over the 64 lanes
__device__ inline uint32_t wave_incl_scan_u32(uint32_t v)
{
	v += dpp0<DPP_ROW_SHR1>(v);
	v += dpp0<DPP_ROW_SHR2>(v);
	v += dpp0<DPP_ROW_SHR4>(v);
	v += dpp0<DPP_ROW_SHR8>(v);
	v += dpp0<DPP_ROW_BCAST15, 0xa>(v);  // row 0 -> row 1, row 2 -> row 3
	v += dpp0<DPP_ROW_BCAST31, 0xc>(v);  // lane 31 -> rows 2, 3
	return v;
}

__device__ inline uint32_t wave_read(uint32_t v, int lane) { return (uint32_t)__builtin_amdgcn_readlane((int)v, lane); }

__device__ inline uint32_t wave_sum_u32(uint32_t v) { return wave_read(wave_incl_scan_u32(v), 63); }

// sum over the 64 lanes (lane 63 holds it after the scan; read as a scalar)
__device__ inline unsigned long long wave_sum_u64(unsigned long long v)
{
	v += dpp0_u64<DPP_ROW_SHR1>(v);
	v += dpp0_u64<DPP_ROW_SHR2>(v);
	v += dpp0_u64<DPP_ROW_SHR4>(v);
	v += dpp0_u64<DPP_ROW_SHR8>(v);
	v += dpp0_u64<DPP_ROW_BCAST15, 0xa>(v);
	v += dpp0_u64<DPP_ROW_BCAST31, 0xc>(v);
	const uint32_t lo = wave_read((uint32_t)v, 63), hi = wave_read((uint32_t)(v >> 32), 63);
	return ((unsigned long long)hi << 32) | lo;
}

// Size in pixels of a cell of level `level`, res^-level in FP64 (res^level is
// an exact integer: one rounding, where libprofit's level-by-level
// (x1 - x0) / resolution takes one per level)
__device__ inline double cell_size_fp64(int res, int level)
{
	double p = 1.0;
	for (int l = 0; l < level; l++) p *= (double)res;
	return 1.0 / p;
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

// Entry of pass-0 claim index i: with spreading (L = spread_log2 > 0) the
// entry list is cut into runs of 2^L entries, and consecutive runs of claim
// indices come from 64 places of the list in turn (run b -> (b % 64) * sstride
// + b / 64), so that the pixels of one deep galaxy centre go to many
// wavefronts.  Returns an index >= n for claim indices past the list.
__device__ inline uint32_t r32_entry_of(uint32_t i, uint32_t spread_log2, uint32_t sstride)
{
	if (!spread_log2) return i;
	const uint32_t b = i >> spread_log2, w = i & ((1u << spread_log2) - 1u);
	return ((((b & 63u) * sstride) + (b >> 6)) << spread_log2) | w;
}

// Expands the refinement entries of claim indices [next, end) into level-0
// cells on the LIFO while they fit, one pixel per lane (one cell per refined
// component).  Needs at least 32 free slots, so that one pixel always fits.
__device__ inline void r32_expand(const PkJob *__restrict__ jobs, const PkComp *__restrict__ comps,
                                  const RefineEntry *__restrict__ entries, uint32_t n_entries, uint32_t spread_log2,
                                  uint32_t sstride, RCell *cells, int cap, int &top, uint32_t &next, uint32_t end,
                                  uint32_t lane)
{
	const uint32_t avail = end - next;
	const uint32_t space = (uint32_t)(cap - top);
	const uint32_t e = r32_entry_of(next + lane, spread_log2, sstride);
	const bool valid = lane < avail && e < n_entries;
	RefineEntry en{0, 0, 0};
	if (valid) en = entries[e];
	const uint32_t nb = valid ? (uint32_t)__popc(en.cmask) : 0u;
	const uint32_t incl = wave_incl_scan_u32(nb);
	// claim indices past the list hold no entry (nb = 0): the lanes taken are
	// a prefix of the available ones
	const int take = __popcll(__ballot(lane < avail && incl <= space));
	// cells in component-major order -- every lane's cell of its first
	// refined component, then of its second, ... -- so that runs of cells
	// on the LIFO share a component (and a resolution)
	const bool mine = (int)lane < take && valid;
	double x = 0, y = 0, hx = 0, hy = 0;
	uint32_t comp_off = 0;
	if (mine) {
		const PkJob &jb = jobs[en.job];
		const uint32_t i = en.pix % jb.dw, j = en.pix / jb.dw;
		hx = jb.sx / 2.;
		hy = jb.sy / 2.;
		x = hx + i * jb.sx;  // pixel centre (src/radial.cpp:313-314)
		y = hy + j * jb.sy;
		comp_off = jb.comp_off;
	}
	uint32_t m = mine ? en.cmask : 0u;
	int pos = top;
	for (;;) {
		const uint64_t b = __ballot(m != 0);
		if (!b) break;
		if (m) {
			const int c = __ffs(m) - 1;
			m &= m - 1;
			const uint32_t gc = comp_off + (uint32_t)c;
			const PkComp &cp = comps[gc];
			uint32_t res = (uint32_t)cp.res, mr = (uint32_t)cp.max_rec;
			if (cp.center_res && fabs(x - cp.xcen) < cp.win_x && fabs(y - cp.ycen) < cp.win_y) {
				// SersicProfile::subsampling_params (src/sersic.cpp:331-347)
				res = (uint32_t)cp.center_res;
				mr = (uint32_t)cp.center_max_rec;
			}
			RCell cl;
			cl.dx0 = (x - hx) - cp.xcen;
			cl.dy0 = (y - hy) - cp.ycen;
			cl.gcomp = gc;
			cl.entry = e;
			cl.meta = (res << 8) | (mr << 16) | ((cp.convolve ? 1u : 0u) << 24);
			cells[pos + __popcll(b & ((1ull << lane) - 1ull))] = cl;
		}
		pos += __popcll(b);
	}
	top = pos;
	next += (uint32_t)take;
}

// Claims the next chunk of a work queue of n items; the chunk shrinks as the
// queue runs dry (`seen` is the head this wavefront saw last).
__device__ inline bool r32_claim(uint32_t *head, uint32_t n, uint32_t chunk, uint32_t &seen, uint32_t &next,
                                 uint32_t &end, uint32_t lane)
{
	const uint32_t left = n > seen ? n - seen : 0u;
	const uint32_t want = max(1u, min(chunk, left / (2u * gridDim.x)));
	uint32_t base = 0;
	if (lane == 0) base = atomicAdd(head, want);
	base = wave_read(base, 0);
	seen = base + want;
	if (base >= n) return false;
	next = base;
	end = min(base + want, n);
	return true;
}

// Copies handed-over cells of claim indices [next, end) onto the LIFO, as many
// as fit (at most 64); claim index i holds cell r32_entry_of(i, R32_CELL_RUN,
// stride) if that is below n.
__device__ inline void r32_take_cells(const RCell *__restrict__ qin, uint32_t n, uint32_t stride, RCell *cells,
                                      int cap, int &top, uint32_t &next, uint32_t end, uint32_t lane)
{
	const uint32_t k = min(min(end - next, (uint32_t)(cap - top)), 64u);
	const uint32_t i = next + lane;
	const uint32_t c = r32_entry_of(i, R32_CELL_RUN, stride);
	const bool valid = lane < k && c < n;
	const uint64_t b = __ballot(valid);
	if (valid) cells[top + __popcll(b & ((1ull << lane) - 1ull))] = qin[c];
	top += __popcll(b);
	next += k;
}

// Adds the adjoint terms of this round's accepted sub-cells (lanes with `ok`)
// to gacc[component]: one wavefront sum per distinct component in the round
// (usually one).
__device__ inline void r32_flush_grad(bool ok, uint32_t gcomp, double ga[GA_SLOTS], double *__restrict__ gacc,
                                      uint32_t lane)
{
	uint64_t todo = __ballot(ok);
	while (todo) {
		const int l0 = __ffsll((long long)todo) - 1;
		const uint32_t g0 = wave_read(gcomp, l0);
		const bool mine = ok && gcomp == g0;
		todo &= ~__ballot(mine);
#pragma unroll
		for (int k = 0; k < GA_SLOTS; k++) {
			const double v = wave_sum(mine ? ga[k] : 0.0);
			if (lane == 0 && v != 0) atomicAdd(&gacc[(size_t)g0 * GA_SLOTS + k], v);
		}
	}
}

// Adds the per-lane running fixed-point sums of pixel accumulator t
// (wavefront-uniform) to acc and clears them; t = ~0: nothing held.
__device__ inline void r32_flush_fx(unsigned long long &run, uint64_t t, unsigned long long *__restrict__ acc,
                                    uint32_t lane)
{
	if (t == ~0ull) return;
	const unsigned long long v = wave_sum_u64(run);
	if (lane == 0 && v) atomicAdd(&acc[t], v);
	run = 0;
}

// Adds the per-lane running adjoint sums of component g (wavefront-uniform)
// to gacc and clears them; g = ~0u: nothing held.
__device__ inline void r32_flush_run(double run[GA_SLOTS], uint32_t g, double *__restrict__ gacc, uint32_t lane)
{
	if (g == ~0u) return;
#pragma unroll
	for (int k = 0; k < GA_SLOTS; k++) {
		const double v = wave_sum(run[k]);
		if (lane == 0 && v != 0) atomicAdd(&gacc[(size_t)g * GA_SLOTS + k], v);
		run[k] = 0;
	}
}

// SOURCE 0: pixels from the coarse pass (pass 0); 1: cells of the previous pass.
// LAST: the final pass, which hands nothing over.
// GRAD: the chi^2 adjoint of the refinement -- the same cells and decisions as
// the forward pass, but every accepted sub-cell adds G w flux_scale f dL/dθ
// to gacc instead of its value to the pixel.  gent[2 e + convolved] holds
// G / fx_scale of refinement entry e (G: the adjoint image at its pixel).
// Rounds whose accepted sub-cells all belong to one component add to
// per-lane running sums, summed over the wavefront only when the component
// changes.
template <int SOURCE, bool LAST, bool GRAD = false>
__global__ __launch_bounds__(64) void refine32_kernel(const PkJob *__restrict__ jobs, const PkComp *__restrict__ comps,
                                                      const CompLite *__restrict__ lite,
                                                      const RefineEntry *__restrict__ entries,
                                                      PkCounters *__restrict__ cnt, const RCell *__restrict__ qin,
                                                      RCell *__restrict__ qout, uint32_t qcap, int pass,
                                                      unsigned long long *__restrict__ acc,
                                                      R32Tune tune,
                                                      const double *__restrict__ gent = nullptr,
                                                      double *__restrict__ gacc = nullptr)
{
	constexpr int CAP = LAST ? R32_CAP_LAST : R32_CAP_MAIN;
	__shared__ R32Stack<CAP> st;
	const uint32_t lane = lane_id();
	if (lane < (uint32_t)R32_TASKS) {
		st.lite_g[lane] = ~0u;
		st.csum[lane] = 0;
	}
	int top = 0;
	bool source_done = false, handover = false;
	const uint32_t n_cells_in = SOURCE == 0 ? 0u : min(cnt->n_cells[pass - 1], qcap);
	// Handed-over cells are claimed in runs of 2^R32_CELL_RUN taken from 64
	// places of the queue in turn, so that the cells one wavefront handed
	// over, often parts of the same deep pixel, go to many wavefronts, while
	// each run (cells handed over together: mostly one component and
	// resolution) stays together on one LIFO.
	const uint32_t stride = ((((n_cells_in + (1u << R32_CELL_RUN) - 1u) >> R32_CELL_RUN) + 63u) / 64u);
	const uint32_t n_entries = cnt->n_entries;
	const uint32_t spread = SOURCE == 0 ? tune.spread_log2 : 0u;
	const uint32_t sstride = ((((n_entries + (1u << spread) - 1u) >> spread) + 63u) / 64u);
	const uint32_t n_source =
	    SOURCE == 0 ? (spread ? (64u * sstride) << spread : n_entries) : (64u * stride) << R32_CELL_RUN;
	uint32_t *head = &cnt->claim[pass];
	uint32_t overflow = 0, bad = 0;
	unsigned long long nev = 0;
	uint32_t chunk_next = 0, chunk_end = 0, seen = 0;
	int poll = 0;
	// per-lane running fixed-point sum of the pixel accumulator run_t (rounds
	// of one cell whose pixel is the last one's add to it)
	unsigned long long run_fx = 0;
	uint64_t run_t = ~0ull;
	// GRAD: per-lane running sums of two components, run_g and run_g2
	double run[GA_SLOTS] = {0, 0, 0, 0, 0, 0, 0, 0}, run2[GA_SLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
	uint32_t run_g = ~0u, run_g2 = ~0u;
	__syncthreads();
	for (;;) {
		// ---- refill while the cells on the LIFO cannot fill the lanes (one
		// row of sub-cells per lane) ----
		for (;;) {
			if (top >= 64) break;  // at least one row per cell
			uint32_t rows = 0;
			if ((int)lane < top) rows = cm_res(st.cell[top - 1 - (int)lane].meta);
			if (wave_sum_u32(rows) >= 64u || source_done) break;
			if (CAP - top < (SOURCE == 0 ? 32 : 1)) break;
			if (chunk_next >= chunk_end &&
			    !r32_claim(head, n_source, SOURCE == 0 ? tune.chunk : tune.chunk_cells, seen, chunk_next, chunk_end,
			               lane)) {
				source_done = true;
				break;
			}
			if (SOURCE == 0) {
				r32_expand(jobs, comps, entries, n_entries, spread, sstride, st.cell, CAP, top, chunk_next, chunk_end,
				           lane);
			} else {
				r32_take_cells(qin, n_cells_in, stride, st.cell, CAP, top, chunk_next, chunk_end, lane);
			}
			__syncthreads();
		}
		if (top == 0) break;
		if (!LAST && ++poll >= (int)tune.poll) {
			poll = 0;
			uint32_t d = 0;
			if (lane == 0) d = __hip_atomic_load(&cnt->n_done[pass], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			if (wave_read(d, 0) * tune.handover_div >= gridDim.x) {
				handover = true;
				break;
			}
		}

		// ---- cells on top of the LIFO -> lanes ----
		// A group is the run of cells from the top that share the top cell's
		// resolution res: 64 / res of them (at most R32_TASKS).  Each lane
		// takes one row si of one cell and walks that row's res sub-cells
		// (si, 0 .. res - 1), so the per-group work (task map, component
		// record, cell geometry, sums) is shared by res sub-cells.  A cell of
		// more than 64 rows takes several rounds of 64 rows.
		const uint32_t res = cm_res(st.cell[top - 1].meta);
		bool same = false;
		if ((int)lane < top && lane < (uint32_t)R32_TASKS) same = cm_res(st.cell[top - 1 - (int)lane].meta) == res;
		const uint64_t not_same = ~__ballot(same);
		const int run_len = not_same ? __ffsll((long long)not_same) - 1 : 64;  // >= 1: the top cell itself
		const int ntask = min(run_len, res >= 64u ? 1 : (int)(64u / max(res, 1u)));
		const int rounds = res > 64u ? (int)((res + 63u) / 64u) : 1;
		// lane -> task t_local (cell top - 1 - t_local), first row si0
		int t_local = 0, si0 = (int)lane;
		if (res <= 64u) {
			t_local = (int)((float)lane * frcp((float)res));
			if ((uint32_t)t_local * res > lane) t_local--;
			else if ((uint32_t)(t_local + 1) * res <= lane) t_local++;
			si0 = (int)lane - t_local * (int)res;
		}
		const bool lane_task = t_local < ntask;
		if (!lane_task) t_local = ntask - 1;  // beyond the last task: inactive, mirrors the last task
		const RCell cell = st.cell[top - 1 - t_local];
		// component records of the tasks -> LDS (kept while unchanged); lane
		// t < ntask keeps the accumulator of task t
		size_t task_target = 0;
		{
			uint32_t g = 0;
			bool miss = false;
			if ((int)lane < ntask) {
				const RCell &tc = st.cell[top - 1 - (int)lane];
				g = tc.gcomp;
				task_target = 2 * (size_t)tc.entry + cm_conv(tc.meta);
				miss = st.lite_g[lane] != g;
			}
			if (__ballot(miss)) {
				if (miss) {
					st.lite[lane] = lite[g];
					st.lite_g[lane] = g;
				}
				__syncthreads();
			}
		}
		top -= ntask;
		// the task's component record, copied into registers: read inside the
		// sub-cell loop from LDS it would be reloaded at every sub-cell
		const CompLite cl = st.lite[t_local];
		const uint32_t meta = cell.meta;
		const int level = (int)cm_level(meta), mr = (int)cm_maxrec(meta);
		// sub-cell size in pixels, in FP64 as libprofit forms the sub-pixel
		// grid: an FP32 size would move the sub-cell centres by ~1e-8 px, which
		// next to a cusp (coresersic, f ~ r^-b) changes the innermost values
		// by ~1e-4
		const double fr = cell_size_fp64((int)res, level + 1);
		const double hx = fr * cl.hsx, hy = fr * cl.hsy;
		const float wq = (float)(fr * fr * cl.wq);
		const bool recurse = res > 1u && level < mr;
		const float acc_thr = (float)cl.acc;
		const float band = 1e-4f * (1.f + acc_thr);
		const double ddy = (-2 * hx * cl.sin_ang + 2 * hy * cl.cos_ang) * cl.inv_axrat;
		const size_t target = 2 * (size_t)cell.entry + cm_conv(meta);
		const double gw = GRAD ? gent[target] : 0.0;  // G / fx_scale of the cell's pixel
		const int kind = cl.fp32 ? (int)cl.kind : -1;
		// wavefront-uniform facts of the group (lanes beyond the last task
		// share the last task's)
		const bool any_rec = __ballot(recurse) != 0;
		const int kind0 = (int)wave_read((uint32_t)kind, 0);
		const bool one_kind = __ballot(kind != kind0) == 0;
		// all tasks add to one pixel accumulator (one cell, or siblings)
		const uint64_t tu = (uint64_t)wave_read((uint32_t)(task_target >> 32), 0) << 32 |
		                    wave_read((uint32_t)task_target, 0);
		const bool one_target = __ballot((int)lane < ntask && task_target != tu) == 0;
		const uint32_t nev_sub = recurse ? 2u : 1u;
		for (int rd = 0; rd < rounds; rd++) {
			const int si = si0 + 64 * rd;
			const bool active = lane_task && si < (int)res;
			const double dx = fma((double)(2 * si + 1), hx, cell.dx0);
			const double dxc = dx * cl.cos_ang, ndxs = -dx * cl.sin_ang;
			unsigned long long fxv = 0;  // fixed-point sum of this lane's accepted sub-cells
			double ga[GA_SLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
			bool any_ok = false;
			for (int sj = 0; sj < (int)res; sj++) {
				const double dy = fma((double)(2 * sj + 1), hy, cell.dy0);
				const double xp = fma(dy, cl.sin_ang, dxc);
				const double yp = fma(dy, cl.cos_ang, ndxs) * cl.inv_axrat;
				float v32 = -1.f, t32 = -1.f;
				if (one_kind) {
					// the usual case: one profile kind in all lanes, scalar branch
					if (any_rec) evaluate_pair32<true>(cl, kind0, xp, yp, fabs(xp), fabs(yp) + fabs(ddy), v32, t32);
					else evaluate_pair32<false>(cl, kind0, xp, yp, 0., 0., v32, t32);
				} else {
					evaluate_pair32<true>(cl, kind, xp, yp, fabs(xp), fabs(yp) + fabs(ddy), v32, t32);
				}
				// accuracy test (src/radial.cpp:86-109) in FP32, decided outside a
				// guard band around the threshold; written without branches
				const float m = fabsf(t32 * frcp(v32) - 1.f);
				const bool dec_rec = (v32 > 0.f) & (v32 <= 3e38f) & (m <= 3e38f) & (fabsf(m - acc_thr) > band);
				const bool decided = recurse ? dec_rec : (v32 >= 0.f) & (v32 <= 3e38f);
				bool fail = recurse & dec_rec & (m > acc_thr);
				if (!GRAD) {
					// accepted value as a 64-bit fixed-point integer (truncated)
					const bool take = active & decided & !fail;
					const float q = v32 * wq;
					const bool q_ok = (q >= 0.f) & (q < 9.2e18f);
					const float qc = q_ok ? q : 0.f;
					const float hi = truncf(qc * 0x1p-32f);
					const float lo = fmaf(hi, -0x1p32f, qc);  // exact
					const unsigned long long f = ((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo;
					fxv += (take & q_ok) ? f : 0ull;
					bad += (take & !q_ok) ? 1u : 0u;
				}
				if (__ballot(active && !decided)) {
					// FP64 value and decision within the guard band (or without an FP32 form)
					if (active && !decided) {
						const double val = refine_fp64_decision(comps[cell.gcomp], xp, yp, ddy, recurse, &fail);
						if (!GRAD && !fail) {
							const double qv = val * (fr * fr * cl.wflux);
							if (qv >= 0 && qv < 9.2e18) fxv += (unsigned long long)__double2ll_rn(qv);
							else bad++;
						}
					}
				}
				fail = fail && active;
				if (GRAD && active && !fail) {
					any_ok = true;
					float gt[GT_COUNT];
					const float vg = sersic_grad32(cl, (float)xp, (float)yp, gt);
					grad_accumulate(ga, gw * (fr * fr * cl.wq) * (double)vg, gt);
				}
				// failing sub-cells: onto the LIFO, or to the next pass when it is full
				const uint64_t ballot = __ballot(fail);
				if (ballot) {
					const int n = __popcll(ballot);
					const int rank = __popcll(ballot & ((1ull << lane) - 1ull));
					RCell c;
					c.dx0 = dx - hx;
					c.dy0 = dy - hy;
					c.gcomp = cell.gcomp;
					c.entry = cell.entry;
					c.meta = (meta & ~255u) | (uint32_t)(level + 1);
					if (top + n <= CAP) {
						if (fail) st.cell[top + rank] = c;
						top += n;
					} else if (!LAST) {
						const int leader = __ffsll((long long)ballot) - 1;
						uint32_t base = 0;
						if ((int)lane == leader) base = atomicAdd(&cnt->n_cells[pass], (uint32_t)n);
						base = wave_read(base, leader);
						if (fail) {
							if (base + (uint32_t)rank < qcap) qout[base + rank] = c;
							else overflow = 1;
						}
					} else {
						overflow = 1;
					}
				}
			}
			if (active) nev += (unsigned long long)(nev_sub * res);
			if (GRAD) {
				const uint64_t okb = __ballot(any_ok);
				if (okb) {
					const uint32_t g0 = wave_read(cell.gcomp, __ffsll((long long)okb) - 1);
					const uint64_t ob = __ballot(any_ok && cell.gcomp != g0);
					const uint32_t g1 = ob ? wave_read(cell.gcomp, __ffsll((long long)ob) - 1) : ~0u;
					if (!ob || __ballot(any_ok && cell.gcomp != g0 && cell.gcomp != g1) == 0) {
						// at most two components: keep adding to the running sums,
						// after handing on a slot held by a third one
						auto hold = [&](uint32_t g, uint32_t keep) {
							if (g == ~0u || g == run_g || g == run_g2) return;
							if (run_g != keep) {
								r32_flush_run(run, run_g, gacc, lane);
								run_g = g;
							} else {
								r32_flush_run(run2, run_g2, gacc, lane);
								run_g2 = g;
							}
						};
						hold(g0, g1);
						hold(g1, g0);
						const bool in1 = any_ok && cell.gcomp == run_g, in2 = any_ok && cell.gcomp == run_g2;
#pragma unroll
						for (int k = 0; k < GA_SLOTS; k++) {
							run[k] += in1 ? ga[k] : 0.0;
							run2[k] += in2 ? ga[k] : 0.0;
						}
					} else {
						r32_flush_grad(any_ok, cell.gcomp, ga, gacc, lane);
					}
				}
			} else if (one_target) {
				// one pixel accumulator for all lanes: exact per-lane sums, added
				// to it (lane 0) when the pixel changes
				if (tu != run_t) {
					r32_flush_fx(run_fx, run_t, acc, lane);
					run_t = tu;
				}
				run_fx += fxv;
			} else {
				// several pixels: exact integer sums per cell in LDS, then one
				// global atomic per cell (lane t adds task t's sum)
				if (fxv) atomicAdd(&st.csum[t_local], fxv);
				__builtin_amdgcn_wave_barrier();
				if ((int)lane < ntask) {
					const unsigned long long v = st.csum[lane];
					st.csum[lane] = 0;
					if (v) atomicAdd(&acc[task_target], v);
				}
			}
			__syncthreads();
		}
	}
	if (GRAD) {
		r32_flush_run(run, run_g, gacc, lane);
		r32_flush_run(run2, run_g2, gacc, lane);
	} else {
		r32_flush_fx(run_fx, run_t, acc, lane);
	}
	if (!LAST && handover) {
		// everything this wavefront holds goes to the next pass
		for (;;) {
			if (top > 0) {
				uint32_t base = 0;
				if (lane == 0) base = atomicAdd(&cnt->n_cells[pass], (uint32_t)top);
				base = wave_read(base, 0);
				for (int i = (int)lane; i < top; i += 64) {
					if (base + (uint32_t)i < qcap) qout[base + i] = st.cell[i];
					else overflow = 1;
				}
				top = 0;
				__syncthreads();
			}
			if (chunk_next >= chunk_end) break;
			if (SOURCE == 0) {
				r32_expand(jobs, comps, entries, n_entries, spread, sstride, st.cell, CAP, top, chunk_next, chunk_end,
				           lane);
			} else {
				r32_take_cells(qin, n_cells_in, stride, st.cell, CAP, top, chunk_next, chunk_end, lane);
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
		if (!LAST) atomicAdd(&cnt->n_done[pass], 1u);
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
