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
// Load balance: the refinement runs as n launches (4 <= n <= R32_PASSES, a
// tuned launch setting), one level of the sub-pixel recursion per launch but
// the last.  Pass 0 takes pixels from the
// coarse pass's entry list; pass p > 0 takes the cells of level p that pass
// p - 1 produced.  The failing sub-cells of a level-pass go to the next
// pass's queue (staged in LDS and written 64 at a time, one atomic per 64),
// so every pass starts from work spread evenly over all wavefronts, whose
// items cost at most res^2 sub-cells each: no wavefront is left with a deep
// pixel's whole subtree while the others idle.  Work is claimed in chunks
// that shrink as the source runs dry.  The last pass refines what remains
// (levels >= n - 1: the Sersic centre pixels' deep recursion) depth-first on
// a LIFO.  A queue that fills up keeps the cells on the
// wavefront's LIFO instead.
//
// Guard band: sub-cells whose FP32 accuracy test falls within the guard band
// of the threshold are decided in FP64.  Outside the last pass they are
// listed and decided by r32_resolve_kernel between the passes (the FP64
// evaluation's call would raise the main kernels' register count by a
// third); a failing one joins the next pass's queue.
//
// Latency: lane-crossing scans and whole-wavefront sums use DPP; a round
// holding several cells sums each cell's sub-cells with LDS integer atomics
// (one slot per cell) and adds the slots to the pixels.
#pragma once

namespace pk {

// LIFO capacity: in the level-passes it holds the claimed cells (the
// children go to the next pass); the last pass refines depth-first, which
// holds at most about res^2 * (levels left) cells: 1024 covers the Sersic
// centre pixels (res 8, 10 levels).
constexpr int R32_CAP_MAIN = 96;
constexpr int R32_CAP_LAST = 1024;
// children staged in LDS before they are written to the next pass's queue:
// written once R32_FLUSH are staged, so one more row (<= 64) always fits
constexpr int R32_FLUSH = 64;
constexpr int R32_OUTB = R32_FLUSH + 64;
// workgroups of the launches that decide handed-over sub-cells in FP64
constexpr uint32_t R32_RESOLVE_GRID = 1024;
constexpr uint32_t R32_CELL_RUN = 4;       // log2 of the runs in which handed-over cells are claimed
constexpr uint32_t R32_LEAST_CLAIM = 16;   // items of the smallest claim (when the tuned chunk is not smaller)
constexpr int R32_TASKS = 16;              // cells per round at most

// Wavefronts per workgroup of a refinement launch.  The wavefronts of a
// workgroup share nothing (each has its own LIFO), so any number works; one
// per workgroup measured fastest (two: 10 % slower on the galaxy batch).
constexpr int r32_waves_per_group(bool, bool) { return 1; }

// Orders this wavefront's LDS accesses (the wavefronts of a workgroup do not
// wait for each other).
__device__ inline void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
	__builtin_amdgcn_wave_barrier();
}
// untuned settings: claims of up to 256 pixels / 256 cells, in entry order
// (smaller claims contend on the queue counters), five launches
constexpr R32Tune R32_TUNE_DEFAULT = {4u, 256u, 256u, 5u, 0u};
constexpr uint32_t R32_PASSES_MIN = 4;  // the last launch refines levels >= passes - 1

template <int CAP, int OUTB>
struct R32Stack {
	unsigned long long csum[R32_TASKS];  // fixed-point sum of task t in the current round (0 between rounds)
	RCell cell[CAP];
	RCell outb[OUTB];                    // staged children (level-passes)
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

// Keeps x in vector registers from here on (an empty asm that takes and
// returns it in a VGPR).
template <typename T>
__device__ inline void pin_vgpr(T &x)
{
	asm volatile("" : "+v"(x));
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
// queue runs dry (`seen` is the head this wavefront saw at its last claim),
// down to `least` items: claims wait on a counter that every wavefront of the
// launch contends for, so the last items must not cost one claim each.
// (Reading the head's current value before each claim was measured slower.)
__device__ inline bool r32_claim(uint32_t *head, uint32_t n, uint32_t chunk, uint32_t least, uint32_t nwaves,
                                 uint32_t &seen, uint32_t &next, uint32_t &end, uint32_t lane)
{
	const uint32_t left = n > seen ? n - seen : 0u;
	const uint32_t want = max(least, min(chunk, left / (2u * nwaves)));
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

// Writes the n staged children to the next pass's queue (one atomic for all of
// them) and clears the staging buffer.  Children past the queue's capacity go
// onto the LIFO (refined by this wavefront) while it has room.  Copied as
// 32-bit words: one register per lane in flight, coalesced stores.
template <int CAP>
__device__ inline void r32_flush_out(RCell *outb, int &n, RCell *__restrict__ qout, uint32_t qcap, uint32_t *count,
                                     RCell *cells, int cap, int &top, uint32_t &overflow, uint32_t lane)
{
	constexpr int WORDS = (int)(sizeof(RCell) / 4);
	if (n == 0) return;
	uint32_t base = 0;
	if (lane == 0) base = atomicAdd(count, (uint32_t)n);
	base = wave_read(base, 0);
	const int fit = base >= qcap ? 0 : (int)min((uint32_t)n, qcap - base);  // children that go to the queue
	const uint32_t *src = reinterpret_cast<const uint32_t *>(outb);
	uint32_t *dst = reinterpret_cast<uint32_t *>(qout + base);
	for (int i = (int)lane; i < fit * WORDS; i += 64) dst[i] = src[i];
	if (fit < n) {
		const int k = n - fit;
		if (top + k <= cap) {
			uint32_t *lifo = reinterpret_cast<uint32_t *>(cells + top);
			for (int i = (int)lane; i < k * WORDS; i += 64) lifo[i] = src[fit * WORDS + i];
			top += k;
		} else {
			overflow = 1;
		}
	}
	n = 0;
	wave_sync();
}

// SOURCE 0: pixels from the coarse pass (pass 0); 1: cells of the previous pass.
// LAST: the final pass, which refines depth-first and hands nothing over.
// Level-passes write their children to qout and their guard-band sub-cells
// to pq (decided by r32_resolve_kernel before the next pass).
// GRAD: the chi^2 adjoint of the refinement -- the same cells and decisions as
// the forward pass, but every accepted sub-cell adds G w flux_scale f dL/dθ
// to gacc instead of its value to the pixel.  gent[2 e + convolved] holds
// G / fx_scale of refinement entry e (G: the adjoint image at its pixel).
// Rounds whose accepted sub-cells all belong to one component add to
// per-lane running sums, summed over the wavefront only when the component
// changes.
template <int SOURCE, bool LAST, bool GRAD = false>
__global__ __launch_bounds__(128) void refine32_kernel(const PkJob *__restrict__ jobs, const PkComp *__restrict__ comps,
                                                      const CompLite *__restrict__ lite,
                                                      const RefineEntry *__restrict__ entries,
                                                      PkCounters *__restrict__ cnt, const RCell *__restrict__ qin,
                                                      RCell *__restrict__ qout, uint32_t qcap, int pass,
                                                      unsigned long long *__restrict__ acc,
                                                      R32Tune tune, RCell *__restrict__ pq,
                                                      const double *__restrict__ gent = nullptr,
                                                      double *__restrict__ gacc = nullptr)
{
	constexpr int CAP = LAST ? R32_CAP_LAST : R32_CAP_MAIN;
	constexpr int WPG = r32_waves_per_group(LAST, GRAD);
	__shared__ R32Stack<CAP, LAST ? 1 : R32_OUTB> stacks[WPG];
	// (the wavefront's index in its workgroup, as a uniform value)
	const uint32_t wig = WPG == 1 ? 0u : (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x / 64);
	R32Stack<CAP, LAST ? 1 : R32_OUTB> &st = stacks[wig];
	const uint32_t lane = lane_id();
	const uint32_t wid = blockIdx.x * WPG + wig;  // this wavefront among the launch's
	const uint32_t nwaves = gridDim.x * WPG;
	if (lane < (uint32_t)R32_TASKS) st.csum[lane] = 0;
	int top = 0;
	int nout = 0;  // children staged in st.outb
	bool source_done = false;
	const uint32_t n_cells_in = SOURCE == 0 ? 0u : min(cnt->n_cells[pass - 1], qcap);
	// Cells of the previous pass are claimed in runs of 2^R32_CELL_RUN taken
	// from 64 places of the queue in turn, so that the children of one deep
	// pixel go to many wavefronts, while each run (written together: mostly
	// one component and resolution) stays together on one LIFO.
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
	const uint32_t chunk_max = SOURCE == 0 ? tune.chunk : tune.chunk_cells;
	const uint32_t least = min(R32_LEAST_CLAIM, chunk_max);
	// wavefronts past the work's smallest claims have nothing to claim: they
	// stop without touching the queue's counter (a pass over a short queue
	// would otherwise spend its time on that counter)
	if ((size_t)wid * least >= n_source) source_done = true;
	// per-lane running fixed-point sum of the pixel accumulator run_t (rounds
	// of one cell whose pixel is the last one's add to it)
	unsigned long long run_fx = 0;
	uint64_t run_t = ~0ull;
	// GRAD: per-lane running sums of component run_g (a group belongs to one
	// component; the sums are handed on when the component changes)
	double run[GA_SLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
	uint32_t run_g = ~0u;
	wave_sync();
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
			    !r32_claim(head, n_source, chunk_max, least, nwaves, seen, chunk_next, chunk_end, lane)) {
				source_done = true;
				break;
			}
			if (SOURCE == 0) {
				r32_expand(jobs, comps, entries, n_entries, spread, sstride, st.cell, CAP, top, chunk_next, chunk_end,
				           lane);
			} else {
				r32_take_cells(qin, n_cells_in, stride, st.cell, CAP, top, chunk_next, chunk_end, lane);
			}
			wave_sync();
		}
		if (top == 0) {
			// the source is exhausted: the staged children go to the next pass
			// (those a full queue leaves on the LIFO are refined here)
			if (!LAST && nout > 0)
				r32_flush_out<CAP>(st.outb, nout, qout, qcap, &cnt->n_cells[pass], st.cell, CAP, top, overflow, lane);
			if (top == 0) break;
		}

		// ---- cells on top of the LIFO -> lanes ----
		// A group is the run of cells from the top that share the top cell's
		// component and metadata (resolution res, recursion depth, convolved
		// flag; their levels may differ): 64 / res of them at most (and at
		// most R32_TASKS).  The component record of a group is wavefront-
		// uniform (scalar loads), so the profile kind and the box flag are
		// branched on once per group, outside the sub-cell loop.  Each lane
		// takes one row si of one cell and walks that row's res sub-cells
		// (si, 0 .. res - 1), so the per-group work (task map, cell geometry,
		// sums) is shared by res sub-cells.  A cell of more than 64 rows takes
		// several rounds of 64 rows.
		const uint32_t g0 = wave_read(st.cell[top - 1].gcomp, 0);
		const uint32_t key = wave_read(st.cell[top - 1].meta >> 8, 0);
		bool same = false;
		size_t task_target = 0;  // lane t < ntask: pixel accumulator of task t
		if ((int)lane < top && lane < (uint32_t)R32_TASKS) {
			const RCell &tc = st.cell[top - 1 - (int)lane];
			same = tc.gcomp == g0 && (tc.meta >> 8) == key;
			task_target = 2 * (size_t)tc.entry + ((key >> 16) & 1u);
		}
		const uint64_t not_same = ~__ballot(same);
		const int run_len = not_same ? __ffsll((long long)not_same) - 1 : 64;  // >= 1: the top cell itself
		const uint32_t res = key & 255u;
		const int mr = (int)((key >> 8) & 255u);
		const uint32_t conv = (key >> 16) & 1u;
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
		top -= ntask;
		// the group's component record: one scalar load per group
		const CompLite L = lite[g0];
		const int level = (int)cm_level(cell.meta);
		// sub-cell size in pixels, in FP64 as libprofit forms the sub-pixel
		// grid: an FP32 size would move the sub-cell centres by ~1e-8 px, which
		// next to a cusp (coresersic, f ~ r^-b) changes the innermost values
		// by ~1e-4
		const double fr = cell_size_fp64((int)res, level + 1);
		const double hx = fr * L.hsx, hy = fr * L.hsy;
		const double wrow = fr * fr * L.wq;  // weight of an accepted FP32 value
		const bool recurse = res > 1u && level < mr;
		const float acc_thr = (float)L.acc;
		const float band = 1e-4f * (1.f + acc_thr);
		const double ddy = (-2 * hx * L.sin_ang + 2 * hy * L.cos_ang) * L.inv_axrat;
		const size_t target = 2 * (size_t)cell.entry + conv;
		const double gw = GRAD ? gent[target] : 0.0;  // G / fx_scale of the cell's pixel
		const bool any_rec = __ballot(recurse) != 0;
		// all tasks add to one pixel accumulator (one cell, or siblings)
		const uint64_t tu = (uint64_t)wave_read((uint32_t)(target >> 32), 0) << 32 | wave_read((uint32_t)target, 0);
		const bool one_target = __ballot(lane_task && target != tu) == 0;
		const uint32_t nev_sub = recurse ? 2u : 1u;
		if (GRAD && g0 != run_g) {
			r32_flush_run(run, run_g, gacc, lane);
			run_g = g0;
		}
		// the rows of the group for profile kind KIND (-1: no FP32 form, every
		// sub-cell decided in FP64) with the box flag BOXY
		auto rows = [&](auto kind_c, auto boxy_c) __attribute__((always_inline)) {
			constexpr int KIND = decltype(kind_c)::value;
			// the constants the sub-cell loop reads, held in vector registers:
			// in scalar registers they would not fit beside the kernel's other
			// uniform state and be reloaded from spill lanes at every sub-cell
			CompLiteB<decltype(boxy_c)::value> cl(L);
			pin_vgpr(cl.cos_ang);
			pin_vgpr(cl.sin_ang);
			pin_vgpr(cl.inv_axrat);
#pragma unroll
			for (int k = 0; k < 8; k++) pin_vgpr(cl.f[k]);
			if (KIND == PK_PROFILE_KING || KIND == PK_PROFILE_FERRER) {
				pin_vgpr(cl.q4);
				pin_vgpr(cl.q5);
				pin_vgpr(cl.b2);
			}
			float acc_v = acc_thr, band_v = band;
			pin_vgpr(acc_v);
			pin_vgpr(band_v);
			for (int rd = 0; rd < rounds; rd++) {
				const int si = si0 + 64 * rd;
				const bool active = lane_task && si < (int)res;
				const double dx = fma((double)(2 * si + 1), hx, cell.dx0);
				const double dxc = dx * cl.cos_ang, ndxs = -dx * cl.sin_ang;
				// accepted values of this lane's row: FP32 values summed in FP64 (in
				// sub-cell order, so the sum does not depend on scheduling) and
				// turned into fixed point once per row; FP64 values in fixed point
				double rs = 0;
				unsigned long long fxv = 0;
				// the sub-cell of this lane's row at column centre dy as a cell of
				// the next level
				auto child = [&](double dy) __attribute__((always_inline)) {
					RCell c;
					c.dx0 = dx - hx;
					c.dy0 = dy - hy;
					c.gcomp = g0;
					c.entry = cell.entry;
					c.meta = (key << 8) | (uint32_t)(level + 1);
					return c;
				};
				// failing sub-cells: staged for the next pass (level-passes), or
				// onto the LIFO (last pass)
				auto push_failing = [&](bool fail, double dy) __attribute__((always_inline)) {
					const uint64_t ballot = __ballot(fail);
					if (!ballot) return;
					const int n = __popcll(ballot);
					const int rank = __popcll(ballot & ((1ull << lane) - 1ull));
					if (!LAST) {
						// nout < R32_FLUSH here, so n <= 64 more fit
						if (fail) st.outb[nout + rank] = child(dy);
						nout += n;
						if (nout >= R32_FLUSH)
							r32_flush_out<CAP>(st.outb, nout, qout, qcap, &cnt->n_cells[pass], st.cell, CAP, top,
							                   overflow, lane);
					} else if (top + n <= CAP) {
						if (fail) st.cell[top + rank] = child(dy);
						top += n;
					} else {
						overflow = 1;
					}
				};
				// the row in chunks of 64 sub-cells.  Sub-cells whose FP32 decision
				// fell within the guard band are listed for r32_resolve_kernel
				// (level-passes) or, in the last pass (bit sj - sj0 of `pend`),
				// decided here after the chunk.
				for (int sj0 = 0; sj0 < (int)res; sj0 += 64) {
					const int sj1 = min((int)res, sj0 + 64);
					uint64_t pend = 0;
					for (int sj = sj0; sj < sj1; sj++) {
						const double dy = fma((double)(2 * sj + 1), hy, cell.dy0);
						const double xp = fma(dy, cl.sin_ang, dxc);
						const double yp = fma(dy, cl.cos_ang, ndxs) * cl.inv_axrat;
						float v32 = -1.f, t32 = -1.f;
						if (any_rec) evaluate_pair32<true>(cl, KIND, xp, yp, fabs(xp), fabs(yp) + fabs(ddy), v32, t32);
						else evaluate_pair32<false>(cl, KIND, xp, yp, 0., 0., v32, t32);
						// accuracy test (src/radial.cpp:86-109) |t / v - 1| > acc, for
						// v > 0 the sign of e = |t - v| - acc v; decided in FP32 outside a
						// guard band |e| <= 1e-4 (1 + acc) v; written without branches
						const float e = fmaf(-acc_v, v32, fabsf(t32 - v32));
						const bool dec_rec = (v32 > 0.f) & (v32 <= 3e38f) & (t32 <= 3e38f) & (fabsf(e) > band_v * v32);
						const bool decided = recurse ? dec_rec : (v32 >= 0.f) & (v32 <= 3e38f);
						const bool fail = active & recurse & dec_rec & (e > 0.f);
						if (!GRAD) rs += (active & decided & !fail) ? (double)v32 : 0.0;
						if (LAST) {
							pend |= (active & !decided) ? 1ull << (sj - sj0) : 0ull;
						} else {
							const bool pnd = active & !decided;
							const uint64_t pb = __ballot(pnd);
							if (pb) {
								const int leader = __ffsll((long long)pb) - 1;
								uint32_t base = 0;
								if ((int)lane == leader) base = atomicAdd(&cnt->n_pend[pass], (uint32_t)__popcll(pb));
								base = wave_read(base, leader);
								const uint32_t slot = base + (uint32_t)__popcll(pb & ((1ull << lane) - 1ull));
								if (pnd) {
									if (slot < qcap) pq[slot] = child(dy);
									else overflow = 1;
								}
							}
						}
						if (GRAD && active && decided && !fail) {
							float gt[GT_COUNT];
							const float vg = sersic_grad32(cl, (float)xp, (float)yp, gt);
							grad_accumulate(run, gw * wrow * (double)vg, gt);
						}
						push_failing(fail, dy);
					}
					// FP64 value and decision of the sub-cells within the guard band
					// (or of every sub-cell without an FP32 form), lowest column first
					while (LAST && __ballot(pend != 0)) {
						bool fail = false;
						double dy = 0;
						if (pend) {
							const int sj = sj0 + __ffsll((long long)pend) - 1;
							pend &= pend - 1;
							dy = fma((double)(2 * sj + 1), hy, cell.dy0);
							const double xp = fma(dy, cl.sin_ang, dxc);
							const double yp = fma(dy, cl.cos_ang, ndxs) * cl.inv_axrat;
							const double val = refine_fp64_decision(comps[g0], xp, yp, ddy, recurse, &fail);
							if (!fail) {
								if (GRAD) {
									float gt[GT_COUNT];
									const float vg = sersic_grad32(cl, (float)xp, (float)yp, gt);
									grad_accumulate(run, gw * wrow * (double)vg, gt);
								} else {
									const double qv = val * (fr * fr * cl.wflux);
									if (qv >= 0 && qv < 9.2e18) fxv += (unsigned long long)__double2ll_rn(qv);
									else bad++;
								}
							}
						}
						push_failing(fail, dy);
					}
				}
				if (!GRAD) {
					// the row's FP32 values in fixed point (truncated)
					const double q = rs * wrow;
					if (q >= 0 && q < 9.2e18) fxv += (unsigned long long)q;
					else bad++;
				}
				if (active) nev += (unsigned long long)(nev_sub * res);
				if (GRAD) {
					// the adjoint terms went to the running sums
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
				wave_sync();
			}
		};
		using BoxN = std::false_type;
		using BoxY = std::true_type;
		const int kind = L.fp32 ? (int)L.kind : -1;
		if (GRAD) {
			// gradient batches hold sersic profiles only
			if (L.boxy) rows(std::integral_constant<int, PK_PROFILE_SERSIC>{}, BoxY{});
			else rows(std::integral_constant<int, PK_PROFILE_SERSIC>{}, BoxN{});
		} else {
#define PK_R32_KIND(K)                                                        \
	case K:                                                               \
		if (L.boxy) rows(std::integral_constant<int, K>{}, BoxY{});     \
		else rows(std::integral_constant<int, K>{}, BoxN{});            \
		break;
			switch (kind) {
				PK_R32_KIND(PK_PROFILE_SERSIC)
				PK_R32_KIND(PK_PROFILE_MOFFAT)
				PK_R32_KIND(PK_PROFILE_CORESERSIC)
				PK_R32_KIND(PK_PROFILE_BROKENEXP)
			// king and ferrer: the FP32 forms of round profiles only (their
			// boxy radius needs FP64 powers, left to the FP64 evaluation)
			case PK_PROFILE_KING:
				if (!L.boxy) rows(std::integral_constant<int, PK_PROFILE_KING>{}, BoxN{});
				else rows(std::integral_constant<int, -1>{}, BoxN{});
				break;
			case PK_PROFILE_FERRER:
				if (!L.boxy) rows(std::integral_constant<int, PK_PROFILE_FERRER>{}, BoxN{});
				else rows(std::integral_constant<int, -1>{}, BoxN{});
				break;
			default: rows(std::integral_constant<int, -1>{}, BoxN{}); break;
			}
#undef PK_R32_KIND
		}
	}
	if (GRAD) {
		r32_flush_run(run, run_g, gacc, lane);
	} else {
		r32_flush_fx(run_fx, run_t, acc, lane);
	}
	for (int off = 32; off > 0; off >>= 1) nev += __shfl_xor(nev, off);
	const uint64_t ov = __ballot(overflow != 0);
	const uint32_t nbad = wave_sum_u32(bad);
	if (lane == 0) {
		if (nev) atomicAdd(&cnt->n_subcells, nev);
		if (ov) atomicOr(&cnt->overflow, 1u);
		if (nbad) atomicAdd(&cnt->n_bad, nbad);
	}
}

// Between two passes: decides in FP64 the guard-band sub-cells that pass p
// listed in pq (as cells of their own level).  An accepted one adds its value
// to its pixel's accumulator (GRAD: its adjoint terms to its component); a
// failing one joins the queue of pass p + 1.  The sub-cell's centre is its
// corner plus half its size.
template <bool GRAD>
__global__ __launch_bounds__(256) void r32_resolve_kernel(const PkComp *__restrict__ comps,
                                                          const CompLite *__restrict__ lite,
                                                          PkCounters *__restrict__ cnt, const RCell *__restrict__ pq,
                                                          RCell *__restrict__ qout, uint32_t qcap, int pass,
                                                          unsigned long long *__restrict__ acc,
                                                          const double *__restrict__ gent = nullptr,
                                                          double *__restrict__ gacc = nullptr)
{
	const uint32_t n = min(cnt->n_pend[pass], qcap);
	uint32_t bad = 0, overflow = 0;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		const RCell c = pq[i];
		const uint32_t meta = c.meta;
		const CompLite &L = lite[c.gcomp];
		const int lv = (int)cm_level(meta);  // the sub-cell's level as a cell (its parent's + 1)
		const uint32_t res = cm_res(meta);
		const double fr = cell_size_fp64((int)res, lv);
		const double hx = fr * L.hsx, hy = fr * L.hsy;
		const double dx = c.dx0 + hx, dy = c.dy0 + hy;
		const double xp = fma(dy, L.sin_ang, dx * L.cos_ang);
		const double yp = fma(dy, L.cos_ang, -dx * L.sin_ang) * L.inv_axrat;
		const double ddy = (-2 * hx * L.sin_ang + 2 * hy * L.cos_ang) * L.inv_axrat;
		const bool recurse = res > 1u && lv - 1 < (int)cm_maxrec(meta);
		bool fail = false;
		const double val = refine_fp64_decision(comps[c.gcomp], xp, yp, ddy, recurse, &fail);
		if (fail) {
			const uint32_t slot = atomicAdd(&cnt->n_cells[pass], 1u);
			if (slot < qcap) qout[slot] = c;
			else overflow = 1;
			continue;
		}
		const size_t target = 2 * (size_t)c.entry + cm_conv(meta);
		if (GRAD) {
			float gt[GT_COUNT];
			const float vg = sersic_grad32(L, (float)xp, (float)yp, gt);
			double ga[GA_SLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
			grad_accumulate(ga, gent[target] * (fr * fr * L.wq) * (double)vg, gt);
			for (int k = 0; k < GA_SLOTS; k++)
				if (ga[k] != 0) atomicAdd(&gacc[(size_t)c.gcomp * GA_SLOTS + k], ga[k]);
		} else {
			const double qv = val * (fr * fr * L.wflux);
			if (qv >= 0 && qv < 9.2e18) {
				const unsigned long long f = (unsigned long long)__double2ll_rn(qv);
				if (f) atomicAdd(&acc[target], f);
			} else {
				bad++;
			}
		}
	}
	if (bad) atomicAdd(&cnt->n_bad, bad);
	if (overflow) atomicOr(&cnt->overflow, 1u);
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
