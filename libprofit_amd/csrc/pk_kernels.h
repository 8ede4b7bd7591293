// Device work descriptors shared by the host orchestration and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pk_profiles.h"

namespace pk {

constexpr uint64_t NONE = ~0ull;

// One model image to produce (a libprofit Model::evaluate call).
struct PkJob {
	uint32_t dw, dh;             // drawing dimensions (src/model.cpp:183)
	uint32_t comp_off, ncomp;    // radial / sky components in the component table
	uint64_t img_off;            // offset of the drawing image in the image buffers
	uint64_t mask_off;           // drawing-dims calculation mask (NONE = no mask)
	double sx, sy;               // pixel scale handed to the profiles (scale / finesampling)
	uint32_t tile_off, tiles_x;  // coarse tiles of this job (prefix over jobs)
	uint32_t has_conv, has_nc;   // any convolved / non-convolved component
	// convolution (BRUTE family): output = drawing dims
	uint64_t psf_off;            // normalised PSF in the PSF buffer
	uint32_t kw, kh;
	uint32_t ctile_off, ctiles_x;  // convolution tiles (prefix over jobs)
	uint32_t stile_off;            // strip tiles of the matrix-core strip form (ctiles_x per tile row)
	uint32_t st_rows;              // output rows of its strip tiles (the image split evenly, <= ST_H)
	uint32_t nc_sky;               // FP32 batch job whose unconvolved components are all skies, without
	                               // mask, whose convolution writes the returned image: the unconvolved
	                               // image is one value, taken from the components, and never stored
	uint64_t conv_off;           // convolved image (cw x ch) in the conv buffer
	uint32_t cw, ch;             // dims of the convolved image
	int32_t csx, csy;            // convolved(X,Y) = same-size convolution at (X+csx, Y+csy)
	int32_t cmx, cmy;            // calculation mask pixel of convolved (X,Y) is (X-cmx, Y-cmy)
	int32_t mox, moy;            // position of the drawing image inside the convolved frame
	uint32_t wrap_y;             // FFT convolver: rows of its circular buffer (0: no wrap)
	uint32_t fuse_out;           // batches: the convolution writes the returned image itself
	                             // (convolved frame = returned image, + the unconvolved part)
	// composition of the returned image (crop, SUM downsample, final mask)
	int32_t cox, coy;            // crop origin in the convolved frame
	uint32_t crw, crh;           // crop dims
	uint32_t fs;                 // downsampling factor (1: none)
	uint32_t ow, oh;             // returned dims
	uint64_t out_off;            // offset of the returned image in the output buffer
	uint64_t omask_off;          // user's mask applied at the end (NONE = none)
	uint32_t omask_w, omask_h, omask_f;  // its dims and upsampling factor
	uint32_t otile_off, otiles_x;  // compose tiles (prefix over jobs)
	// FP32-mode refinement accumulates pixel contributions as exact 64-bit
	// fixed-point integers: value * fx_scale, fx_scale = 2^60 / (sum of the
	// total fluxes of the job's components), so the order of the additions
	// cannot change the result
	double fx_scale;
};

struct RefineEntry {
	uint32_t job, pix, cmask;
};

// The FP32-mode refinement runs in at most R32_PASSES launches (a launch
// setting, R32Tune::passes): one per level of the sub-pixel recursion (the
// failing sub-cells of a level are the next launch's work, spread evenly over
// its wavefronts), the last one refining all deeper levels.
constexpr int R32_PASSES = 7;

// Run-time settings of the FP32 refinement launches (chosen per batch by the
// autotuner; any values >= 1 give the same result, only the speed changes).
struct R32Tune {
	uint32_t handover_div;  // hand over once 1/handover_div of a pass's wavefronts stopped
	uint32_t chunk;         // pixels per claim at most (pass 0)
	uint32_t chunk_cells;   // handed-over cells per claim at most (later passes)
	uint32_t passes;        // launches of the refinement (R32_PASSES_MIN .. R32_PASSES)
	uint32_t spread_log2;   // pass 0 claims runs of 2^spread_log2 entries from 64 places of the entry
	                        // list in turn (0: in entry order)
};

struct PkCounters {
	uint32_t n_entries;    // refinement entries appended by the coarse pass
	uint32_t next_entry;   // work-queue head of the FP64 refinement pass
	uint32_t overflow;     // set when a refinement work queue exceeded its capacity
	uint32_t n_bad;        // non-finite contributions (dropped)
	unsigned long long n_subcells;  // profile evaluations made by the refinement
	uint32_t claim[R32_PASSES];    // work-queue head of FP32 refinement pass p
	uint32_t n_cells[R32_PASSES];  // cells pass p hands to pass p + 1
	uint32_t n_pend[R32_PASSES];   // guard-band sub-cells pass p listed for the FP64 decision
};

// A cell waiting to be subdivided (FP32-mode refinement), 32 bytes.  Its size
// in pixels, resolution^-level, follows from `meta` (cell_size_fp64).
struct alignas(16) RCell {
	double dx0, dy0;       // lower corner relative to the component centre
	uint32_t gcomp;        // index in the component table
	uint32_t entry;        // pixel accumulator
	uint32_t meta;         // level | resolution << 8 | max recursions << 16 | convolved << 24
};

// PSF-profile work item (src/psf.cpp:58-145), geometry resolved on the host.
struct PkPsfItem {
	uint32_t job;
	uint32_t to_conv;
	uint32_t x0, y0, x1, y1;     // image pixel range (inclusive)
	double origin_x, origin_y, scale;
	double psx, psy;
	uint32_t pw, ph;
	uint64_t psf_off;
};

}  // namespace pk
