// Launches of the matrix-core strip convolution (conv_strip_kernel), one per
// class of PSF column steps NJ = ceil(kw / 4), row blocks C and tail rows TM.
// The kernels are instantiated in pk_strips.hip, which is compiled once per
// part of the NJ range (PK_STRIP_PART) so that the parts build in parallel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pk_kernels.h"

namespace pk {

constexpr int ST_NJ_PER_PART = 2;  // PSF column steps per part
constexpr int ST_PARTS = 6;        // parts: NJ = 1 .. ST_PARTS * ST_NJ_PER_PART (PSF sides <= 48)

// Arguments of one strip-form launch: the jobs of one class.
struct StripLaunch {
	const PkJob *jobs;
	uint32_t njobs, count, start;
	const float *src, *kt;
	const uint8_t *masks;
	float *dst;
	const uint32_t *tile_job;
	float *out_all;
	const float *img_nc;
	const PkComp *comps;
};

// The launches of part P (PSF column steps P * ST_NJ_PER_PART + 1 ...); false
// where (nj, c, tm) is no class of that part.
template <int P>
bool launch_strips_part(hipStream_t s, uint32_t nj, uint32_t c, uint32_t tm, const StripLaunch &a);
template <> bool launch_strips_part<0>(hipStream_t, uint32_t, uint32_t, uint32_t, const StripLaunch &);
template <> bool launch_strips_part<1>(hipStream_t, uint32_t, uint32_t, uint32_t, const StripLaunch &);
template <> bool launch_strips_part<2>(hipStream_t, uint32_t, uint32_t, uint32_t, const StripLaunch &);
template <> bool launch_strips_part<3>(hipStream_t, uint32_t, uint32_t, uint32_t, const StripLaunch &);
template <> bool launch_strips_part<4>(hipStream_t, uint32_t, uint32_t, uint32_t, const StripLaunch &);
template <> bool launch_strips_part<5>(hipStream_t, uint32_t, uint32_t, uint32_t, const StripLaunch &);

inline bool launch_strips(hipStream_t s, uint32_t nj, uint32_t c, uint32_t tm, const StripLaunch &a)
{
	switch (nj ? (nj - 1) / ST_NJ_PER_PART : ST_PARTS) {
	case 0: return launch_strips_part<0>(s, nj, c, tm, a);
	case 1: return launch_strips_part<1>(s, nj, c, tm, a);
	case 2: return launch_strips_part<2>(s, nj, c, tm, a);
	case 3: return launch_strips_part<3>(s, nj, c, tm, a);
	case 4: return launch_strips_part<4>(s, nj, c, tm, a);
	case 5: return launch_strips_part<5>(s, nj, c, tm, a);
	default: return false;
	}
}

}  // namespace pk
