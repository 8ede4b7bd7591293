// Strip-form convolution launches of one part of the PSF column steps
// (compiled with -DPK_STRIP_PART=p for p = 0 .. ST_PARTS - 1).
// the strip-form kernels without the other kernels of pk_device.h (they are
// compiled in pk_api.hip)
#define PK_STRIP_KERNELS_ONLY
#include "pk_device.h"
#include "pk_strips.h"

#ifndef PK_STRIP_PART
#error "PK_STRIP_PART not set"
#endif

namespace pk {

static_assert(ST_PARTS * ST_NJ_PER_PART * 4 >= ST_KMAX, "strip parts cover every PSF width");

// classes (NJ, C, TM) of this part in turn: TM = 0 .. ST_TAIL where a tail
// fits (16 C + TM <= ST_KMAX), C = 1 .. 3, NJ over the part
template <int NJ, int C, int TM>
static bool launch_strip_class(hipStream_t s, uint32_t nj, uint32_t c, uint32_t tm, const StripLaunch &a)
{
	if (nj == (uint32_t)NJ && c == (uint32_t)C && tm == (uint32_t)TM) {
		hipLaunchKernelGGL((conv_strip_kernel<NJ, C, TM>), dim3(a.count), dim3(256), 0, s, a.jobs, a.njobs, a.src,
		                   a.kt, a.masks, a.dst, a.tile_job, a.out_all, a.img_nc, a.start, a.comps);
		return true;
	}
	constexpr int NJ_END = (PK_STRIP_PART + 1) * ST_NJ_PER_PART;
	if constexpr (TM < ST_TAIL && 16 * C + TM + 1 <= ST_KMAX) return launch_strip_class<NJ, C, TM + 1>(s, nj, c, tm, a);
	else if constexpr (C < (ST_KMAX + 15) / 16) return launch_strip_class<NJ, C + 1, 0>(s, nj, c, tm, a);
	else if constexpr (NJ < NJ_END && NJ < (ST_KMAX + 3) / 4) return launch_strip_class<NJ + 1, 1, 0>(s, nj, c, tm, a);
	else return false;
}

template <>
bool launch_strips_part<PK_STRIP_PART>(hipStream_t s, uint32_t nj, uint32_t c, uint32_t tm, const StripLaunch &a)
{
	return launch_strip_class<PK_STRIP_PART * ST_NJ_PER_PART + 1, 1, 0>(s, nj, c, tm, a);
}

}  // namespace pk
