/*
 * profit/convolve.h -- convolvers of the C++ host API
 * (libprofit: include/profit/convolve.h:44-206, src/convolve.cpp:43-99,521-584).
 * Every type runs on the device through pk_convolve.
 */
#ifndef PROFIT_AMD_CONVOLVE_H
#define PROFIT_AMD_CONVOLVE_H

#include <memory>
#include <string>

#include "profit/image.h"
#include "profit/library.h"

namespace profit {

enum ConvolverType { BRUTE_OLD = 0, BRUTE, OPENCL, FFT };

enum class effort_t { ESTIMATE = 0, MEASURE, PATIENT, EXHAUSTIVE };

/* Options of create_convolver (libprofit's ConvolverCreationPreferences);
 * omp_threads and effort are accepted and have no effect on the device. */
class ConvolverCreationPreferences {
public:
	ConvolverCreationPreferences() = default;
	ConvolverCreationPreferences(Dimensions src_dims, Dimensions krn_dims, unsigned int omp_threads,
	                             GpuEnvPtr gpu_env, effort_t effort, bool reuse_krn_fft)
	    : src_dims(src_dims), krn_dims(krn_dims), omp_threads(omp_threads), gpu_env(gpu_env), effort(effort),
	      reuse_krn_fft(reuse_krn_fft)
	{
	}
	Dimensions src_dims, krn_dims;
	unsigned int omp_threads = 1;
	GpuEnvPtr gpu_env;
	effort_t effort = effort_t::ESTIMATE;
	bool reuse_krn_fft = false;
};

class Convolver {
public:
	Convolver(ConvolverType type, const ConvolverCreationPreferences &prefs);
	virtual ~Convolver() = default;

	/* Convolver::convolve(src, krn, mask, crop, offset_out): the result has
	 * src's dimensions if crop, the convolver's padded ones otherwise, with
	 * the position of src inside it in offset_out */
	Image convolve(const Image &src, const Image &krn, const Mask &mask, bool crop = true,
	               Point &offset_out = NO_OFFSET);
	/* padding (lower-left, upper-right) added around src by this convolver */
	PointPair padding(const Dimensions &src_dims, const Dimensions &krn_dims) const;

	ConvolverType type() const { return type_; }
	bool reuse_krn_fft() const { return prefs_.reuse_krn_fft; }
	static Point NO_OFFSET;

private:
	ConvolverType type_;
	ConvolverCreationPreferences prefs_;
};

typedef std::shared_ptr<Convolver> ConvolverPtr;

ConvolverPtr create_convolver(const ConvolverType type,
                              const ConvolverCreationPreferences &prefs = ConvolverCreationPreferences());
/* by name: "brute-old", "brute", "opencl", "fft" */
ConvolverPtr create_convolver(const std::string &type,
                              const ConvolverCreationPreferences &prefs = ConvolverCreationPreferences());

}  // namespace profit

#endif
