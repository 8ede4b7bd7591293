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

/* Options of create_convolver (libprofit's ConvolverCreationPreferences,
 * include/profit/convolve.h:133-182); omp_threads, effort and
 * instruction_set are accepted and have no effect on the device.
 * opencl_env selects the device and precision (a GpuEnv; default: device 0,
 * FP64). */
class ConvolverCreationPreferences {
public:
	ConvolverCreationPreferences() = default;
	ConvolverCreationPreferences(Dimensions src_dims, Dimensions krn_dims, unsigned int omp_threads,
	                             OpenCLEnvPtr opencl_env, effort_t effort, bool reuse_krn_fft,
	                             simd_instruction_set instruction_set = AUTO)
	    : src_dims(src_dims), krn_dims(krn_dims), omp_threads(omp_threads), opencl_env(opencl_env), effort(effort),
	      reuse_krn_fft(reuse_krn_fft), instruction_set(instruction_set)
	{
	}
	Dimensions src_dims, krn_dims;
	unsigned int omp_threads = 1;
	OpenCLEnvPtr opencl_env;
	effort_t effort = effort_t::ESTIMATE;
	bool reuse_krn_fft = false;
	simd_instruction_set instruction_set = AUTO;
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
