/*
 * profit/model.h -- profit::Model of the C++ host API
 * (libprofit: include/profit/model.h:60-408, src/model.cpp): a set of
 * profiles drawn on an image, optionally convolved with a PSF, evaluated on
 * the device through pk_evaluate.
 */
#ifndef PROFIT_AMD_MODEL_H
#define PROFIT_AMD_MODEL_H

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "profit/convolve.h"
#include "profit/exceptions.h"
#include "profit/image.h"
#include "profit/library.h"
#include "profit/profile.h"

namespace profit {

class Model {
public:
	Model(unsigned int width = 0, unsigned int height = 0);
	explicit Model(Dimensions dimensions);

	/* "sersic", "moffat", "ferrer" ("ferrers"), "coresersic", "king",
	 * "brokenexp", "sky", "psf", "null"; invalid_parameter otherwise */
	ProfilePtr add_profile(const std::string &profile_name);
	bool has_profiles() const { return !profiles.empty(); }

	/* Model::evaluate(offset_out) (src/model.cpp:259-320) */
	Image evaluate(Point &offset_out = NO_OFFSET);
	std::map<std::string, std::shared_ptr<ProfileStats>> get_stats() const;

	void set_dimensions(const Dimensions &d) { dimensions = d; }
	Dimensions get_dimensions() const { return dimensions; }
	void set_finesampling(unsigned int fs)
	{
		if (fs == 0) throw invalid_parameter("finesampling must be > 0");
		finesampling = fs;
	}
	unsigned int get_finesampling() const { return finesampling; }
	/* the PSF is normalised to unit sum (include/profit/model.h:148-158) */
	void set_psf(const Image &p) { psf = p.normalize(); }
	void set_psf(Image &&p) { psf = std::move(p); psf.normalize(); }
	const Image &get_psf() const { return psf; }
	void set_image_pixel_scale(const PixelScale &s) { scale = s; }
	PixelScale get_image_pixel_scale() const { return scale; }
	void set_psf_pixel_scale(const PixelScale &s) { psf_scale = s; }
	PixelScale get_psf_pixel_scale() const { return psf_scale; }
	void set_magzero(double m) { magzero = m; }
	double get_magzero() const { return magzero; }
	void set_mask(const Mask &m) { mask = m; }
	void set_mask(Mask &&m) { mask = std::move(m); }
	const Mask &get_mask() const { return mask; }
	void set_adjust_mask(bool v) { adjust_mask = v; }
	void set_convolver(ConvolverPtr c) { convolver = std::move(c); }
	ConvolverPtr get_convolver() const { return convolver; }
	void set_crop(bool v) { crop = v; }
	void set_dry_run(bool v) { dry_run = v; }
	void set_return_finesampled(bool v) { return_finesampled = v; }
	/* accepted and unused: the evaluation runs on the device */
	void set_omp_threads(unsigned int n) { omp_threads = n; }
	unsigned int get_omp_threads() { return omp_threads; }
	/* device and precision of the evaluation (default: device 0, FP64) */
	void set_gpu_env(const GpuEnvPtr &env) { gpu_env = env; }
	GpuEnvPtr get_gpu_env() const { return gpu_env; }
	/* the same through libprofit's accelerator handle
	 * (include/profit/model.h:305-311); invalid_parameter for an
	 * environment that is not a GpuEnv */
	void set_opencl_env(const OpenCLEnvPtr &env);
	OpenCLEnvPtr get_opencl_env() const { return gpu_env; }
	/* Sub-pixel integrations of the last evaluation per profile name, by
	 * recursion level (libprofit's Model::get_profile_integrations,
	 * include/profit/model.h:116, a PROFIT_DEBUG build's): level 0, the
	 * pixels whose integration started (subsample_pixel calls at level 0);
	 * deeper levels are not counted on the device. */
	std::map<std::string, std::map<int, int>> get_profile_integrations() const;

	/* Model::adjust: the calculation mask a Model derives from `mask` to
	 * keep the flux of the convolution (src/model.cpp:230-257) */
	static void adjust(Mask &mask, const Dimensions &dims, const Image &psf, unsigned int finesampling = 1);

	static Point NO_OFFSET;

	/* the C-ABI description of this model; `comps` receives the components
	 * the returned pk_model points to */
	pk_model to_pk_model(std::vector<pk_component> &comps, std::vector<unsigned char> &mask_bytes) const;

private:
	Dimensions dimensions;
	unsigned int finesampling = 1;
	Image psf;
	PixelScale scale{1, 1}, psf_scale{1, 1};
	double magzero = 0;
	Mask mask;
	bool adjust_mask = true, crop = true, dry_run = false, return_finesampled = true;
	unsigned int omp_threads = 0;
	ConvolverPtr convolver;
	GpuEnvPtr gpu_env;
	std::vector<ProfilePtr> profiles;
	std::map<std::string, std::map<int, int>> integrations;
};

}  // namespace profit

#endif
