// A caller written only against libprofit's own public API (names and
// signatures of the reference's include/profit/*.h), in the manner of
// profit-cli (src/profit-cli.cpp:333-412: the statistics printout and the
// timed evaluation loop): it must compile unchanged against this library's
// headers.  It evaluates a Sersic + Moffat model convolved with a PSF,
// prints the statistics and writes the image to <out>
// (uint32 width, height, int32 offset x, y, then width*height doubles).
//   ref_api_check <out> [iterations]
#include <chrono>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include <profit/profit.h>

using namespace profit;

static void print_stats_line(std::ostream &os, const std::string &prefix, const std::string &name, nsecs_t t)
{
	os << prefix << std::setw(40) << std::left << name << ": " << t / 1e6 << " [ms]" << std::endl;
}

static void print_cl_stats(std::ostream &os, const std::string &prefix, const OpenCL_times &stats)
{
	std::ostringstream ops;
	ops << "OpenCL operations (" << stats.nwork_items << " work items)";
	print_stats_line(os, prefix, ops.str(), stats.total);
	print_stats_line(os, prefix + "  ", "Kernel preparation", stats.kernel_prep);
	print_stats_line(os, prefix + "  ", "Kernel execution", stats.kernel_times.exec);
}

static void print_stats(std::ostream &os, const Model &m)
{
	for (const auto &profile_integrations : m.get_profile_integrations()) {
		int total = 0;
		os << "Integrations per recursion level for profile " << std::get<0>(profile_integrations) << std::endl;
		for (const auto level_integrations : std::get<1>(profile_integrations)) {
			total += std::get<1>(level_integrations);
			os << " Level " << std::get<0>(level_integrations) << ": " << std::get<1>(level_integrations)
			   << " integrations" << std::endl;
		}
		os << " Total: " << total << " integrations" << std::endl;
	}
	for (auto const &stat_pair : m.get_stats()) {
		ProfileStats *profile_stats = std::get<1>(stat_pair).get();
		if (!profile_stats) continue;
		os << "Stats for profile " << std::get<0>(stat_pair) << std::endl;
		auto *rprofile_stats = dynamic_cast<RadialProfileStats *>(profile_stats);
		auto opencl_env = m.get_opencl_env();
		if (rprofile_stats && opencl_env) {
			os << "  environment " << opencl_env->get_platform_name() << " / " << opencl_env->get_device_name()
			   << " (OpenCL version " << opencl_env->get_version() << ")" << std::endl;
			print_cl_stats(os, "", rprofile_stats->cl_times);
			print_stats_line(os, "", "Subsampling loop", rprofile_stats->subsampling.total);
			print_stats_line(os, "", "Final image", rprofile_stats->final_image);
		}
		print_stats_line(os, "", "Total", profile_stats->total);
	}
}

int main(int argc, char **argv)
{
	if (argc < 2) {
		std::cerr << "usage: ref_api_check <out> [iterations]" << std::endl;
		return 2;
	}
	const unsigned iterations = argc > 2 ? (unsigned)std::stoul(argv[2]) : 3;
	if (!init()) {
		std::cerr << "init failed: " << init_diagnose() << std::endl;
		return 1;
	}
	std::cout << "libprofit " << version() << "; OpenCL " << has_opencl() << " (" << opencl_version_major() << "."
	          << opencl_version_minor() << "); SIMD AUTO " << has_simd_instruction_set(AUTO) << ", AVX "
	          << has_simd_instruction_set(AVX) << std::endl;
	for (const auto &plat : get_opencl_info())
		for (const auto &dev : std::get<1>(plat).dev_info)
			std::cout << "platform " << std::get<1>(plat).name << " device " << std::get<0>(dev) << ": "
			          << std::get<1>(dev).name << std::endl;

	Model m(80, 70);
	auto sersic = m.add_profile("sersic");
	for (const char *p : {"xcen=40.3", "ycen=34.7", "mag=15", "re=6", "nser=2.5", "ang=30", "axrat=0.6", "box=0.1",
	                      "convolve=1"})
		sersic->parameter(std::string(p));
	auto moffat = m.add_profile("moffat");
	moffat->parameter("xcen", 30.2);
	moffat->parameter("ycen", 41.6);
	moffat->parameter("mag", 16.0);
	moffat->parameter("fwhm", 4.0);
	moffat->parameter("con", 3.0);
	moffat->parameter("convolve", true);
	auto sky = m.add_profile("sky");
	sky->parameter("bg", 1e-9);

	// a 15 x 15 Moffat PSF drawn by a Model of its own
	Model psf_model(15, 15);
	auto psf_profile = psf_model.add_profile("moffat");
	psf_profile->parameter("xcen", 7.5);
	psf_profile->parameter("ycen", 7.5);
	psf_profile->parameter("mag", 0.0);
	psf_profile->parameter("fwhm", 3.0);
	psf_profile->parameter("con", 2.5);
	m.set_psf(psf_model.evaluate());

	OpenCLEnvPtr env = get_opencl_environment(0, 0, true, false);
	m.set_opencl_env(env);
	ConvolverCreationPreferences prefs(m.get_dimensions(), Dimensions(15, 15), 1, env, effort_t::ESTIMATE, false,
	                                   simd_instruction_set::AUTO);
	m.set_convolver(create_convolver(ConvolverType::BRUTE, prefs));

	Point offset;
	Image result;
	auto start = std::chrono::system_clock::now();
	for (unsigned i = 0; i != iterations; i++) result = m.evaluate(offset);
	auto end = std::chrono::system_clock::now();
	auto duration = std::chrono::duration_cast<std::chrono::milliseconds>(end - start).count();
	std::cout << "Ran " << iterations << " iterations in " << std::fixed << std::setprecision(3)
	          << (double)duration / 1000 << " [s] (" << (double)duration / iterations << " [ms] per iteration)"
	          << std::endl;
	print_stats(std::cout, m);

	std::ofstream f(argv[1], std::ios::binary);
	const uint32_t w = result.getWidth(), h = result.getHeight();
	const int32_t ox = (int32_t)offset.x, oy = (int32_t)offset.y;
	f.write((const char *)&w, 4);
	f.write((const char *)&h, 4);
	f.write((const char *)&ox, 4);
	f.write((const char *)&oy, 4);
	for (uint32_t j = 0; j < h; j++)
		for (uint32_t i = 0; i < w; i++) {
			const double v = result[Point(i, j)];
			f.write((const char *)&v, 8);
		}
	finish();
	return 0;
}
