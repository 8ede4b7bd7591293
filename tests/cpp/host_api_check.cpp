// Checks of the C++ host API (include/profit/*.h), in the manner of
// libprofit's own CxxTest suites (tests/test_profile.h, tests/test_model.h,
// tests/test_image.h, tests/test_sky.h, tests/test_psf.h).
//   host_api_check cpu          host-only checks (no device needed)
//   host_api_check gpu <dir>    evaluations on device 0; writes <dir>/<case>.bin
//                               (uint32 width, height, int32 offset x, y, then
//                               width*height doubles) for comparison with the oracle
// Prints one line per check and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iostream>
#include <string>

#include "profit/fits.h"
#include "profit/profit.h"

using namespace profit;

static int n_checks = 0;

#define CHECK(cond)                                                                        \
	do {                                                                                   \
		n_checks++;                                                                        \
		if (!(cond)) {                                                                     \
			std::cerr << "FAILED " << __FILE__ << ":" << __LINE__ << ": " #cond << std::endl; \
			std::exit(1);                                                                  \
		}                                                                                  \
	} while (0)

template <typename E>
static bool throws(const std::function<void()> &f, const std::string &contains = "")
{
	try {
		f();
	} catch (const E &e) {
		return contains.empty() || std::string(e.what()).find(contains) != std::string::npos;
	} catch (...) {
		return false;
	}
	return false;
}

static void cpu_checks()
{
	// ---- parameters (tests/test_profile.h) ----
	Model m(10, 10);
	auto s = m.add_profile("sersic");
	s->parameter("xcen", 2.0);
	s->parameter("box=0.5");
	s->parameter("resolution", 4u);
	s->parameter("rough", true);
	s->parameter("max_recursions=3");
	CHECK(s->get_component().xcen == 2.0);
	CHECK(s->get_component().box == 0.5);
	CHECK(s->get_component().resolution == 4u);
	CHECK(s->get_component().rough == 1);
	CHECK(s->get_component().max_recursions == 3u);
	CHECK(throws<invalid_parameter>([&] { s->parameter("xcen", true); }, "Unknown bool parameter"));
	CHECK(throws<invalid_parameter>([&] { s->parameter("rough", 1.0); }, "Unknown double parameter"));
	CHECK(throws<invalid_parameter>([&] { s->parameter("nser", 2u); }, "Unknown unsigned int parameter"));
	CHECK(throws<unknown_parameter>([&] { s->parameter("foo=1"); }, "doesn't support parameter foo"));
	CHECK(throws<invalid_parameter>([&] { s->parameter("xcen=abc"); }, "cannot be parsed as double"));
	CHECK(throws<invalid_parameter>([&] { s->parameter("xcen"); }, "missing = in parameter"));
	CHECK(throws<invalid_parameter>([&] { m.add_profile("unknown"); }, "Unknown profile name"));
	CHECK(m.add_profile("ferrers")->get_kind() == PK_PROFILE_FERRER);
	auto sky = Model(2, 2).add_profile("sky");
	sky->parameter("bg=3");
	CHECK(sky->get_component().p[PK_SKY_BG] == 3.0);
	CHECK(throws<unknown_parameter>([&] { sky->parameter("xcen=1"); }));
	std::cout << "ok parameters" << std::endl;

	// ---- model validation (tests/test_model.h:35-160), host only ----
	{
		Model z;
		z.add_profile("null");
		CHECK(throws<invalid_parameter>([&] { z.evaluate(); }, "requested dimensions are 0"));
		Model sc(10, 10);
		sc.add_profile("null");
		sc.set_image_pixel_scale({0, 1});
		CHECK(throws<invalid_parameter>([&] { sc.evaluate(); }, "scale_x"));
		Model mk(10, 10);
		mk.add_profile("null");
		mk.set_mask(Mask(true, 9, 10));
		CHECK(throws<invalid_parameter>([&] { mk.evaluate(); }, "Mask dimensions"));
		Model cv(10, 10);
		cv.add_profile("sersic")->parameter("convolve", true);
		CHECK(throws<invalid_parameter>([&] { cv.evaluate(); }, "No psf provided"));
		Model bad(10, 10);
		bad.add_profile("sersic")->parameter("axrat", 0.0);
		CHECK(throws<invalid_parameter>([&] { bad.evaluate(); }, "axrat"));
		CHECK(throws<invalid_parameter>([&] { Model f; f.set_finesampling(0); }));
		Model dry(10, 10);
		dry.add_profile("sersic");
		dry.set_dry_run(true);
		Point off(3, 3);
		Image im = dry.evaluate(off);
		CHECK(im.getDimensions() == Dimensions(10, 10) && im.total() == 0 && off == Point(0, 0));
		Model dryp(10, 10);
		dryp.add_profile("sersic")->parameter("convolve", true);
		dryp.set_psf(Image(1.0, 3, 5));
		dryp.set_dry_run(true);
		CHECK(dryp.evaluate().getDimensions() == Dimensions(12, 14));  // drawing dims: 10 + 2 * (3/2, 5/2)
	}
	CHECK(throws<invalid_parameter>([] { create_convolver("foo"); }, "Convolver of type foo is not supported"));
	CHECK(throws<invalid_parameter>([] { create_convolver((ConvolverType)7); }, "Unsupported convolver type"));
	CHECK(create_convolver("fft")->type() == FFT);
	{
		auto fft = create_convolver(FFT);
		auto pad = fft->padding(Dimensions(100, 100), Dimensions(25, 25));
		// ext = 2 max(src, krn) = 200, offset 100/2 - 1 (tests/test_model.h:365-374)
		CHECK(pad.first == Point(49, 49) && pad.second == Point(51, 51));
		auto brute = create_convolver(BRUTE);
		CHECK(brute->padding(Dimensions(100, 100), Dimensions(25, 25)).first == Point(0, 0));
	}
	std::cout << "ok model validation" << std::endl;

	// ---- containers (tests/test_image.h) ----
	Image a(std::vector<double>{1, 2, 3, 4, 5, 6}, 3, 2);
	CHECK(a.total() == 21);
	CHECK(a[Point(2, 1)] == 6);
	Image e = a.extend(Dimensions(5, 4), Point(1, 1));
	CHECK(e.total() == 21 && e[Point(1, 1)] == 1 && e[Point(3, 2)] == 6);
	CHECK(e.crop(Dimensions(3, 2), Point(1, 1)) == a);
	CHECK(a.reverse()[0] == 6);
	Image d = Image(1.0, 4, 4).downsample(2);
	CHECK(d.getDimensions() == Dimensions(2, 2) && d[0] == 4);
	CHECK(Image(1.0, 5, 5).downsample(2).getDimensions() == Dimensions(3, 3));
	CHECK(Image(1.0, 5, 5).downsample(2, Image::AVERAGE)[8] == 1);
	Image u = a.upsample(2);
	CHECK(u.getDimensions() == Dimensions(6, 4) && u[0] == 0.25 && u.total() == 21);
	CHECK(a.upsample(2, Image::COPY).total() == 84);
	CHECK(std::fabs(static_cast<const Image &>(a).normalize().total() - 1.0) < 1e-15);
	Mask mk(false, 5, 5);
	mk[Point(2, 2)] = true;
	Mask ex = mk.expand_by(Dimensions(1, 1));
	int n = 0;
	for (bool b : ex) n += b;
	CHECK(n == 9 && ex[Point(1, 1)] && !ex[Point(0, 0)]);
	Box bb = mk.bounding_box();
	CHECK(bb.first == Point(2, 2) && bb.second == Point(3, 3));
	Image masked = Image(2.0, 5, 5) & ex;
	CHECK(masked.total() == 18);
	CHECK(throws<std::invalid_argument>([&] { a.extend(Dimensions(2, 2)); }));
	std::cout << "ok containers" << std::endl;

	// ---- Model::adjust (src/model.cpp:230-257) ----
	{
		Mask m5(false, 10, 10);
		m5[Point(5, 5)] = true;
		Model::adjust(m5, Dimensions(10, 10), Image(1.0, 3, 3), 1);
		int c = 0;
		for (bool b : m5) c += b;
		CHECK(m5.getDimensions() == Dimensions(10, 10) && c == 9);
		Mask edge(false, 10, 10);
		edge[Point(0, 0)] = true;
		Model::adjust(edge, Dimensions(10, 10), Image(1.0, 3, 3), 2);
		CHECK(edge.getDimensions() == Dimensions(22, 22));  // upsampled, padded by 1
	}
	std::cout << "ok adjust" << std::endl;

	// ---- FITS round trip ----
	{
		Image f(std::vector<double>{1.5, -2.25, 3e-300, 4e10, 0, 7}, 2, 3);
		to_fits(f, Point(1, 2), {0.5, 2.0}, "/tmp/pk_host_api_check");
		PixelScale ps;
		Image g = from_fits("/tmp/pk_host_api_check.fits", ps);
		CHECK(g == f);
		CHECK(ps.first == 0.5 && ps.second == 2.0);
		std::ifstream in("/tmp/pk_host_api_check.fits", std::ios::binary | std::ios::ate);
		CHECK(in.tellg() % 2880 == 0);
	}
	std::cout << "ok fits" << std::endl;
	CHECK(version() == "1.9.3");
}

static void dump(const std::string &dir, const std::string &name, const Image &im, Point off)
{
	std::ofstream f(dir + "/" + name + ".bin", std::ios::binary);
	uint32_t w = im.getWidth(), h = im.getHeight();
	int32_t ox = (int32_t)off.x, oy = (int32_t)off.y;
	f.write((const char *)&w, 4);
	f.write((const char *)&h, 4);
	f.write((const char *)&ox, 4);
	f.write((const char *)&oy, 4);
	f.write((const char *)im._get().data(), (std::streamsize)(8 * im.size()));
	std::cout << "wrote " << name << " " << w << "x" << h << std::endl;
}

static Image moffat_psf(unsigned int k)
{
	Model p(k, k);
	auto mf = p.add_profile("moffat");
	mf->parameter("xcen", k / 2.0);
	mf->parameter("ycen", k / 2.0);
	mf->parameter("fwhm=3");
	mf->parameter("con=2.5");
	mf->parameter("mag=0");
	return p.evaluate();
}

static void gpu_checks(const std::string &dir)
{
	CHECK(init());
	Point off;
	{
		// profit-cli defaults: one Sersic, 100x100, FP64
		Model m(100, 100);
		auto s = m.add_profile("sersic");
		for (const char *spec : {"xcen=50.3", "ycen=49.7", "mag=15", "re=10", "nser=4", "ang=30", "axrat=0.6", "box=0"})
			s->parameter(spec);
		dump(dir, "c1", m.evaluate(off), off);
		CHECK(m.get_stats()["sersic"]->refined_pixels > 0);
	}
	dump(dir, "psf15", moffat_psf(15), Point());
	for (const char *conv : {"brute", "brute-old", "fft"}) {
		for (int crop = 1; crop >= 0; crop--) {
			Model m(96, 80);
			auto s = m.add_profile("sersic");
			for (const char *spec : {"xcen=47.1", "ycen=41.4", "re=7", "nser=2", "axrat=0.6", "ang=10", "convolve=1"})
				s->parameter(spec);
			m.add_profile("sky")->parameter("bg", 1e-9);
			m.set_psf(moffat_psf(15));
			m.set_convolver(create_convolver(conv));
			m.set_crop(crop != 0);
			dump(dir, std::string("conv_") + conv + (crop ? "" : "_nocrop"), m.evaluate(off), off);
		}
	}
	{
		// masks and finesampling
		Model m(50, 40);
		auto s = m.add_profile("sersic");
		for (const char *spec : {"xcen=20.2", "ycen=22.7", "re=6", "nser=3", "axrat=0.5", "ang=70"}) s->parameter(spec);
		Mask mk(true, 50, 40);
		for (unsigned int j = 0; j < 40; j += 3)
			for (unsigned int i = 0; i < 50; i += 2) mk[Point(i, j)] = false;
		m.set_mask(mk);
		m.set_finesampling(2);
		m.set_return_finesampled(false);
		dump(dir, "mask_fs2", m.evaluate(off), off);
	}
	{
		// direct use of a convolver
		Image src(std::vector<double>(37 * 41), 37, 41), krn(std::vector<double>(9 * 6), 9, 6);
		for (unsigned int i = 0; i < src.size(); i++) src[i] = std::fmod(i * 0.37, 1.0);
		for (unsigned int i = 0; i < krn.size(); i++) krn[i] = std::fmod(i * 0.61, 1.0);
		dump(dir, "convolver_brute", create_convolver(BRUTE)->convolve(src, krn, Mask()), Point());
		Point o;
		dump(dir, "convolver_fft_nocrop", create_convolver(FFT)->convolve(src, krn, Mask(), false, o), o);
	}
	{
		// FP32 context
		Model m(100, 100);
		auto s = m.add_profile("sersic");
		for (const char *spec : {"xcen=50.3", "ycen=49.7", "mag=15", "re=10", "nser=4", "ang=30", "axrat=0.6"})
			s->parameter(spec);
		m.set_gpu_env(get_gpu_environment(0, Precision::FP32));
		dump(dir, "c1_fp32", m.evaluate(off), off);
	}
	finish();
}

int main(int argc, char *argv[])
{
	const std::string mode = argc > 1 ? argv[1] : "cpu";
	try {
		if (mode == "cpu") cpu_checks();
		else if (mode == "gpu" && argc > 2) gpu_checks(argv[2]);
		else {
			std::cerr << "usage: host_api_check cpu | gpu <dir>" << std::endl;
			return 2;
		}
	} catch (const std::exception &e) {
		std::cerr << "FAILED with exception: " << e.what() << std::endl;
		return 1;
	}
	std::cout << "all " << n_checks << " checks passed" << std::endl;
	return 0;
}
