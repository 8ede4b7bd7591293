// profit-cli: generates an image from a model and a set of profiles on the
// device, with the options and output of libprofit's command-line tool
// (src/profit-cli.cpp:140-700): -p [count,]name[:k=v:...] profiles,
// -P w:h[:sx:sy]:v1,v2,... or a FITS file as PSF, -T convolver, -i N timing
// loop printing "Ran N iterations in X [s] (Y [ms] per iteration)", -t text
// or -f FITS output, -s statistics.  -C p,d,D selects device d and FP64
// (D = 1) or FP32 (D = 0) arithmetic, in the place of the OpenCL platform,
// device and double-support triple.
#include <unistd.h>

#include <chrono>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "profit/fits.h"
#include "profit/profit.h"

using namespace profit;

namespace {

class invalid_cmdline : public exception {
public:
	explicit invalid_cmdline(const std::string &what) : exception(what) {}
};

std::vector<std::string> split(const std::string &s, const std::string &sep)
{
	std::vector<std::string> out;
	size_t b = 0;
	for (;;) {
		size_t e = s.find(sep, b);
		out.push_back(s.substr(b, e == std::string::npos ? std::string::npos : e - b));
		if (e == std::string::npos) break;
		b = e + sep.size();
	}
	return out;
}

unsigned int stoui(const std::string &s)
{
	unsigned long v = std::stoul(s);
	if (v > 0xffffffffUL) throw std::out_of_range(s);
	return (unsigned int)v;
}

const char *help_msg = R"(
%s: generates an image out of a model and a set of profiles on an MI355X GPU

Usage: %s [options] -p <spec> [-p <spec> ...]

Options:
  -t        Output image as text values on stdout
  -f <file> Output image as fits file
  -i <n>    Output performance information after evaluating the model n times
  -s        Show runtime stats
  -T <conv> Use this type of convolver (brute, brute-old, opencl, fft)
  -u        Return an un-cropped image from the convolver
  -C <p,d,D> Use GPU device d, in FP64 (D=1, default) or FP32 (D=0) arithmetic
  -c        Display information about the GPU devices
  -n <n>    Accepted for compatibility (the evaluation runs on the GPU)
  -e <n>    Accepted for compatibility (FFT plan effort)
  -I <n>    Accepted for compatibility (SIMD instruction set)
  -r        Reuse FFT-transformed PSF across evaluations (if -T fft)
  -x        Image pixel scale in x. Defaults to 1
  -y        Image pixel scale in y. Defaults to 1
  -S <n>    Finesampling factor. Defaults to 1
  -F        Do *not* return finesampled image (if -S <n>)
  -w        Width in pixels. Defaults to 100
  -H        Height in pixels. Defaults to 100
  -m        Zero magnitude. Defaults to 0
  -P        PSF function (specified as w:h:val1,val2..., or as a FITS filename)
  -R        Clear the library's cache and exit
  -h,-?     Show this help and exit
  -V        Show the program version and exit

Profiles are specified as -p [count,]name[:param1=val1:param2=val2:...] with
name one of null, psf, sky, sersic, moffat, ferrer, coresersic, brokenexp, king.
)";

void usage(std::ostream &os, const char *prog)
{
	std::vector<char> buf(std::strlen(help_msg) + 2 * std::strlen(prog) + 1);
	std::snprintf(buf.data(), buf.size(), help_msg, prog, prog);
	os << buf.data();
}

void show_version(std::ostream &os)
{
	os << "libprofit version " << version() << std::endl;
	os << "GPU support: " << (has_gpu() ? "Yes (HIP, gfx950)" : "No") << std::endl;
	os << "OpenCL support: No" << std::endl;
	os << "OpenMP support: No" << std::endl;
	os << "FFTW support: No" << std::endl;
}

void parse_profile(std::ostream &os, Model &model, const std::string &description)
{
	// src/profit-cli.cpp:63-94
	std::string desc = description;
	if (desc.empty()) throw invalid_cmdline("Missing parameter name after -p");
	unsigned int count = 1;
	auto parts = split(desc, ",");
	if (parts.size() > 1) {
		count = (unsigned int)std::stod(parts[0]);
		parts = split(parts[1], ":");
	} else {
		parts = split(desc, ":");
	}
	const std::string name = parts[0];
	for (unsigned int i = 0; i < count; i++) {
		auto p = model.add_profile(name);
		for (size_t k = 1; k < parts.size(); k++) {
			try {
				p->parameter(parts[k]);
			} catch (const unknown_parameter &e) {
				os << e.what();
			}
		}
	}
}

Image parse_psf(const std::string &arg, Model &m)
{
	// w:h[:scale_x:scale_y]:v1,v2,...  (src/profit-cli.cpp:96-137)
	auto tok = split(arg, ":");
	if (tok.size() < 2) throw invalid_cmdline(tok.empty() ? "Missing psf's width" : "Missing psf's height");
	if (tok.size() != 3 && tok.size() != 5) throw invalid_cmdline("Invalid psf format, see -h for help");
	const unsigned int w = stoui(tok[0]), h = stoui(tok[1]);
	if (tok.size() == 5) m.set_psf_pixel_scale({std::stod(tok[2]), std::stod(tok[3])});
	auto vals = split(tok.back(), ",");
	Image psf(w, h);
	if (vals.size() != psf.size()) {
		std::ostringstream os;
		os << "Not enough values provided for PSF. Provided: " << vals.size() << ", expected: " << psf.size();
		throw invalid_cmdline(os.str());
	}
	for (size_t i = 0; i < vals.size(); i++) psf[i] = std::stod(vals[i]);
	return psf;
}

bool file_exists(const std::string &f) { return access(f.c_str(), R_OK) == 0; }

Image run(std::ostream &os, unsigned int iterations, Model &m, Point &offset)
{
	// src/profit-cli.cpp:389-412
	using clock = std::chrono::system_clock;
	Image result;
	auto start = clock::now();
	for (unsigned int i = 0; i != iterations; i++) result = m.evaluate(offset);
	auto end = clock::now();
	const double ms = std::chrono::duration_cast<std::chrono::microseconds>(end - start).count() / 1e3;
	os << std::fixed << std::setprecision(3) << "Ran " << iterations << " iterations in " << ms / 1e3 << " [s] ("
	   << ms / iterations << " [ms] per iteration)" << std::endl;
	return result;
}

int parse_and_run(int argc, char *argv[], std::ostream &cout, std::ostream &cerr)
{
	unsigned int width = 100, height = 100, iterations = 1, finesampling = 1;
	double scale_x = 1, scale_y = 1;
	enum { none, text, fits } output = none;
	std::string fits_output, convolver_type = "brute";
	ConvolverCreationPreferences prefs;
	bool show_stats = false;
	int device = 0;
	Precision precision = Precision::FP64;
	Model m;
	int opt;
	while ((opt = getopt(argc, argv, "h?VsRP:p:w:H:x:y:X:Y:m:tf:i:T:uS:C:ce:rn:FI:")) != -1) {
		switch (opt) {
		case 'h':
		case '?': usage(cout, argv[0]); return 0;
		case 'V': show_version(cout); return 0;
		case 'R': clear_cache(); return 0;
		case 's': show_stats = true; break;
		case 'P': {
			Image psf;
			if (file_exists(optarg)) {
				PixelScale ps;
				try {
					psf = from_fits(optarg, ps);
				} catch (const invalid_file &e) {
					throw invalid_cmdline(std::string("Error loading PSF from ") + optarg + ": " + e.what());
				}
				m.set_psf_pixel_scale(ps);
			} else {
				psf = parse_psf(optarg, m);
			}
			prefs.krn_dims = psf.getDimensions();
			m.set_psf(std::move(psf));
			break;
		}
		case 'p': parse_profile(cerr, m, optarg); break;
		case 'w': width = stoui(optarg); break;
		case 'H': height = stoui(optarg); break;
		case 'x': scale_x = std::stod(optarg); break;
		case 'y': scale_y = std::stod(optarg); break;
		case 'X':
		case 'Y': break;
		case 'm': m.set_magzero(std::stod(optarg)); break;
		case 't': output = text; break;
		case 'f':
			fits_output = optarg;
			output = fits;
			break;
		case 'i': iterations = stoui(optarg); break;
		case 'T': convolver_type = optarg; break;
		case 'u': m.set_crop(false); break;
		case 'S': finesampling = stoui(optarg); break;
		case 'F': m.set_return_finesampled(false); break;
		case 'C': {
			auto t = split(optarg, ",");
			if (t.size() != 3) throw invalid_cmdline("-C argument must be of the form 'p,d,D' (e.g., -C 0,0,1)");
			device = (int)stoui(t[1]);
			precision = stoui(t[2]) ? Precision::FP64 : Precision::FP32;
			break;
		}
		case 'c': {
			int n = 0;
			if (pk_device_count(&n) != PK_OK) n = 0;
			cout << "GPU devices (HIP): " << n << std::endl;
			return 0;
		}
		case 'n': m.set_omp_threads(stoui(optarg)); break;
		case 'e':
		case 'I': break;
		case 'r': prefs.reuse_krn_fft = true; break;
		default: usage(cerr, argv[0]); return 1;
		}
	}
	if (!m.has_profiles()) {
		usage(cerr, argv[0]);
		return 1;
	}
	Dimensions dims(width, height);
	m.set_dimensions(dims);
	m.set_image_pixel_scale({scale_x, scale_y});
	m.set_finesampling(finesampling);
	prefs.src_dims = dims * finesampling;
	auto env = get_gpu_environment(device, precision);
	m.set_gpu_env(env);
	prefs.opencl_env = env;

	using clock = std::chrono::system_clock;
	auto start = clock::now();
	m.set_convolver(create_convolver(convolver_type, prefs));
	auto end = clock::now();
	cout << std::fixed << std::setprecision(3) << "Created convolver in "
	     << std::chrono::duration_cast<std::chrono::milliseconds>(end - start).count() << " [ms]" << std::endl;

	Point offset;
	Image image = run(cout, iterations, m, offset);
	if (output == text) {
		cout << std::setprecision(17) << std::scientific;
		for (unsigned int j = 0; j < image.getHeight(); j++) {
			for (unsigned int i = 0; i < image.getWidth(); i++) cout << image[j * image.getWidth() + i] << " ";
			cout << std::endl;
		}
	} else if (output == fits) {
		to_fits(image, offset, m.get_image_pixel_scale(), fits_output);
	}
	if (show_stats) {
		cout << std::fixed << std::setprecision(3);
		for (auto &kv : m.get_stats())
			cout << "Profile " << kv.first << ": device time " << kv.second->total_ms << " [ms], "
			     << kv.second->refined_pixels << " refined pixels, " << kv.second->subcells << " sub-cells"
			     << std::endl;
	}
	return 0;
}

}  // namespace

int main(int argc, char *argv[])
{
	try {
		return parse_and_run(argc, argv, std::cout, std::cerr);
	} catch (const invalid_cmdline &e) {
		std::cerr << "Error on command line: " << e.what() << std::endl;
		return 1;
	} catch (const invalid_parameter &e) {
		std::cerr << "Invalid parameter: " << e.what() << std::endl;
		return 1;
	} catch (const std::exception &e) {
		std::cerr << "Error: " << e.what() << std::endl;
		return 1;
	}
}
