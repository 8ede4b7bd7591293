// C++ host API (include/profit/*.h) over the C-ABI of include/profit_hip.h.
//
// profit::Model keeps libprofit's state and setters (include/profit/model.h);
// evaluate() turns it into a pk_model and calls pk_evaluate on the device.
// Profile keeps libprofit's named-parameter registry
// (src/profile.cpp:69-163, src/radial.cpp:738-751 and the profile
// constructors), bound to the fields of a pk_component.  Errors of the C-ABI
// come back as the exception libprofit throws for the same condition.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <sstream>

#include "profit/profit.h"

namespace profit {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void throw_status(int status)
{
	const std::string msg = pk_last_error();
	switch (status) {
	case PK_OK: return;
	case PK_EUNKNOWN: throw unknown_parameter(msg);
	case PK_EINVAL: throw invalid_parameter(msg);
	case PK_EFFT: throw fft_error(msg);
	default: throw gpu_error(msg);
	}
}

static void check(int status) { throw_status(status); }

// ---------------------------------------------------------------------------
// library
// ---------------------------------------------------------------------------
std::string version() { return pk_version(); }
unsigned short version_major() { return PK_VERSION_MAJOR; }
unsigned short version_minor() { return PK_VERSION_MINOR; }
unsigned short version_patch() { return PK_VERSION_PATCH; }
std::string version_suffix() { return ""; }

static std::string g_init_diagnose;

bool init()
{
	g_init_diagnose.clear();
	if (pk_init() != PK_OK) {
		g_init_diagnose = pk_last_error();
		return false;
	}
	return true;
}
std::string init_diagnose() { return g_init_diagnose; }
void finish() { pk_finish(); }
std::string finish_diagnose() { return ""; }
bool has_openmp() { return false; }
bool has_fftw() { return false; }
bool has_fftw_with_openmp() { return false; }
bool has_opencl() { return false; }
bool has_gpu()
{
	int n = 0;
	return pk_device_count(&n) == PK_OK && n > 0;
}
void clear_cache() { check(pk_clear_cache()); }
bool has_simd_instruction_set(simd_instruction_set instruction_set)
{
	return instruction_set == AUTO || instruction_set == NONE;
}
unsigned short opencl_version_major() { return 0; }
unsigned short opencl_version_minor() { return 0; }

std::string GpuEnv::get_device_name() { return pk_ctx_device_name(static_cast<const pk_ctx *>(ctx_)); }

std::map<int, OpenCL_plat_info> get_opencl_info()
{
	std::map<int, OpenCL_plat_info> info;
	int n = 0;
	if (pk_device_count(&n) != PK_OK || n <= 0) return info;
	OpenCL_plat_info plat{"HIP", 0, {}};
	for (int d = 0; d < n; d++) plat.dev_info[d] = OpenCL_dev_info{get_gpu_environment(d)->get_device_name(), 0, true};
	info[0] = plat;
	return info;
}

OpenCLEnvPtr get_opencl_environment(unsigned int platform_idx, unsigned int device_idx, bool use_double, bool)
{
	if (platform_idx != 0) throw invalid_parameter("platform " + std::to_string(platform_idx) + " does not exist");
	return get_gpu_environment((int)device_idx, use_double ? Precision::FP64 : Precision::FP32);
}

static GpuEnvPtr as_gpu_env(const OpenCLEnvPtr &env)
{
	if (!env) return nullptr;
	auto g = std::dynamic_pointer_cast<GpuEnv>(env);
	if (!g) throw invalid_parameter("the environment is not a device environment of this library (GpuEnv)");
	return g;
}

GpuEnv::GpuEnv(int device, Precision precision) : device_(device), precision_(precision)
{
	pk_ctx *c = nullptr;
	check(pk_ctx_create(device, precision == Precision::FP32 ? PK_FP32 : PK_FP64, &c));
	ctx_ = c;
}

GpuEnv::~GpuEnv()
{
	if (ctx_) pk_ctx_destroy(static_cast<pk_ctx *>(ctx_));
}

GpuEnvPtr get_gpu_environment(int device, Precision precision)
{
	static std::mutex mu;
	static std::map<std::pair<int, int>, std::weak_ptr<GpuEnv>> envs;
	std::lock_guard<std::mutex> lock(mu);
	auto key = std::make_pair(device, (int)precision);
	if (auto e = envs[key].lock()) return e;
	auto e = std::make_shared<GpuEnv>(device, precision);
	envs[key] = e;
	return e;
}

// ---------------------------------------------------------------------------
// images
// ---------------------------------------------------------------------------
Mask Mask::expand_by(Dimensions pad, int) const
{
	// separable dilation by a (2 pad.x + 1) x (2 pad.y + 1) box, clipped at
	// the edges: row pass, then column pass
	const unsigned int w = getWidth(), h = getHeight();
	std::vector<bool> rows(data_.size());
	for (unsigned int j = 0; j < h; j++) {
		int last = -(int)pad.x - 1;
		std::vector<int> next(w + 1, (int)(w + pad.x + 1));
		for (int i = (int)w - 1; i >= 0; i--) next[i] = data_[i + (size_t)j * w] ? i : next[i + 1];
		for (unsigned int i = 0; i < w; i++) {
			if (data_[i + (size_t)j * w]) last = (int)i;
			rows[i + (size_t)j * w] = ((int)i - last) <= (int)pad.x || (next[i] - (int)i) <= (int)pad.x;
		}
	}
	Mask out(getDimensions());
	for (unsigned int i = 0; i < w; i++) {
		int last = -(int)pad.y - 1;
		std::vector<int> next(h + 1, (int)(h + pad.y + 1));
		for (int j = (int)h - 1; j >= 0; j--) next[j] = rows[i + (size_t)j * w] ? j : next[j + 1];
		for (unsigned int j = 0; j < h; j++) {
			if (rows[i + (size_t)j * w]) last = (int)j;
			out[i + (size_t)j * w] = ((int)j - last) <= (int)pad.y || (next[j] - (int)j) <= (int)pad.y;
		}
	}
	return out;
}

template <typename S, typename V>
static S upsample_surface(const S &s, unsigned int factor, V scale)
{
	const Dimensions d = s.getDimensions() * factor;
	S out(d);
	for (unsigned int r = 0; r < d.y; r++)
		for (unsigned int c = 0; c < d.x; c++) out[c + (size_t)r * d.x] = scale(s[Point(c / factor, r / factor)]);
	return out;
}

Mask Mask::upsample(unsigned int factor) const
{
	if (factor == 0) throw std::invalid_argument("upsampling factor is 0");
	if (factor == 1) return *this;
	return upsample_surface(*this, factor, [](bool v) { return v; });
}

std::vector<unsigned char> Mask::bytes() const
{
	std::vector<unsigned char> b(data_.size());
	for (size_t i = 0; i < data_.size(); i++) b[i] = data_[i] ? 1 : 0;
	return b;
}

double Image::total() const
{
	double s = 0;
	for (double v : data_) s += v;
	return s;
}

void Image::normalize()
{
	const double s = total();
	if (s > 0) *this /= s;
}

Image Image::normalize() const
{
	Image r(*this);
	r.normalize();
	return r;
}

Image Image::upsample(unsigned int factor, UpsamplingMode mode) const
{
	if (factor == 0) throw std::invalid_argument("upsampling factor is 0");
	if (factor == 1) return *this;
	const double div = mode == SCALE ? double(factor) * factor : 1.0;
	return upsample_surface(*this, factor, [div](double v) { return v / div; });
}

Image Image::downsample(unsigned int factor, DownsamplingMode mode) const
{
	if (factor == 0) throw std::invalid_argument("downsampling factor is 0");
	if (factor == 1) return *this;
	const Dimensions d = getDimensions();
	const Dimensions dd((d.x + factor - 1) / factor, (d.y + factor - 1) / factor);
	Image out(dd);
	if (mode == SAMPLE) {
		for (unsigned int r = 0; r < dd.y; r++)
			for (unsigned int c = 0; c < dd.x; c++) out[c + (size_t)r * dd.x] = (*this)[Point(c * factor, r * factor)];
	} else if (mode == SUM) {
		for (unsigned int r = 0; r < d.y; r++)
			for (unsigned int c = 0; c < d.x; c++) out[c / factor + (size_t)(r / factor) * dd.x] += (*this)[Point(c, r)];
	} else {
		for (unsigned int rd = 0; rd < dd.y; rd++)
			for (unsigned int cd = 0; cd < dd.x; cd++) {
				double t = 0;
				unsigned int n = 0;
				for (unsigned int r = rd * factor; r < std::min(rd * factor + factor, d.y); r++)
					for (unsigned int c = cd * factor; c < std::min(cd * factor + factor, d.x); c++, n++)
						t += (*this)[Point(c, r)];
				out[cd + (size_t)rd * dd.x] = t / n;
			}
	}
	return out;
}

#define PK_IMAGE_OP(OP)                                                                   \
	Image &Image::operator OP(const Image &o)                                             \
	{                                                                                     \
		if (o.getDimensions() != getDimensions())                                         \
			throw std::invalid_argument("Images have different dimensions");              \
		for (size_t i = 0; i < data_.size(); i++) data_[i] OP o.data_[i];                 \
		return *this;                                                                     \
	}                                                                                     \
	Image &Image::operator OP(double v)                                                   \
	{                                                                                     \
		for (double &x : data_) x OP v;                                                   \
		return *this;                                                                     \
	}
PK_IMAGE_OP(+=)
PK_IMAGE_OP(-=)
PK_IMAGE_OP(*=)
PK_IMAGE_OP(/=)
#undef PK_IMAGE_OP

Image &Image::operator&=(const Mask &m)
{
	if (m.empty()) return *this;  // empty masks are not applied (src/image.cpp:324-328)
	if (m.getDimensions() != getDimensions()) throw std::invalid_argument("Mask and image have different dimensions");
	for (size_t i = 0; i < data_.size(); i++)
		if (!m[i]) data_[i] = 0;
	return *this;
}

// ---------------------------------------------------------------------------
// profiles
// ---------------------------------------------------------------------------
Profile::Profile(const std::string &name_, int kind)
    : name(name_),
      stats(kind == PK_PROFILE_NULL || kind == PK_PROFILE_SKY || kind == PK_PROFILE_PSF
                ? std::make_shared<ProfileStats>()
                : std::static_pointer_cast<ProfileStats>(std::make_shared<RadialProfileStats>()))
{
	check(pk_component_defaults(kind, &component));
	pk_component &c = component;
	register_bool("convolve", c.convolve);  // src/profile.cpp:44
	const bool radial = kind != PK_PROFILE_NULL && kind != PK_PROFILE_SKY && kind != PK_PROFILE_PSF;
	if (radial) {
		// src/radial.cpp:738-751
		register_bool("rough", c.rough);
		register_bool("adjust", c.adjust);
		register_bool("force_cpu", c.force_cpu);
		register_double("xcen", c.xcen);
		register_double("ycen", c.ycen);
		register_double("mag", c.mag);
		register_double("ang", c.ang);
		register_double("axrat", c.axrat);
		register_double("box", c.box);
		register_double("acc", c.acc);
		register_double("rscale_switch", c.rscale_switch);
		register_double("rscale_max", c.rscale_max);
		register_uint("max_recursions", c.max_recursions);
		register_uint("resolution", c.resolution);
	}
	double *p = c.p;
	switch (kind) {
	case PK_PROFILE_SERSIC:  // src/sersic.cpp:357-359
		register_double("re", p[PK_SERSIC_RE]);
		register_double("nser", p[PK_SERSIC_NSER]);
		register_bool("rescale_flux", c.rescale_flux);
		break;
	case PK_PROFILE_MOFFAT:  // src/moffat.cpp:105-106
		register_double("fwhm", p[PK_MOFFAT_FWHM]);
		register_double("con", p[PK_MOFFAT_CON]);
		break;
	case PK_PROFILE_FERRER:  // src/ferrer.cpp:122-124
		register_double("rout", p[PK_FERRER_ROUT]);
		register_double("a", p[PK_FERRER_A]);
		register_double("b", p[PK_FERRER_B]);
		break;
	case PK_PROFILE_CORESERSIC:  // src/coresersic.cpp:135-139
		register_double("re", p[PK_CORESERSIC_RE]);
		register_double("rb", p[PK_CORESERSIC_RB]);
		register_double("nser", p[PK_CORESERSIC_NSER]);
		register_double("a", p[PK_CORESERSIC_A]);
		register_double("b", p[PK_CORESERSIC_B]);
		break;
	case PK_PROFILE_KING:  // src/king.cpp:115-117
		register_double("rc", p[PK_KING_RC]);
		register_double("rt", p[PK_KING_RT]);
		register_double("a", p[PK_KING_A]);
		break;
	case PK_PROFILE_BROKENEXP:  // src/brokenexponential.cpp:145-148
		register_double("h1", p[PK_BROKENEXP_H1]);
		register_double("h2", p[PK_BROKENEXP_H2]);
		register_double("rb", p[PK_BROKENEXP_RB]);
		register_double("a", p[PK_BROKENEXP_A]);
		break;
	case PK_PROFILE_SKY:  // src/sky.cpp:70
		register_double("bg", p[PK_SKY_BG]);
		break;
	case PK_PROFILE_PSF:  // src/psf.cpp:153-155
		register_double("xcen", c.xcen);
		register_double("ycen", c.ycen);
		register_double("mag", c.mag);
		break;
	default: break;
	}
}

template <typename T, typename V>
static void set_typed(std::map<std::string, std::reference_wrapper<T>> &reg, const std::string &name,
                      const std::string &profile, const char *tname, V value)
{
	auto it = reg.find(name);
	if (it == reg.end()) {
		std::ostringstream os;
		os << "Unknown " << tname << " parameter in profile " << profile << ": " << name;
		throw invalid_parameter(os.str());
	}
	it->second.get() = (T)value;
}

void Profile::parameter(const std::string &n, bool v) { set_typed(bools, n, name, "bool", v ? 1 : 0); }
void Profile::parameter(const std::string &n, double v) { set_typed(doubles, n, name, "double", v); }
void Profile::parameter(const std::string &n, unsigned int v) { set_typed(uints, n, name, "unsigned int", v); }

static std::string trim(const std::string &s)
{
	const char *ws = " \t\n\r\f\v";
	const size_t b = s.find_first_not_of(ws);
	if (b == std::string::npos) return "";
	return s.substr(b, s.find_last_not_of(ws) - b + 1);
}

void Profile::parameter(const std::string &spec)
{
	// src/profile.cpp:139-163: the bool, unsigned int and double registries in
	// that order; stoul / stoui / stod conversions
	const size_t eq = spec.find('=');
	if (eq == std::string::npos || spec.find('=', eq + 1) != std::string::npos) {
		std::ostringstream os;
		os << "missing = in parameter: " << spec;
		throw invalid_parameter(os.str());
	}
	const std::string pname = trim(spec.substr(0, eq)), val = trim(spec.substr(eq + 1));
	auto bad = [&](const char *tname) {
		std::ostringstream os;
		os << "Parameter " << pname << " in profile " << name << " is of type " << tname
		   << ", but given value cannot be parsed as " << tname << ": " << val;
		return invalid_parameter(os.str());
	};
	if (bools.count(pname)) {
		try {
			bools.at(pname).get() = std::stoul(val, nullptr, 10) ? 1 : 0;
		} catch (const std::logic_error &) {
			throw bad("bool");
		}
		return;
	}
	if (uints.count(pname)) {
		try {
			const unsigned long v = std::stoul(val, nullptr, 10);
			if (v > 0xffffffffUL) throw std::out_of_range(val);
			uints.at(pname).get() = (uint32_t)v;
		} catch (const std::logic_error &) {
			throw bad("unsigned int");
		}
		return;
	}
	if (doubles.count(pname)) {
		try {
			doubles.at(pname).get() = std::stod(val);
		} catch (const std::logic_error &) {
			throw bad("double");
		}
		return;
	}
	std::ostringstream os;
	os << "Profile " << name << " doesn't support parameter " << pname << ", or parameter has invalid value: " << val;
	throw unknown_parameter(os.str());
}

// ---------------------------------------------------------------------------
// convolvers
// ---------------------------------------------------------------------------
Point Convolver::NO_OFFSET;

Convolver::Convolver(ConvolverType type, const ConvolverCreationPreferences &prefs) : type_(type), prefs_(prefs) {}

Image Convolver::convolve(const Image &src, const Image &krn, const Mask &mask, bool crop, Point &offset_out)
{
	GpuEnvPtr env = prefs_.opencl_env ? as_gpu_env(prefs_.opencl_env) : get_gpu_environment();
	const Dimensions s = src.getDimensions(), k = krn.getDimensions();
	const size_t mx = std::max({s.x, s.y, k.x, k.y});
	std::vector<double> out(std::max<size_t>(4 * mx * mx, (size_t)s.x * s.y));
	std::vector<unsigned char> mb;
	if (!mask.empty()) {
		if (mask.getDimensions() != s) throw invalid_parameter("mask dimensions != source dimensions");
		mb = mask.bytes();
	}
	uint32_t ow = 0, oh = 0;
	int32_t ox = 0, oy = 0;
	check(pk_convolve(static_cast<pk_ctx *>(env->handle()), type_, src._get().data(), s.x, s.y, krn._get().data(), k.x,
	                  k.y, mb.empty() ? nullptr : mb.data(), crop ? 1 : 0, out.data(), out.size(), &ow, &oh, &ox,
	                  &oy));
	out.resize((size_t)ow * oh);
	if (&offset_out != &NO_OFFSET) offset_out = Point((unsigned int)ox, (unsigned int)oy);
	return Image(std::move(out), ow, oh);
}

PointPair Convolver::padding(const Dimensions &src_dims, const Dimensions &krn_dims) const
{
	int32_t lx, ly, hx, hy;
	check(pk_convolver_padding(type_, src_dims.x, src_dims.y, krn_dims.x, krn_dims.y, &lx, &ly, &hx, &hy));
	return {Point((unsigned int)lx, (unsigned int)ly), Point((unsigned int)hx, (unsigned int)hy)};
}

ConvolverPtr create_convolver(const ConvolverType type, const ConvolverCreationPreferences &prefs)
{
	switch (type) {
	case BRUTE_OLD:
	case BRUTE:
	case OPENCL:
	case FFT: return std::make_shared<Convolver>(type, prefs);
	default: throw invalid_parameter("Unsupported convolver type: " + std::to_string((int)type));
	}
}

ConvolverPtr create_convolver(const std::string &type, const ConvolverCreationPreferences &prefs)
{
	// src/convolve.cpp:562-584
	if (type == "brute-old") return create_convolver(BRUTE_OLD, prefs);
	if (type == "brute") return create_convolver(BRUTE, prefs);
	if (type == "opencl") return create_convolver(OPENCL, prefs);
	if (type == "fft") return create_convolver(FFT, prefs);
	std::ostringstream os;
	os << "Convolver of type " << type << " is not supported";
	throw invalid_parameter(os.str());
}

// ---------------------------------------------------------------------------
// model
// ---------------------------------------------------------------------------
Point Model::NO_OFFSET;

Model::Model(unsigned int width, unsigned int height) : dimensions(width, height) {}
Model::Model(Dimensions d) : dimensions(d) {}

ProfilePtr Model::add_profile(const std::string &profile_name)
{
	// src/model.cpp:103-135
	static const std::map<std::string, int> kinds = {
	    {"sersic", PK_PROFILE_SERSIC},   {"moffat", PK_PROFILE_MOFFAT},       {"ferrer", PK_PROFILE_FERRER},
	    {"ferrers", PK_PROFILE_FERRER},  {"coresersic", PK_PROFILE_CORESERSIC}, {"king", PK_PROFILE_KING},
	    {"brokenexp", PK_PROFILE_BROKENEXP}, {"sky", PK_PROFILE_SKY},           {"psf", PK_PROFILE_PSF},
	    {"null", PK_PROFILE_NULL}};
	auto it = kinds.find(profile_name);
	if (it == kinds.end()) throw invalid_parameter("Unknown profile name: " + profile_name);
	auto p = std::make_shared<Profile>(profile_name, it->second);
	profiles.push_back(p);
	return p;
}

std::map<std::string, std::shared_ptr<ProfileStats>> Model::get_stats() const
{
	std::map<std::string, std::shared_ptr<ProfileStats>> s;
	for (auto &p : profiles) s[p->get_name()] = p->get_stats();
	return s;
}

pk_model Model::to_pk_model(std::vector<pk_component> &comps, std::vector<unsigned char> &mask_bytes) const
{
	pk_model m;
	check(pk_model_defaults(dimensions.x, dimensions.y, &m));
	m.scale_x = scale.first;
	m.scale_y = scale.second;
	m.magzero = magzero;
	m.finesampling = finesampling;
	m.crop = crop;
	m.return_finesampled = return_finesampled;
	m.adjust_mask = adjust_mask;
	m.dry_run = dry_run;
	if (!mask.empty()) {
		mask_bytes = mask.bytes();
		m.mask = mask_bytes.data();
		m.mask_width = mask.getWidth();
		m.mask_height = mask.getHeight();
	}
	if (!psf.empty()) {
		m.psf = psf._get().data();
		m.psf_width = psf.getWidth();
		m.psf_height = psf.getHeight();
	}
	m.psf_scale_x = psf_scale.first;
	m.psf_scale_y = psf_scale.second;
	m.convolver = convolver ? (int32_t)convolver->type() : (int32_t)PK_CONV_BRUTE;
	m.reuse_krn_fft = convolver ? convolver->reuse_krn_fft() : 0;
	comps.clear();
	for (auto &p : profiles) comps.push_back(p->get_component());
	m.n_components = (uint32_t)comps.size();
	m.components = comps.empty() ? nullptr : comps.data();
	return m;
}

Image Model::evaluate(Point &offset_out)
{
	std::vector<pk_component> comps;
	std::vector<unsigned char> mb;
	pk_model m = to_pk_model(comps, mb);
	// validation and drawing dimensions (Model::analyze_inputs) on the host
	uint32_t dw = 0, dh = 0;
	check(pk_model_output_dims(&m, &dw, &dh));
	if (dry_run) {
		// src/model.cpp:263-267
		if (&offset_out != &NO_OFFSET) offset_out = Point(0, 0);
		return Image(Dimensions(dw, dh));
	}
	if (!gpu_env) gpu_env = get_gpu_environment();
	const size_t mx = std::max({(size_t)dw, (size_t)dh, (size_t)m.psf_width, (size_t)m.psf_height});
	std::vector<double> out(std::max<size_t>(4 * mx * mx, (size_t)dw * dh));
	pk_result r;
	check(pk_evaluate(static_cast<pk_ctx *>(gpu_env->handle()), &m, out.data(), out.size(), &r));
	out.resize((size_t)r.width * r.height);
	if (&offset_out != &NO_OFFSET) offset_out = Point((unsigned int)r.offset_x, (unsigned int)r.offset_y);
	std::vector<uint64_t> integr(profiles.size());
	check(pk_ctx_last_integrations(static_cast<pk_ctx *>(gpu_env->handle()), integr.data(),
	                               (uint32_t)integr.size()));
	integrations.clear();
	for (size_t k = 0; k < profiles.size(); k++) {
		auto &p = profiles[k];
		p->stats->total = (nsecs_t)(r.ms_total * 1e6);
		p->stats->total_ms = r.ms_total;
		p->stats->refined_pixels = r.n_refined_pixels;
		p->stats->subcells = r.n_subcells;
		if (auto *rs = dynamic_cast<RadialProfileStats *>(p->stats.get())) {
			rs->subsampling.total = (nsecs_t)(r.ms_profiles * 1e6);
			integrations[p->get_name()][0] += (int)integr[k];
		}
	}
	return Image(std::move(out), r.width, r.height);
}

void Model::set_opencl_env(const OpenCLEnvPtr &env) { gpu_env = as_gpu_env(env); }

std::map<std::string, std::map<int, int>> Model::get_profile_integrations() const { return integrations; }

void Model::adjust(Mask &mask, const Dimensions &dims, const Image &psf, unsigned int finesampling)
{
	// src/model.cpp:152-184 and 230-257
	if (mask.empty()) return;
	const bool conv = psf.size() > 0;
	bool needs_conv = false, needs_pad = false;
	if (conv) {
		const Box b = mask.bounding_box();
		const Point lo = b.first * finesampling;
		const Point up = mask.getDimensions() * finesampling - b.second * finesampling;
		const Dimensions need = psf.getDimensions() / 2;
		needs_pad = !(lo >= need) || !(up >= need);
		needs_conv = true;
	}
	const Dimensions pad = needs_pad ? psf.getDimensions() / 2 : Dimensions();
	if (!(needs_conv || needs_pad || finesampling > 1)) return;
	if (finesampling > 1) mask = mask.upsample(finesampling);
	if (needs_pad) mask = mask.extend(dims * finesampling + pad * 2, pad);
	if (needs_conv) mask = mask.expand_by(psf.getDimensions() / 2);
}

}  // namespace profit
