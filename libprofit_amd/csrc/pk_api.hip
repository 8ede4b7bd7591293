// C-ABI of the gfx950 image-generation path (include/profit_hip.h).
//
// Host side of one Model::evaluate (src/model.cpp:259-360): input validation
// and geometry (analyze_inputs / analyze_expansion_requirements), calculation
// mask adjustment (Model::adjust, Mask::upsample/extend/expand_by), per
// component FP64 set-up, then one stream-ordered sequence of kernels:
//   coarse -> refine -> psf profile -> convolution -> compose.
// Inputs travel in one pinned upload, the result in one pinned download.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "pk_device.h"
#include "pk_fft.h"
#include "pk_strips.h"
#include "profit_hip.h"

using namespace pk;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
// integrator of the luminosity integrals (pk_set_integrator)
static std::atomic<int> g_integrator{PK_INTEGRATOR_QUADPACK};

static int set_err(int code, const char *fmt, ...)
{
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	g_err = buf;
	return code;
}

#define PK_HIP(expr)                                                                                   \
	do {                                                                                               \
		hipError_t e_ = (expr);                                                                        \
		if (e_ != hipSuccess) return set_err(PK_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
	} while (0)

#define PK_TRY(expr)              \
	do {                          \
		int st_ = (expr);         \
		if (st_ != PK_OK) return st_; \
	} while (0)

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct pk_ctx {
	int device = 0;
	int precision = PK_FP64;
	hipStream_t stream = nullptr;
	// a second stream (default priority, as `stream`) for the strip-form
	// convolution classes that run beside the others, forked from and joined
	// to `stream` by events (in graph capture: parallel branches)
	hipStream_t side = nullptr;
	hipEvent_t fork = nullptr, join = nullptr;
	std::mutex mu;
	void *dmem = nullptr;
	size_t dcap = 0;
	void *hmem = nullptr;  // pinned staging
	size_t hcap = 0;
	hipEvent_t ev[5] = {};
	// grids of the FP32 refinement passes: as many wavefronts as fit on the
	// device at once (the hand-over rule counts them)
	uint32_t r32_grid_main = 0, r32_grid_last = 0;
	uint32_t r32g_grid_main = 0, r32g_grid_last = 0;  // the same for the adjoint variants
	std::string arch;  // device name and CU count (autotuner cache key)
	// set while pk_evaluate_device_noalloc runs: dmem / hmem are then the
	// library's fixed workspace, and nothing may be (re)allocated
	bool fixed_ws = false;
	// refined pixels per model component of the last single evaluation
	std::vector<uint64_t> last_integrations;
};

static int ensure_dev(pk_ctx *ctx, size_t bytes)
{
	if (bytes <= ctx->dcap) return PK_OK;
	if (ctx->fixed_ws)
		return set_err(PK_ERANGE, "the model needs %zu bytes of device workspace, the allocation-free entry point "
		                          "has %zu (PK_NOALLOC_SCRATCH_BYTES)", bytes, ctx->dcap);
	if (ctx->dmem) PK_HIP(hipFree(ctx->dmem));
	ctx->dmem = nullptr;
	ctx->dcap = 0;
	size_t cap = std::max(bytes + bytes / 4, (size_t)1 << 20);
	if (hipMalloc(&ctx->dmem, cap) != hipSuccess) {
		(void)hipGetLastError();
		return set_err(PK_ENOMEM, "device allocation of %zu bytes failed", cap);
	}
	ctx->dcap = cap;
	return PK_OK;
}

static int ensure_host(pk_ctx *ctx, size_t bytes)
{
	if (bytes <= ctx->hcap) return PK_OK;
	if (ctx->fixed_ws)
		return set_err(PK_ERANGE, "the model needs %zu bytes of host staging, the allocation-free entry point has "
		                          "%zu (PK_NOALLOC_STAGE_BYTES)", bytes, ctx->hcap);
	if (ctx->hmem) PK_HIP(hipHostFree(ctx->hmem));
	ctx->hmem = nullptr;
	ctx->hcap = 0;
	size_t cap = std::max(bytes + bytes / 4, (size_t)1 << 20);
	if (hipHostMalloc(&ctx->hmem, cap, hipHostMallocDefault) != hipSuccess) {
		(void)hipGetLastError();
		return set_err(PK_ENOMEM, "pinned host allocation of %zu bytes failed", cap);
	}
	ctx->hcap = cap;
	return PK_OK;
}

// Bump layout of one call's device memory: the first `in_bytes` mirror the
// pinned staging buffer and are uploaded in one copy.
struct Layout {
	size_t off = 0;
	size_t take(size_t bytes)
	{
		size_t o = off;
		off += (bytes + 255) & ~(size_t)255;
		return o;
	}
};

namespace {

// FP32 refinement launches a component needs: the last launch refines levels
// passes - 1 .. max_recursions depth-first on an LDS stack of R32_CAP_LAST
// cells, which holds up to 128 claimed cells plus res^2 children per level
// (the Sersic centre override, resolution 8 and 10 levels, always fits).
uint32_t refine_passes_needed(const pk_component &c)
{
	if (c.kind == PK_PROFILE_SKY || c.kind == PK_PROFILE_NULL || c.kind == PK_PROFILE_PSF || c.rough)
		return R32_PASSES_MIN;
	const uint64_t res = c.adjust ? 16u : std::max(1u, c.resolution);
	const int levels = (int)((R32_CAP_LAST - 128) / (res * res));
	const int need = (int)std::min(c.max_recursions, 64u) + 2 - levels;
	return (uint32_t)std::min(std::max(need, (int)R32_PASSES_MIN), R32_PASSES);
}

// ---- autotuner cache -----------------------------------------------------------
// One text line per tuned workload: "<key>\t<conv> <occupancy> <handover> <claim
// pixels> <claim cells> <refinement launches> <spread> <ms>", in
// $PROFIT_HOME/hip_tuning (a line whose settings are not valid is skipped).
std::mutex g_tune_mu;
std::map<std::string, pk_tuning> g_tune_cache;
bool g_tune_dirty = false;

// $PROFIT_HOME, else $HOME/.profit (src/utils.cpp:471-495); "" if neither is set.
std::string profit_home()
{
	if (const char *p = std::getenv("PROFIT_HOME")) return p;
	if (const char *h = std::getenv("HOME")) return std::string(h) + "/.profit";
	return "";
}

std::string tune_cache_file()
{
	const std::string h = profit_home();
	return h.empty() ? h : h + "/hip_tuning";
}

bool tuning_valid(const pk_tuning &t)
{
	return t.conv_tiles <= PK_CONV_TILES_STRIPS && t.refine_occupancy >= 1 && t.refine_occupancy <= 100 &&
	       t.handover_div >= 1 && t.claim_pixels >= 1 && t.claim_cells >= 1 && t.refine_passes >= R32_PASSES_MIN &&
	       t.handover_div <= 1024 && t.claim_pixels <= 4096 && t.claim_cells <= 4096 && t.refine_passes <= R32_PASSES &&
	       t.claim_spread <= 4096 && (t.claim_spread & (t.claim_spread - 1)) == 0;
}

void tune_cache_load()
{
	const std::string f = tune_cache_file();
	if (f.empty()) return;
	FILE *fp = std::fopen(f.c_str(), "r");
	if (!fp) return;
	char line[1024];
	while (std::fgets(line, sizeof line, fp)) {
		char *tab = std::strchr(line, '\t');
		if (!tab) continue;
		*tab = 0;
		pk_tuning t{};
		if (std::sscanf(tab + 1, "%u %u %u %u %u %u %u %lf", &t.conv_tiles, &t.refine_occupancy, &t.handover_div,
		                &t.claim_pixels, &t.claim_cells, &t.refine_passes, &t.claim_spread, &t.ms_evaluate) == 8 &&
		    tuning_valid(t))
			g_tune_cache[line] = t;
	}
	std::fclose(fp);
}

int tune_cache_save()
{
	const std::string h = profit_home();
	if (h.empty()) return PK_OK;
	(void)::mkdir(h.c_str(), 0755);
	const std::string f = h + "/hip_tuning", tmp = f + ".tmp";
	FILE *fp = std::fopen(tmp.c_str(), "w");
	if (!fp) return set_err(PK_EINVAL, "cannot write %s", tmp.c_str());
	for (const auto &kv : g_tune_cache) {
		const pk_tuning &t = kv.second;
		std::fprintf(fp, "%s\t%u %u %u %u %u %u %u %.6f\n", kv.first.c_str(), t.conv_tiles, t.refine_occupancy,
		             t.handover_div, t.claim_pixels, t.claim_cells, t.refine_passes, t.claim_spread, t.ms_evaluate);
	}
	const bool ok = std::fclose(fp) == 0;
	if (!ok || std::rename(tmp.c_str(), f.c_str()) != 0) {
		(void)std::remove(tmp.c_str());
		return set_err(PK_EINVAL, "cannot write %s", f.c_str());
	}
	return PK_OK;
}

}  // namespace

extern "C" {

const char *pk_version(void) { return "1.9.3"; }

int pk_init(void)
{
	{
		std::lock_guard<std::mutex> lock(g_tune_mu);
		tune_cache_load();
	}
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess) return set_err(PK_EDEVICE, "hipGetDeviceCount failed: %s", hipGetErrorString(e));
	return PK_OK;
}

void pk_finish(void)
{
	std::lock_guard<std::mutex> lock(g_tune_mu);
	if (g_tune_dirty && tune_cache_save() == PK_OK) g_tune_dirty = false;
}

int pk_clear_cache(void)
{
	std::lock_guard<std::mutex> lock(g_tune_mu);
	g_tune_cache.clear();
	g_tune_dirty = false;
	const std::string f = tune_cache_file();
	if (!f.empty()) (void)std::remove(f.c_str());
	return PK_OK;
}

const char *pk_last_error(void) { return g_err.c_str(); }

int pk_device_count(int *count)
{
	if (!count) return set_err(PK_EINVAL, "count is NULL");
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess) {
		*count = 0;
		return set_err(PK_EDEVICE, "hipGetDeviceCount failed: %s", hipGetErrorString(e));
	}
	*count = n;
	return PK_OK;
}

int pk_ctx_create(int device, int precision, pk_ctx **ctx_out)
{
	if (!ctx_out) return set_err(PK_EINVAL, "ctx_out is NULL");
	if (precision != PK_FP64 && precision != PK_FP32) return set_err(PK_EINVAL, "unknown precision %d", precision);
	int n = 0;
	PK_HIP(hipGetDeviceCount(&n));
	if (device < 0 || device >= n) return set_err(PK_EINVAL, "device %d out of range (%d devices)", device, n);
	PK_HIP(hipSetDevice(device));
	pk_ctx *ctx = new pk_ctx();
	ctx->device = device;
	ctx->precision = precision;
	if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
		delete ctx;
		return set_err(PK_EDEVICE, "hipStreamCreate failed");
	}
	if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess) {
		(void)hipStreamDestroy(ctx->stream);
		delete ctx;
		return set_err(PK_EDEVICE, "hipStreamCreate failed");
	}
	for (auto &e : ctx->ev) PK_HIP(hipEventCreate(&e));
	PK_HIP(hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
	PK_HIP(hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming));
	hipDeviceProp_t prop;
	PK_HIP(hipGetDeviceProperties(&prop, device));
	int nb_main = 0, nb_last = 0;
	PK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_main, refine32_kernel<0, false>,
	                                                    64 * r32_waves_per_group(false, false), 0));
	PK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_last, refine32_kernel<1, true>,
	                                                    64 * r32_waves_per_group(true, false), 0));
	ctx->r32_grid_main = (uint32_t)std::max(1, nb_main * prop.multiProcessorCount);
	ctx->r32_grid_last = (uint32_t)std::max(1, nb_last * prop.multiProcessorCount);
	PK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_main, refine32_kernel<1, false, true>,
	                                                    64 * r32_waves_per_group(false, true), 0));
	PK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_last, refine32_kernel<1, true, true>,
	                                                    64 * r32_waves_per_group(true, true), 0));
	ctx->r32g_grid_main = (uint32_t)std::max(1, nb_main * prop.multiProcessorCount);
	ctx->r32g_grid_last = (uint32_t)std::max(1, nb_last * prop.multiProcessorCount);
	ctx->arch = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount);
	for (auto &ch : ctx->arch)
		if (ch == ' ' || ch == '\t' || ch == '\n') ch = '_';
	*ctx_out = ctx;
	return PK_OK;
}

int pk_ctx_destroy(pk_ctx *ctx)
{
	if (!ctx) return PK_OK;
	(void)hipSetDevice(ctx->device);
	(void)hipStreamSynchronize(ctx->stream);
	(void)hipStreamSynchronize(ctx->side);
	for (auto &e : ctx->ev)
		if (e) (void)hipEventDestroy(e);
	if (ctx->fork) (void)hipEventDestroy(ctx->fork);
	if (ctx->join) (void)hipEventDestroy(ctx->join);
	if (ctx->dmem) (void)hipFree(ctx->dmem);
	if (ctx->hmem) (void)hipHostFree(ctx->hmem);
	(void)hipStreamDestroy(ctx->stream);
	if (ctx->side) (void)hipStreamDestroy(ctx->side);
	delete ctx;
	return PK_OK;
}

const char *pk_ctx_device_name(const pk_ctx *ctx) { return ctx ? ctx->arch.c_str() : ""; }

int pk_ctx_last_integrations(const pk_ctx *ctx, uint64_t *counts, uint32_t n)
{
	if (!ctx || (!counts && n)) return set_err(PK_EINVAL, "NULL argument");
	for (uint32_t k = 0; k < n; k++) counts[k] = k < ctx->last_integrations.size() ? ctx->last_integrations[k] : 0;
	return PK_OK;
}

int pk_ctx_precision(const pk_ctx *ctx) { return ctx ? ctx->precision : -1; }

int pk_component_defaults(int kind, pk_component *c)
{
	if (!c) return set_err(PK_EINVAL, "component is NULL");
	if (kind < PK_PROFILE_NULL || kind > PK_PROFILE_PSF) return set_err(PK_EINVAL, "Unknown profile kind %d", kind);
	std::memset(c, 0, sizeof(*c));
	c->kind = kind;
	// RadialProfile constructor (src/radial.cpp:724-737)
	c->mag = 15;
	c->axrat = 1;
	c->acc = 0.1;
	c->rscale_switch = 1;
	c->resolution = 9;
	c->max_recursions = 2;
	c->adjust = 1;
	switch (kind) {
	case PK_PROFILE_SERSIC: c->p[PK_SERSIC_RE] = 1; c->p[PK_SERSIC_NSER] = 1; break;
	case PK_PROFILE_MOFFAT: c->p[PK_MOFFAT_FWHM] = 3; c->p[PK_MOFFAT_CON] = 2; break;
	case PK_PROFILE_FERRER:
		c->p[PK_FERRER_ROUT] = 3; c->p[PK_FERRER_A] = 1; c->p[PK_FERRER_B] = 1;
		c->acc = 1;  // src/ferrer.cpp:120
		break;
	case PK_PROFILE_CORESERSIC:
		c->p[PK_CORESERSIC_RE] = 1; c->p[PK_CORESERSIC_RB] = 1; c->p[PK_CORESERSIC_NSER] = 4;
		c->p[PK_CORESERSIC_A] = 1; c->p[PK_CORESERSIC_B] = 1;
		break;
	case PK_PROFILE_KING: c->p[PK_KING_RC] = 1; c->p[PK_KING_RT] = 3; c->p[PK_KING_A] = 2; break;
	case PK_PROFILE_BROKENEXP:
		c->p[PK_BROKENEXP_H1] = 1; c->p[PK_BROKENEXP_H2] = 1; c->p[PK_BROKENEXP_RB] = 1; c->p[PK_BROKENEXP_A] = 1;
		break;
	case PK_PROFILE_PSF: c->mag = 0; break;  // src/psf.cpp:147-156
	default: break;
	}
	return PK_OK;
}

int pk_model_defaults(uint32_t width, uint32_t height, pk_model *m)
{
	if (!m) return set_err(PK_EINVAL, "model is NULL");
	std::memset(m, 0, sizeof(*m));
	// Model constructor (src/model.cpp:54-73)
	m->width = width;
	m->height = height;
	m->scale_x = m->scale_y = 1;
	m->finesampling = 1;
	m->crop = 1;
	m->return_finesampled = 1;
	m->adjust_mask = 1;
	m->psf_scale_x = m->psf_scale_y = 1;
	m->convolver = PK_CONV_BRUTE;
	return PK_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// validation and geometry (src/model.cpp:152-219)
// ---------------------------------------------------------------------------
namespace {

struct Analysis {
	uint32_t dw = 0, dh = 0, padx = 0, pady = 0;
	bool conv_required = false, mask_needs_psf_padding = false, mask_needs_convolution = false,
	     mask_needs_adjustment = false;
};

bool has_mask(const pk_model *m) { return m->mask && m->mask_width > 0 && m->mask_height > 0; }
bool has_psf(const pk_model *m) { return m->psf && m->psf_width > 0 && m->psf_height > 0; }

int validate_component(const pk_component &c, bool psf_present)
{
	switch (c.kind) {
	case PK_PROFILE_NULL:
	case PK_PROFILE_SKY: return PK_OK;
	case PK_PROFILE_PSF:
		if (!psf_present) return set_err(PK_EINVAL, "No psf present in the model, cannot produce a psf profile");
		return PK_OK;
	case PK_PROFILE_SERSIC:
	case PK_PROFILE_MOFFAT:
	case PK_PROFILE_FERRER:
	case PK_PROFILE_CORESERSIC:
	case PK_PROFILE_KING:
	case PK_PROFILE_BROKENEXP: break;
	default: return set_err(PK_EINVAL, "Unknown profile kind %d", c.kind);
	}
	// RadialProfile::validate (src/radial.cpp:204-214)
	if (c.axrat <= 0) return set_err(PK_EINVAL, "axrat <= 0, must have axrat > 0");
	if (c.axrat > 1) return set_err(PK_EINVAL, "axrat > 1, must have axrat <= 1");
	if (c.box <= -2) return set_err(PK_EINVAL, "box <= -2, must have box > -2");
	const double *p = c.p;
	switch (c.kind) {
	case PK_PROFILE_SERSIC:  // src/sersic.cpp:207-218
		if (p[PK_SERSIC_RE] <= 0) return set_err(PK_EINVAL, "re <= 0, must have re > 0");
		if (p[PK_SERSIC_NSER] <= 0) return set_err(PK_EINVAL, "nser <= 0, must have nser > 0");
		break;
	case PK_PROFILE_MOFFAT:  // src/moffat.cpp:60-70
		if (p[PK_MOFFAT_FWHM] <= 0) return set_err(PK_EINVAL, "fwhm <= 0, must have fwhm > 0");
		if (p[PK_MOFFAT_CON] < 0) return set_err(PK_EINVAL, "con < 0, must have con >= 0");
		break;
	case PK_PROFILE_FERRER:  // src/ferrer.cpp:63-78
		if (p[PK_FERRER_ROUT] <= 0) return set_err(PK_EINVAL, "rout <= 0, must have rout >= 0");
		if (p[PK_FERRER_A] < 0) return set_err(PK_EINVAL, "a < 0, must have a >= 0");
		if (p[PK_FERRER_B] > 2) return set_err(PK_EINVAL, "b > 2, must have b <= 2");
		break;
	case PK_PROFILE_KING:  // src/king.cpp:61-76
		if (p[PK_KING_RC] <= 0) return set_err(PK_EINVAL, "rc <= 0, must have rc > 0");
		if (p[PK_KING_RT] <= 0) return set_err(PK_EINVAL, "rt <= 0, must have rt > 0");
		if (p[PK_KING_A] < 0) return set_err(PK_EINVAL, "a < 0, must have a >=0");
		break;
	case PK_PROFILE_BROKENEXP:  // src/brokenexponential.cpp:87-101
		if (p[PK_BROKENEXP_H1] <= 0) return set_err(PK_EINVAL, "h1 <= 0, must have h1 > 0");
		if (p[PK_BROKENEXP_H2] <= 0) return set_err(PK_EINVAL, "h2 <= 0, must have h2 > 0");
		if (p[PK_BROKENEXP_RB] <= 0) return set_err(PK_EINVAL, "rb <= 0, must have rb > 0");
		break;
	case PK_PROFILE_CORESERSIC:  // src/coresersic.cpp:62-80
		if (p[PK_CORESERSIC_RE] <= 0) return set_err(PK_EINVAL, "re <= 0, must have re > 0");
		if (p[PK_CORESERSIC_RB] <= 0) return set_err(PK_EINVAL, "rb <= 0, must have rb > 0");
		if (p[PK_CORESERSIC_NSER] <= 0) return set_err(PK_EINVAL, "nser <= 0, must have nser > 0");
		if (p[PK_CORESERSIC_A] <= 0) return set_err(PK_EINVAL, "a <= 0, must have a > 0");
		if (p[PK_CORESERSIC_B] > 1.999) return set_err(PK_EINVAL, "b > 1.999, must have b < 1.999");
		break;
	default: break;
	}
	return PK_OK;
}

// Mask::bounding_box (include/profit/image.h:451-483)
void mask_bbox(const uint8_t *m, uint32_t w, uint32_t h, uint32_t b[4])
{
	uint32_t lx = w, ly = h, ux = 0, uy = 0;
	bool any = false;
	for (uint32_t j = 0; j < h; j++)
		for (uint32_t i = 0; i < w; i++) {
			if (!m[i + (size_t)j * w]) continue;
			lx = std::min(lx, i);
			ly = std::min(ly, j);
			ux = std::max(ux, i + 1);
			uy = std::max(uy, j + 1);
			any = true;
		}
	if (!any) lx = ly = ux = uy = 0;
	b[0] = lx;
	b[1] = ly;
	b[2] = ux;
	b[3] = uy;
}

int analyze(const pk_model *m, Analysis &a)
{
	// Model::analyze_inputs (src/model.cpp:186-219)
	if (!(m->width > 0 && m->height > 0)) return set_err(PK_EINVAL, "Model's requested dimensions are 0");
	if (m->scale_x <= 0) return set_err(PK_EINVAL, "Model's scale_x cannot be negative or zero");
	if (m->scale_y <= 0) return set_err(PK_EINVAL, "Model's scale_y cannot be negative or zero");
	if (m->finesampling == 0) return set_err(PK_EINVAL, "finesampling must be > 0");
	if (m->n_components && !m->components) return set_err(PK_EINVAL, "components is NULL");
	if (has_mask(m) && m->adjust_mask && (m->mask_width != m->width || m->mask_height != m->height))
		return set_err(PK_EINVAL, "Mask dimensions != model dimensions");
	a.conv_required = false;
	for (uint32_t c = 0; c < m->n_components; c++)
		if (m->components[c].convolve) a.conv_required = true;
	if (a.conv_required && !has_psf(m))
		return set_err(PK_EINVAL, "No psf provided but profile(s) requested convolution");
	if (a.conv_required && m->convolver != PK_CONV_BRUTE_OLD && m->convolver != PK_CONV_BRUTE &&
	    m->convolver != PK_CONV_OPENCL && m->convolver != PK_CONV_FFT)
		return set_err(PK_EINVAL, "Unsupported convolver type: %d", m->convolver);
	for (uint32_t c = 0; c < m->n_components; c++) PK_TRY(validate_component(m->components[c], has_psf(m)));
	// analyze_expansion_requirements (src/model.cpp:152-184)
	const uint32_t fs = m->finesampling;
	bool need_pad = a.conv_required;
	if (has_mask(m) && !m->adjust_mask) {
		need_pad = m->mask_width > m->width * fs && m->mask_height > m->height * fs;
	} else if (has_mask(m) && a.conv_required) {
		uint32_t b[4];
		mask_bbox(m->mask, m->mask_width, m->mask_height, b);
		const uint32_t lx = b[0] * fs, ly = b[1] * fs, upx = m->mask_width * fs - b[2] * fs,
		               upy = m->mask_height * fs - b[3] * fs;
		const uint32_t nx = m->psf_width / 2, ny = m->psf_height / 2;
		need_pad = !(lx >= nx && ly >= ny) || !(upx >= nx && upy >= ny);
		a.mask_needs_convolution = true;
		a.mask_needs_psf_padding = need_pad;
	}
	a.padx = need_pad ? m->psf_width / 2 : 0;
	a.pady = need_pad ? m->psf_height / 2 : 0;
	a.mask_needs_adjustment = has_mask(m) && (a.mask_needs_convolution || a.mask_needs_psf_padding || fs > 1);
	a.dw = m->width * fs + 2 * a.padx;
	a.dh = m->height * fs + 2 * a.pady;
	return PK_OK;
}

// Mask::expand_by (src/image.cpp:112-162) as a separable dilation by a
// (2px+1) x (2py+1) box.
std::vector<uint8_t> mask_expand(const std::vector<uint8_t> &m, uint32_t w, uint32_t h, uint32_t px, uint32_t py)
{
	std::vector<uint8_t> rows(m.size()), out(m.size());
	for (uint32_t j = 0; j < h; j++) {
		const uint8_t *r = &m[(size_t)j * w];
		int last = -1000000000;  // last set column seen
		std::vector<int> next(w + 1, 1000000000);
		for (int i = (int)w - 1; i >= 0; i--) next[i] = r[i] ? i : next[i + 1];
		for (uint32_t i = 0; i < w; i++) {
			if (r[i]) last = (int)i;
			bool hit = ((int)i - last) <= (int)px || (next[i] - (int)i) <= (int)px;
			rows[(size_t)j * w + i] = hit;
		}
	}
	for (uint32_t i = 0; i < w; i++) {
		int last = -1000000000;
		std::vector<int> next(h + 1, 1000000000);
		for (int j = (int)h - 1; j >= 0; j--) next[j] = rows[(size_t)j * w + i] ? j : next[j + 1];
		for (uint32_t j = 0; j < h; j++) {
			if (rows[(size_t)j * w + i]) last = (int)j;
			bool hit = ((int)j - last) <= (int)py || (next[j] - (int)j) <= (int)py;
			out[(size_t)j * w + i] = hit || m[(size_t)j * w + i];
		}
	}
	return out;
}

// Calculation mask of a model in its drawing frame (0/1 per pixel), adjusted
// as Model::evaluate does (src/model.cpp:230-246, 265-285): upsampled by the
// finesampling, placed in the PSF padding, expanded by the PSF half-size.
// `adjusted`: the user's mask is applied again to the returned image.
// Work item of a psf profile: PsfProfile::evaluate's geometry
// (src/psf.cpp:58-103), including its mixed x/y use of the offsets and
// dimensions; job and psf_off are left to the caller.
PkPsfItem psf_item(const pk_model *m, const pk_component &pc, const Analysis &a, double sx, double sy)
{
	PkPsfItem it{};
	it.to_conv = pc.convolve ? 1 : 0;
	const double psx = m->psf_scale_x, psy = m->psf_scale_y;
	const uint32_t pw = m->psf_width, ph = m->psf_height;
	const double origin_x = pc.xcen + a.padx * sx - pw * psx / 2;
	const double end_x = pc.xcen + a.pady * sy + pw * psx / 2;
	const double origin_y = pc.ycen + a.padx * sx - ph * psy / 2;
	const double end_y = pc.ycen + a.pady * sy + ph * psy / 2;
	auto bind = [](double v, uint32_t mx) -> uint32_t {
		int iv = (int)std::floor(v);
		if (iv < 0) return 0;
		return std::min((uint32_t)iv, mx);
	};
	it.x0 = bind(origin_x / sx, a.dw - 1);
	it.y0 = bind(origin_y / sy, a.dh - 1);
	it.x1 = bind(end_x / sx, a.dw - 1);
	it.y1 = bind(end_y / sy, a.dh - 1);
	it.origin_x = origin_x;
	it.origin_y = origin_y;
	it.scale = std::pow(10, -0.4 * (pc.mag - m->magzero));
	it.psx = psx;
	it.psy = psy;
	it.pw = pw;
	it.ph = ph;
	return it;
}

int calc_mask(const pk_model *m, const Analysis &a, std::vector<uint8_t> &dmask, bool &adjusted)
{
	const uint32_t fs = m->finesampling;
	adjusted = false;
	if (m->adjust_mask && a.mask_needs_adjustment) {
		uint32_t w = m->mask_width, h = m->mask_height;
		std::vector<uint8_t> cur(m->mask, m->mask + (size_t)w * h);
		for (auto &v : cur) v = v ? 1 : 0;
		if (fs > 1) {
			std::vector<uint8_t> up((size_t)w * fs * h * fs);
			for (uint32_t r = 0; r < h * fs; r++)
				for (uint32_t c = 0; c < w * fs; c++) up[c + (size_t)r * w * fs] = cur[c / fs + (size_t)(r / fs) * w];
			cur.swap(up);
			w *= fs;
			h *= fs;
		}
		if (a.mask_needs_psf_padding) {
			std::vector<uint8_t> ext((size_t)a.dw * a.dh, 0);
			for (uint32_t j = 0; j < h; j++)
				for (uint32_t i = 0; i < w; i++) ext[(i + a.padx) + (size_t)(j + a.pady) * a.dw] = cur[i + (size_t)j * w];
			cur.swap(ext);
			w = a.dw;
			h = a.dh;
		}
		if (a.mask_needs_convolution) cur = mask_expand(cur, w, h, m->psf_width / 2, m->psf_height / 2);
		dmask.swap(cur);
		adjusted = true;
		return PK_OK;
	}
	if (!m->adjust_mask && (m->mask_width != a.dw || m->mask_height != a.dh))
		return set_err(PK_EINVAL, "Mask dimensions != drawing dimensions: [%u, %u] != [%u, %u]", m->mask_width,
		               m->mask_height, a.dw, a.dh);
	dmask.assign(m->mask, m->mask + (size_t)a.dw * a.dh);
	for (auto &v : dmask) v = v ? 1 : 0;
	return PK_OK;
}

// FFTConvolver::offset_after_convolution (src/convolve.cpp:270-281)
void fft_offset(uint32_t sw, uint32_t sh, uint32_t kw, uint32_t kh, int &ox, int &oy)
{
	ox = (int)(sw / 2);
	oy = (int)(sh / 2);
	if (sw % 2 == 0 || kw % 2 == 0) ox -= 1;
	if (sh % 2 == 0 || kh % 2 == 0) oy -= 1;
}

// Convolver::padding (src/convolve.cpp:50-53, 263-268); max() of
// include/profit/image.h:204-207 takes the x component for both axes.
void conv_padding(int type, uint32_t sw, uint32_t sh, uint32_t kw, uint32_t kh, int p[4])
{
	if (type != PK_CONV_FFT) {
		p[0] = p[1] = p[2] = p[3] = 0;
		return;
	}
	const uint32_t mx = std::max(sw, kw), ex = 2 * mx, ey = 2 * mx;
	int ox, oy;
	fft_offset(sw, sh, kw, kh, ox, oy);
	p[0] = ox;
	p[1] = oy;
	p[2] = (int)(ex - ex / 2) - ox;
	p[3] = (int)(ey - ey / 2) - oy;
}

// Geometry of Convolver::convolve on a (sw x sh) source for a given type:
// dims of the result, the same-size-frame shift of its pixels, the position
// of the source inside the result, and (FFT) the offset reported back.
struct ConvGeom {
	uint32_t cw, ch;
	int csx, csy;  // result(X,Y) = same(X + csx, Y + csy)
	int mx, my;    // source pixel (X - mx, Y - my) is the mask pixel of result (X,Y)
	int offx, offy;
	uint32_t wrap_y;  // FFT: rows of the circular buffer
};

int conv_geometry(int type, uint32_t sw, uint32_t sh, uint32_t kw, uint32_t kh, bool crop, ConvGeom &g)
{
	g = ConvGeom{sw, sh, 0, 0, 0, 0, 0, 0, 0};
	if (!(sw >= kw && sh >= kh)) {
		// src/convolve.cpp:65-75: extend, convolve, crop the extension back
		const uint32_t mxd = std::max(sw, kw);
		if (mxd < sh) return set_err(PK_EINVAL, "new_height should be >= height");
		if (type != PK_CONV_FFT) return PK_OK;  // same-size result, cropped back to sw x sh
		const uint32_t nw = mxd, nh = mxd;
		PK_TRY(conv_geometry(type, nw, nh, kw, kh, crop, g));
		// the crop keeps the top-left (result - (new - src)) block
		g.cw -= nw - sw;
		g.ch -= nh - sh;
		// same-size frame of the extended source equals the source's own frame
		return PK_OK;
	}
	if (type != PK_CONV_FFT) return PK_OK;
	int ox, oy;
	fft_offset(sw, sh, kw, kh, ox, oy);
	const uint32_t mxd = std::max(sw, kw);
	// libprofit's circular buffer has 2 max(src, krn) rows, the x component
	// (include/profit/image.h:204-207), and the crop of mask_and_crop fails
	// when it leaves it (include/profit/image.h:433-448)
	g.wrap_y = 2 * mxd;
	if (crop && ((uint32_t)ox + sw > 2 * mxd || (uint32_t)oy + sh > 2 * mxd))
		return set_err(PK_EINVAL, "start_y + new_height should be <= image.height");
	if (crop) return PK_OK;  // crop(src_dims, ext_offset) is the same-size convolution
	g.cw = 2 * mxd;
	g.ch = 2 * mxd;
	g.csx = -ox;
	g.csy = -oy;
	g.mx = ox;
	g.my = oy;
	g.offx = ox;
	g.offy = oy;
	return PK_OK;
}

template <typename T>
T *dptr(pk_ctx *ctx, size_t off) { return reinterpret_cast<T *>(static_cast<char *>(ctx->dmem) + off); }

template <typename T>
T *hptr(pk_ctx *ctx, size_t off) { return reinterpret_cast<T *>(static_cast<char *>(ctx->hmem) + off); }

uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Frames of Model::evaluate (src/model.cpp:286-318) for a job whose dw, dh
// are set: the convolver's result frame, the crop of the PSF padding inside
// it, the returned size after downsampling, and the offset evaluate(Point &)
// reports.
int frame_geometry(const pk_model *m, const Analysis &a, PkJob &jb, int &offx, int &offy)
{
	const uint32_t fs = m->finesampling;
	ConvGeom g{a.dw, a.dh, 0, 0, 0, 0, 0, 0};
	if (a.conv_required)
		PK_TRY(conv_geometry(m->convolver, a.dw, a.dh, m->psf_width, m->psf_height, m->crop != 0, g));
	jb.cw = a.conv_required ? g.cw : a.dw;
	jb.ch = a.conv_required ? g.ch : a.dh;
	jb.ctiles_x = cdiv(jb.cw, CT_W);
	jb.csx = g.csx;
	jb.wrap_y = g.wrap_y;
	jb.csy = g.csy;
	jb.cmx = g.mx;
	jb.cmy = g.my;
	jb.mox = a.conv_required ? g.offx : 0;
	jb.moy = a.conv_required ? g.offy : 0;
	offx = a.conv_required ? g.offx : 0;
	offy = a.conv_required ? g.offy : 0;
	int cox = 0, coy = 0;
	int crw = (int)jb.cw, crh = (int)jb.ch;
	if (a.padx > 0 && a.pady > 0) {
		cox = (int)a.padx;
		coy = (int)a.pady;
		crw = (int)a.dw - 2 * (int)a.padx;
		crh = (int)a.dh - 2 * (int)a.pady;
		if (!m->crop) {
			int pa[4], pi[4];
			const int ct = a.conv_required ? m->convolver : PK_CONV_BRUTE;
			conv_padding(ct, a.dw, a.dh, m->psf_width, m->psf_height, pa);
			conv_padding(ct, a.dw - 2 * a.padx, a.dh - 2 * a.pady, m->psf_width, m->psf_height, pi);
			const int odx = pa[0] - pi[0], ody = pa[1] - pi[1], ddx = pa[2] - pi[2], ddy = pa[3] - pi[3];
			crw = (int)jb.cw - 2 * (int)a.padx - (odx + ddx);
			crh = (int)jb.ch - 2 * (int)a.pady - (ody + ddy);
			cox += odx;
			coy += ody;
			offx -= odx;
			offy -= ody;
		}
	}
	if (crw <= 0 || crh <= 0) return set_err(PK_EINVAL, "empty image after cropping");
	jb.cox = cox;
	jb.coy = coy;
	jb.crw = (uint32_t)crw;
	jb.crh = (uint32_t)crh;
	jb.fs = 1;
	uint32_t ow = (uint32_t)crw, oh = (uint32_t)crh;
	if (fs > 1 && !m->return_finesampled) {
		jb.fs = fs;
		ow = cdiv(ow, fs);
		oh = cdiv(oh, fs);
		offx /= (int)fs;
		offy /= (int)fs;
	}
	jb.ow = ow;
	jb.oh = oh;
	return PK_OK;
}


size_t conv_lds_bytes(uint32_t kw, uint32_t rows, size_t elem)
{
	return (size_t)(CT_W + kw - 1) * (CT_H + rows - 1) * elem;
}

constexpr size_t LDS_LIMIT = 64 * 1024;

// PSF rows per convolution launch so that the source tile fits in LDS
// (all rows in one launch for every PSF up to ~41x41 in FP64).
int conv_rows_per_launch(uint32_t kw, uint32_t kh, size_t elem, uint32_t &rows)
{
	rows = kh;
	while (rows > 1 && conv_lds_bytes(kw, rows, elem) > LDS_LIMIT) rows = (rows + 1) / 2;
	if (conv_lds_bytes(kw, rows, elem) > LDS_LIMIT)
		return set_err(PK_EINVAL, "PSF width %u exceeds the direct convolution tile", kw);
	return PK_OK;
}

// Reversed, transposed copy of a kw x kh PSF: kt[k * kh + l] = krn(kw-1-k, kh-1-l).
template <typename T>
void psf_conv_layout(const double *krn, uint32_t kw, uint32_t kh, T *kt)
{
	for (uint32_t k = 0; k < kw; k++)
		for (uint32_t l = 0; l < kh; l++) kt[(size_t)k * kh + l] = (T)krn[(kw - 1 - k) + (size_t)(kh - 1 - l) * kw];
}

template <typename T>
int launch_conv(hipStream_t s, const PkJob *djobs, uint32_t njobs, uint32_t ntiles, const T *src, const T *kt,
                const uint8_t *masks, T *dst, uint32_t max_kw, uint32_t max_kh, const uint32_t *tile_job = nullptr,
                uint32_t tiles = PK_CONV_TILES_VALU, T *out_all = nullptr, const T *img_nc = nullptr,
                uint32_t tile_base = 0, const PkComp *comps = nullptr)
{
	if (tiles != PK_CONV_TILES_VALU && sizeof(T) == 4 && conv_mfma_lds_bytes((int)max_kw, (int)max_kh) <= LDS_LIMIT) {
		// FP32 on request: the banded matrix-core form, whole PSF in one launch
		// (also where the strip form was asked for: its tiling is the batch's)
		hipLaunchKernelGGL(conv_mfma_kernel, dim3(ntiles), dim3(256),
		                   conv_mfma_lds_bytes((int)max_kw, (int)max_kh), s, djobs, njobs,
		                   reinterpret_cast<const float *>(src), reinterpret_cast<const float *>(kt), masks,
		                   reinterpret_cast<float *>(dst), tile_job, reinterpret_cast<float *>(out_all),
		                   reinterpret_cast<const float *>(img_nc), tile_base, comps);
		PK_HIP(hipGetLastError());
		return PK_OK;
	}
	uint32_t rows;
	PK_TRY(conv_rows_per_launch(max_kw, max_kh, sizeof(T), rows));
	if (out_all && rows < max_kh) return set_err(PK_EINVAL, "fused composition needs a single convolution launch");
	const size_t lds = conv_lds_bytes(max_kw, rows, sizeof(T));
	for (uint32_t l0 = 0; l0 < max_kh; l0 += rows) {
		// the returned image is written directly only by a single launch
		const bool last = l0 + rows >= max_kh;
		hipLaunchKernelGGL(conv_kernel<T>, dim3(ntiles), dim3(256), lds, s, djobs, njobs, src, kt, masks, dst,
		                   (int)l0, (int)rows, tile_job, (l0 == 0 && last) ? out_all : (T *)nullptr, img_nc, tile_base,
		                   comps);
		PK_HIP(hipGetLastError());
	}
	return PK_OK;
}

// ---- matrix-core strips (launches in pk_strips.hip, pk_strips.h) ----------------
// Rows of a PSF of kh rows in the strip form: c blocks of 16 on the matrix
// cores and tm tail rows on the ALU (where kh exceeds a multiple of 16 by at
// most ST_TAIL rows)
void strip_row_form(uint32_t kh, uint32_t &c, uint32_t &tm)
{
	tm = kh > 16 && kh % 16 <= (uint32_t)ST_TAIL ? kh % 16 : 0u;
	c = tm ? kh / 16 : (kh + 15) / 16;
}

// ---- FFT convolution ---------------------------------------------------------
struct FftGeom {
	uint32_t P = 0, Q = 0;
	int lx = 0, ly = 0;
};

int fft_geometry(uint32_t sw, uint32_t sh, uint32_t kw, uint32_t kh, FftGeom &g)
{
	// smallest powers of two holding the linear convolution
	const uint32_t lw = sw + kw - 1, lh = sh + kh - 1;
	g.lx = 0;
	while ((1u << g.lx) < lw) g.lx++;
	g.ly = 0;
	while ((1u << g.ly) < lh) g.ly++;
	if (g.lx > FFT_MAX_LOG2 || g.ly > FFT_MAX_LOG2)
		return set_err(PK_EFFT, "FFT convolution of %ux%u with %ux%u needs %ux%u points, at most %d per axis", sw, sh,
		               kw, kh, 1u << g.lx, 1u << g.ly, 1 << FFT_MAX_LOG2);
	g.P = 1u << g.lx;
	g.Q = 1u << g.ly;
	return PK_OK;
}

// exp(-2 pi i t / FFT_TWIDDLES), t < FFT_TWIDDLES, in FP64 (host, once)
const double2 *fft_twiddles()
{
	static std::vector<double2> tw;
	static std::once_flag once;
	std::call_once(once, [] {
		tw.resize(FFT_TWIDDLES);
		for (int t = 0; t < FFT_TWIDDLES; t++) {
			const double a = -2 * PK_PI * t / FFT_TWIDDLES;
			tw[t] = make_double2(cos(a), sin(a));
		}
	});
	return tw.data();
}

// One FFT convolution of the job's source (src + jb.img_off, sw x sh = dw x
// dh) with the normalised kw x kh PSF (FP64, natural layout), into the job's
// conv buffer with its geometry; z: P x Q complex work array, tw: twiddles.
// z needs P * Q complex doubles plus 16 bytes (the source's peak).
template <typename TI>
int launch_fft_conv(hipStream_t s, const PkJob &jb, const PkJob *djob, const TI *src, const double *psf,
                    double krn_max, const uint8_t *masks, TI *dst, double2 *z, const double2 *tw)
{
	FftGeom g;
	PK_TRY(fft_geometry(jb.dw, jb.dh, jb.kw, jb.kh, g));
	const FftPlan px = fft_plan(g.lx), py = fft_plan(g.ly);
	const size_t n = (size_t)g.P * g.Q;
	const uint32_t grid = (uint32_t)std::min<size_t>(cdiv((uint32_t)std::min<size_t>(n, 1u << 30), 256), 8192);
	unsigned long long *mx = reinterpret_cast<unsigned long long *>(z + n);
	const size_t ns = (size_t)jb.dw * jb.dh;
	hipLaunchKernelGGL(fft_zero_kernel, dim3(1), dim3(1), 0, s, mx);
	hipLaunchKernelGGL(fft_maxabs_kernel<TI>, dim3((uint32_t)std::min<size_t>(cdiv((uint32_t)std::min<size_t>(ns, 1u << 30), 256), 1024)),
	                   dim3(256), 0, s, src + jb.img_off, ns, mx);
	hipLaunchKernelGGL(fft_pack_kernel<TI>, dim3(grid), dim3(256), 0, s, src + jb.img_off, jb.dw, jb.dh, psf, jb.kw,
	                   jb.kh, krn_max, (const unsigned long long *)mx, z, g.P, g.Q);
	auto pass = [&](bool inverse, bool rows) {
		const FftPlan &pl = rows ? px : py;
		const uint32_t nlines = rows ? g.Q : g.P;
		const uint32_t per = std::min<uint32_t>(std::max<uint32_t>(1, FFT_LDS_COMPLEX / (uint32_t)pl.n), nlines);
		const size_t ls = rows ? g.P : 1, es = rows ? 1 : g.P;
		if (inverse)
			hipLaunchKernelGGL(fft_pass_kernel<true>, dim3(cdiv(nlines, per)), dim3(256), 0, s, z, pl, nlines, per, ls,
			                   es, tw);
		else
			hipLaunchKernelGGL(fft_pass_kernel<false>, dim3(cdiv(nlines, per)), dim3(256), 0, s, z, pl, nlines, per,
			                   ls, es, tw);
	};
	pass(false, true);
	pass(false, false);
	hipLaunchKernelGGL(fft_mul_kernel, dim3(grid), dim3(256), 0, s, z, px, py, krn_max, (const unsigned long long *)mx);
	pass(true, false);
	pass(true, true);
	const uint32_t ng = (uint32_t)std::min<size_t>(cdiv((uint32_t)((size_t)jb.cw * jb.ch), 256), 8192);
	hipLaunchKernelGGL(fft_sample_kernel<TI>, dim3(std::max(1u, ng)), dim3(256), 0, s, djob, (const double2 *)z, g.P,
	                   masks, dst);
	PK_HIP(hipGetLastError());
	return PK_OK;
}

// capacity (cells) of each of the two queues between FP32 refinement passes
// Capacity (cells) of each FP32 refinement queue: the cells of one level
// handed to the next pass (two, used in turn) and the guard-band list.  A
// batch sizes them by its radial components (about 2048 cells of one level
// per component on the galaxy workloads; a full queue leaves cells on the
// wavefront's LIFO, so a short one costs speed, not results).
constexpr uint32_t R32_QCAP_SINGLE = 1u << 19;
inline uint32_t r32_qcap_batch(size_t n_comps)
{
	return (uint32_t)std::min<size_t>(std::max<size_t>(n_comps * 2048, 1u << 20), 1u << 26);
}

// Fixed-point scale of the FP32-mode refinement accumulators of one model:
// 2^60 over the summed total flux of its radial profiles (no pixel can hold
// more than a profile's total flux, so the sums stay below 2^63).
double fixed_point_scale(const pk_model *m)
{
	double f = 0;
	for (uint32_t c = 0; c < m->n_components; c++) {
		const int k = m->components[c].kind;
		if (k == PK_PROFILE_SKY || k == PK_PROFILE_NULL || k == PK_PROFILE_PSF) continue;
		f += std::pow(10, -0.4 * (m->components[c].mag - m->magzero));
	}
	if (!(f > 0) || !std::isfinite(f)) return 1.0;
	return std::ldexp(1.0, 60) / f;
}

// The FP32-mode refinement stores level, resolution and recursion depth in
// 8 bits each.
int check_fp32_limits(const pk_model *m)
{
	for (uint32_t c = 0; c < m->n_components; c++) {
		const pk_component &pc = m->components[c];
		const int k = pc.kind;
		if (k == PK_PROFILE_SKY || k == PK_PROFILE_NULL || k == PK_PROFILE_PSF || pc.adjust) continue;
		if (pc.resolution == 0 || pc.resolution > 255)
			return set_err(PK_EINVAL, "resolution %u: the FP32 path supports 1..255", pc.resolution);
	}
	return PK_OK;
}

// coarse pass + refinement of all radial profiles; FP32 mode uses the
// fixed-point wavefront-packed refinement, FP64 mode the exact one.
template <typename TI>
int launch_profiles(hipStream_t s, bool fp32, const PkJob *djobs, uint32_t njobs, uint32_t ntiles,
                    const PkComp *dcomp, const CompLite *dlite, const uint8_t *dmask, TI *img_nc, TI *img_c,
                    RefineEntry *dent, PkCounters *dcnt, unsigned long long *dacc, RCell *dq, uint32_t qcap,
                    uint32_t ref_blocks, uint32_t grid_main, uint32_t grid_last, bool any_comp, hipEvent_t ev_coarse,
                    const uint32_t *tile_job = nullptr, R32Tune tune = R32_TUNE_DEFAULT)
{
	if (fp32) {
		hipLaunchKernelGGL((coarse32_kernel<TI>), dim3(ntiles), dim3(256), 0, s, djobs, njobs, dcomp, dmask, img_nc,
		                   img_c, dent, dcnt, dacc, tile_job, (uint32_t *)nullptr);
	} else {
		hipLaunchKernelGGL((coarse_kernel<TI>), dim3(ntiles), dim3(256), 0, s, djobs, njobs, dcomp, dmask,
		                   img_nc, img_c, dent, dcnt, (unsigned long long *)nullptr, tile_job, (uint32_t *)nullptr);
	}
	PK_HIP(hipGetLastError());
	if (ev_coarse) PK_HIP(hipEventRecord(ev_coarse, s));
	if (!any_comp) return PK_OK;
	if (fp32) {
		// pass p reads queue (p - 1) & 1 and hands over into queue p & 1; pq
		// holds the guard-band sub-cells of the pass that just ran
		RCell *q0 = dq, *q1 = dq + qcap, *pq = dq + 2 * (size_t)qcap;
		const uint32_t gm = std::min(grid_main, ref_blocks), gl = std::min(grid_last, ref_blocks);
		const int np = (int)std::min<uint32_t>(std::max(tune.passes, R32_PASSES_MIN), R32_PASSES);
		for (int p = 0; p < np - 1; p++) {
			RCell *qin = (p & 1) ? q0 : q1, *qout = (p & 1) ? q1 : q0;
			if (p == 0)
				hipLaunchKernelGGL((refine32_kernel<0, false>), dim3(gm), dim3(64 * r32_waves_per_group(false, false)), 0,
				                   s, djobs, dcomp, dlite, dent,
				                   dcnt, (const RCell *)nullptr, qout, qcap, 0, dacc, tune, pq,
				                   (const double *)nullptr, (double *)nullptr);
			else
				hipLaunchKernelGGL((refine32_kernel<1, false>), dim3(gm), dim3(64 * r32_waves_per_group(false, false)), 0,
				                   s, djobs, dcomp, dlite, dent,
				                   dcnt, (const RCell *)qin, qout, qcap, p, dacc, tune, pq, (const double *)nullptr,
				                   (double *)nullptr);
			hipLaunchKernelGGL(r32_resolve_kernel<false>, dim3(R32_RESOLVE_GRID), dim3(256), 0, s, dcomp, dlite, dcnt,
			                   (const RCell *)pq, qout, qcap, p, dacc, (const double *)nullptr, (double *)nullptr);
		}
		hipLaunchKernelGGL((refine32_kernel<1, true>), dim3(gl), dim3(64), 0, s, djobs, dcomp, dlite, dent, dcnt,
		                   (const RCell *)((np & 1) ? q1 : q0), (RCell *)nullptr, qcap, np - 1, dacc,
		                   tune, (RCell *)nullptr, (const double *)nullptr, (double *)nullptr);
		hipLaunchKernelGGL(finalize32_kernel<TI>, dim3(1024), dim3(256), 0, s, djobs, dent, dcnt, dacc, img_nc, img_c);
	} else {
		hipLaunchKernelGGL((refine_kernel<TI, false>), dim3(ref_blocks), dim3(64), 0, s, djobs, dcomp, dent, dcnt,
		                   img_nc, img_c);
	}
	PK_HIP(hipGetLastError());
	return PK_OK;
}

// Core of Model::evaluate on the device for one model.
template <typename TI>
int evaluate_impl(pk_ctx *ctx, const pk_model *m, const Analysis &a, double *out, size_t out_capacity,
                  pk_result *res, bool device_out = false)
{
	const bool fp32 = ctx->precision == PK_FP32;
	const uint32_t fs = m->finesampling;
	const double sx = m->scale_x / fs, sy = m->scale_y / fs;

	// normalised PSF (Model::set_psf -> Image::normalize, src/image.cpp:208-221)
	std::vector<double> psf;
	if (has_psf(m)) {
		size_t pn = (size_t)m->psf_width * m->psf_height;
		psf.assign(m->psf, m->psf + pn);
		double sum = 0;
		for (double v : psf) sum += v;
		if (sum > 0)
			for (double &v : psf) v /= sum;
	}

	// calculation mask in the drawing frame (src/model.cpp:230-246, 265-285)
	std::vector<uint8_t> dmask;
	bool use_mask = false, adjusted = false;
	if (has_mask(m)) {
		PK_TRY(calc_mask(m, a, dmask, adjusted));
		use_mask = true;
	}

	// components: radial + sky go to the coarse/refine kernels (their FP64
	// set-up runs on the device, setup_kernel), psf profiles become rebinning
	// items, null profiles do nothing
	std::vector<SetupIn> comps;
	std::vector<uint32_t> comp_model_index;  // model component of comps[k]
	std::vector<PkPsfItem> items;
	bool any_nc = false;
	for (uint32_t c = 0; c < m->n_components; c++) {
		const pk_component &pc = m->components[c];
		if (pc.kind == PK_PROFILE_NULL) continue;
		if (pc.kind == PK_PROFILE_PSF) {
			items.push_back(psf_item(m, pc, a, sx, sy));
			if (!pc.convolve) any_nc = true;
			continue;
		}
		SetupIn in{};
		in.c = pc;
		in.magzero = m->magzero;
		in.sx = sx;
		in.sy = sy;
		in.img_sx = m->scale_x;
		in.img_sy = m->scale_y;
		in.pad_x = a.padx;
		in.pad_y = a.pady;
		in.fp32_mode = fp32 ? 1 : 0;
		in.fs = fs;
		in.fx_scale = fixed_point_scale(m);
		in.integrator = g_integrator.load();
		if (!pc.convolve) any_nc = true;
		comps.push_back(in);
		comp_model_index.push_back(c);
	}
	if (comps.size() > 32) return set_err(PK_EINVAL, "at most 32 radial/sky profiles per model are supported");

	// composition: crop padding (src/model.cpp:289-306), downsample, mask
	PkJob jb{};
	jb.dw = a.dw;
	jb.dh = a.dh;
	jb.comp_off = 0;
	jb.ncomp = (uint32_t)comps.size();
	jb.img_off = 0;
	jb.sx = sx;
	jb.sy = sy;
	jb.tile_off = 0;
	jb.tiles_x = cdiv(a.dw, 64);
	jb.has_conv = a.conv_required ? 1 : 0;
	jb.has_nc = (any_nc || !a.conv_required) ? 1 : 0;
	jb.kw = m->psf_width;
	jb.kh = m->psf_height;
	jb.ctile_off = 0;
	int offx = 0, offy = 0;
	PK_TRY(frame_geometry(m, a, jb, offx, offy));
	const uint32_t ow = jb.ow, oh = jb.oh;
	jb.out_off = 0;
	jb.otile_off = 0;
	jb.otiles_x = cdiv(ow, 64);
	jb.omask_off = NONE;
	jb.omask_w = m->mask_width;
	jb.omask_h = m->mask_height;
	jb.omask_f = 1;
	if (adjusted) jb.omask_f = (ow == m->mask_width) ? 1 : fs;
	if ((size_t)ow * oh > out_capacity)
		return set_err(PK_ERANGE, "output buffer holds %zu doubles, %zu needed", out_capacity, (size_t)ow * oh);

	// ---- memory layout: inputs (uploaded) then workspace ----
	const size_t npix = (size_t)a.dw * a.dh;
	const size_t ncpix = (size_t)jb.cw * jb.ch;
	Layout L;
	const size_t o_job = L.take(sizeof(PkJob));
	const size_t o_cnt = L.take(sizeof(PkCounters));
	const size_t o_setup = L.take(sizeof(SetupIn) * std::max<size_t>(1, comps.size()));
	const size_t o_items = L.take(sizeof(PkPsfItem) * std::max<size_t>(1, items.size()));
	const size_t o_psf64 = L.take(sizeof(double) * std::max<size_t>(1, psf.size()));
	const size_t o_psfT = L.take(sizeof(TI) * std::max<size_t>(1, psf.size()));
	const size_t o_mask = L.take(use_mask ? npix : 1);
	const size_t o_omask = L.take(adjusted ? (size_t)m->mask_width * m->mask_height : 1);
	const size_t o_hist = L.take(sizeof(unsigned long long) * 32);
	const size_t in_bytes = L.off;
	const size_t o_comp = L.take(sizeof(PkComp) * std::max<size_t>(1, comps.size()));
	const size_t o_lite = L.take(sizeof(CompLite) * std::max<size_t>(1, comps.size()));
	const size_t o_nc = L.take(sizeof(TI) * npix);
	const size_t o_c = L.take(sizeof(TI) * (a.conv_required ? npix : 1));
	const size_t o_conv = L.take(sizeof(TI) * (a.conv_required ? ncpix : 1));
	const size_t o_ent = L.take(sizeof(RefineEntry) * npix);
	const size_t o_acc = L.take(fp32 ? 2 * sizeof(unsigned long long) * npix : 8);
	const uint32_t qcap = fp32 ? R32_QCAP_SINGLE : 0;
	const size_t o_q = L.take(3 * sizeof(RCell) * std::max<uint32_t>(1, qcap));
	const size_t o_out = L.take(sizeof(double) * (size_t)ow * oh);
	const bool use_fft = a.conv_required && m->convolver == PK_CONV_FFT;
	FftGeom fg;
	if (use_fft) PK_TRY(fft_geometry(a.dw, a.dh, m->psf_width, m->psf_height, fg));
	const size_t o_fz = L.take(use_fft ? sizeof(double2) * ((size_t)fg.P * fg.Q + 1) : 1);
	const size_t o_ftw = L.take(use_fft ? sizeof(double2) * FFT_TWIDDLES : 1);
	PK_TRY(ensure_dev(ctx, L.off));
	PK_TRY(ensure_host(ctx, std::max(in_bytes, (device_out ? 0 : sizeof(double) * (size_t)ow * oh + 256) + 512 +
	                                               sizeof(PkCounters))));

	jb.mask_off = use_mask ? 0 : NONE;
	jb.psf_off = 0;
	jb.conv_off = 0;
	jb.fx_scale = fixed_point_scale(m);
	if (adjusted) jb.omask_off = 0;
	for (auto &it : items) it.psf_off = 0;

	// stage inputs
	std::memset(ctx->hmem, 0, in_bytes);
	*hptr<PkJob>(ctx, o_job) = jb;
	if (!comps.empty()) std::memcpy(hptr<void>(ctx, o_setup), comps.data(), sizeof(SetupIn) * comps.size());
	if (!items.empty()) std::memcpy(hptr<void>(ctx, o_items), items.data(), sizeof(PkPsfItem) * items.size());
	if (!psf.empty()) {
		std::memcpy(hptr<void>(ctx, o_psf64), psf.data(), sizeof(double) * psf.size());
		TI *pt = hptr<TI>(ctx, o_psfT);
		psf_conv_layout<TI>(psf.data(), m->psf_width, m->psf_height, pt);
	}
	if (use_mask) std::memcpy(hptr<void>(ctx, o_mask), dmask.data(), npix);
	if (adjusted) {
		uint8_t *om = hptr<uint8_t>(ctx, o_omask);
		for (size_t t = 0; t < (size_t)m->mask_width * m->mask_height; t++) om[t] = m->mask[t] ? 1 : 0;
	}

	hipStream_t s = ctx->stream;
	PK_HIP(hipMemcpyAsync(ctx->dmem, ctx->hmem, in_bytes, hipMemcpyHostToDevice, s));
	if (use_fft)
		PK_HIP(hipMemcpyAsync(dptr<void>(ctx, o_ftw), fft_twiddles(), sizeof(double2) * FFT_TWIDDLES,
		                      hipMemcpyHostToDevice, s));
	PK_HIP(hipEventRecord(ctx->ev[4], s));
	if (!comps.empty()) {
		// per-component FP64 set-up: bn root-solve, normalisation, adjusted
		// sub-pixel parameters
		hipLaunchKernelGGL(setup_kernel, dim3((uint32_t)comps.size()), dim3(64), 0, s, dptr<SetupIn>(ctx, o_setup),
		                   dptr<PkComp>(ctx, o_comp), dptr<CompLite>(ctx, o_lite), (uint32_t)comps.size());
		PK_HIP(hipGetLastError());
	}
	PK_HIP(hipEventRecord(ctx->ev[0], s));

	const PkJob *djob = dptr<PkJob>(ctx, o_job);
	PkCounters *dcnt = dptr<PkCounters>(ctx, o_cnt);
	TI *img_nc = dptr<TI>(ctx, o_nc);
	TI *img_c = dptr<TI>(ctx, o_c);
	const uint8_t *dm = dptr<uint8_t>(ctx, o_mask);
	RefineEntry *dent = dptr<RefineEntry>(ctx, o_ent);
	const PkComp *dcomp = dptr<PkComp>(ctx, o_comp);

	const uint32_t ntiles = jb.tiles_x * cdiv(a.dh, PX_ROWS);
	const uint32_t nref_blocks = (uint32_t)std::min<size_t>(npix, 2048);
	R32Tune r32 = R32_TUNE_DEFAULT;
	for (uint32_t c = 0; c < m->n_components; c++)
		r32.passes = std::max(r32.passes, refine_passes_needed(m->components[c]));
	PK_TRY(launch_profiles<TI>(s, fp32, djob, 1u, ntiles, dcomp, dptr<CompLite>(ctx, o_lite), dm, img_nc, img_c, dent, dcnt,
	                           dptr<unsigned long long>(ctx, o_acc), dptr<RCell>(ctx, o_q), qcap, nref_blocks,
	                           ctx->r32_grid_main, ctx->r32_grid_last,
	                           !comps.empty(), nullptr, nullptr, r32));
	if (!comps.empty()) {
		hipLaunchKernelGGL(integrations_kernel, dim3(64), dim3(256), 0, s, (const RefineEntry *)dent,
		                   (const PkCounters *)dcnt, dptr<unsigned long long>(ctx, o_hist));
		PK_HIP(hipGetLastError());
	}
	for (size_t k = 0; k < items.size(); k++) {
		const auto &it = items[k];
		if (it.x1 < it.x0 || it.y1 < it.y0) continue;
		const uint32_t n = (it.x1 - it.x0 + 1) * (it.y1 - it.y0 + 1);
		hipLaunchKernelGGL(psf_kernel<TI>, dim3(cdiv(n, 256), 1), dim3(256), 0, s,
		                   dptr<PkPsfItem>(ctx, o_items + k * sizeof(PkPsfItem)), djob, dptr<double>(ctx, o_psf64),
		                   img_nc, img_c);
		PK_HIP(hipGetLastError());
	}
	PK_HIP(hipEventRecord(ctx->ev[1], s));
	if (use_fft) {
		double kmax = 0;
		for (double v : psf) kmax = std::max(kmax, std::fabs(v));
		PK_TRY(launch_fft_conv<TI>(s, jb, djob, img_c, dptr<double>(ctx, o_psf64), kmax, dm, dptr<TI>(ctx, o_conv),
		                           dptr<double2>(ctx, o_fz), dptr<double2>(ctx, o_ftw)));
	} else if (a.conv_required) {
		const uint32_t nct = jb.ctiles_x * cdiv(jb.ch, CT_H);
		PK_TRY(launch_conv<TI>(s, djob, 1u, nct, img_c, dptr<TI>(ctx, o_psfT), dm, dptr<TI>(ctx, o_conv), jb.kw, jb.kh));
	}
	PK_HIP(hipEventRecord(ctx->ev[2], s));
	const uint32_t notiles = jb.otiles_x * cdiv(oh, PX_ROWS);
	hipLaunchKernelGGL((compose_kernel<TI, double>), dim3(notiles), dim3(256), 0, s, djob, 1u, img_nc,
	                   dptr<TI>(ctx, o_conv), dptr<uint8_t>(ctx, o_omask), dptr<double>(ctx, o_out),
	                   (const uint32_t *)nullptr);
	PK_HIP(hipGetLastError());
	PK_HIP(hipEventRecord(ctx->ev[3], s));

	// download result + counters
	const size_t out_bytes = sizeof(double) * (size_t)ow * oh;
	// counters after the staged image (host output), or at the start of the
	// staging area (device output: the input upload before it in stream order
	// has finished by then)
	const size_t ho_cnt = device_out ? 0 : (out_bytes + 255) & ~(size_t)255;
	if (device_out)
		PK_HIP(hipMemcpyAsync(out, dptr<void>(ctx, o_out), out_bytes, hipMemcpyDeviceToDevice, s));
	else
		PK_HIP(hipMemcpyAsync(ctx->hmem, dptr<void>(ctx, o_out), out_bytes, hipMemcpyDeviceToHost, s));
	PK_HIP(hipMemcpyAsync(hptr<void>(ctx, ho_cnt), dcnt, sizeof(PkCounters), hipMemcpyDeviceToHost, s));
	const size_t ho_hist = (ho_cnt + sizeof(PkCounters) + 7) & ~(size_t)7;
	PK_HIP(hipMemcpyAsync(hptr<void>(ctx, ho_hist), dptr<void>(ctx, o_hist), sizeof(unsigned long long) * 32,
	                      hipMemcpyDeviceToHost, s));
	PK_HIP(hipStreamSynchronize(s));
	const PkCounters cnt = *hptr<PkCounters>(ctx, ho_cnt);
	ctx->last_integrations.assign(m->n_components, 0);
	for (size_t k = 0; k < comp_model_index.size(); k++)
		ctx->last_integrations[comp_model_index[k]] = hptr<unsigned long long>(ctx, ho_hist)[k];
	if (cnt.overflow)
		return set_err(PK_ERANGE, "sub-pixel refinement work queue capacity exceeded");
	if (cnt.n_bad) return set_err(PK_ERANGE, "sub-pixel refinement produced %u non-finite values", cnt.n_bad);
	if (!device_out) std::memcpy(out, ctx->hmem, out_bytes);
	if (res) {
		res->width = ow;
		res->height = oh;
		res->offset_x = offx;
		res->offset_y = offy;
		res->n_refined_pixels = cnt.n_entries;
		res->n_subcells = cnt.n_subcells;
		float t40 = 0, t01 = 0, t12 = 0, t23 = 0, t03 = 0;
		(void)hipEventElapsedTime(&t40, ctx->ev[4], ctx->ev[0]);
		(void)hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]);
		(void)hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]);
		(void)hipEventElapsedTime(&t23, ctx->ev[2], ctx->ev[3]);
		(void)hipEventElapsedTime(&t03, ctx->ev[4], ctx->ev[3]);
		res->ms_setup = t40;
		res->ms_profiles = t01;
		res->ms_convolution = t12;
		res->ms_total = t03;
	}
	return PK_OK;
}

// Fills n doubles with v (grid-stride).
__global__ void fill_kernel(double *p, size_t n, double v)
{
	for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = v;
}

// ---- allocation-free single evaluation ---------------------------------------
// The workspace of pk_evaluate_device_noalloc: one array per device that
// loads this library (held until the process ends) and a host staging area.
// Host code uses the address hipGetSymbolAddress reports for the current
// device, never the array's name.
static __device__ __align__(256) unsigned char scratch[PK_NOALLOC_SCRATCH_BYTES];
alignas(256) static unsigned char g_noalloc_stage[PK_NOALLOC_STAGE_BYTES];
static std::mutex g_noalloc_mu;           // one allocation-free evaluation at a time in the process
static std::map<int, void *> g_scratch_ptr;  // device -> address of `scratch` on it

}  // namespace

extern "C" {

int pk_model_output_dims(const pk_model *m, uint32_t *width, uint32_t *height)
{
	if (!m) return set_err(PK_EINVAL, "model is NULL");
	Analysis a;
	PK_TRY(analyze(m, a));
	if (width) *width = a.dw;
	if (height) *height = a.dh;
	return PK_OK;
}

int pk_component_constants(const pk_model *m, uint32_t index, pk_constants *out)
{
	if (!m || !out) return set_err(PK_EINVAL, "NULL argument");
	Analysis a;
	PK_TRY(analyze(m, a));
	if (index >= m->n_components) return set_err(PK_EINVAL, "component index %u out of range", index);
	const pk_component &pc = m->components[index];
	if (pc.kind == PK_PROFILE_NULL || pc.kind == PK_PROFILE_SKY || pc.kind == PK_PROFILE_PSF)
		return set_err(PK_EINVAL, "component %u is not a radial profile", index);
	const uint32_t fs = m->finesampling;
	SetupIn in{};
	in.c = pc;
	in.magzero = m->magzero;
	in.sx = m->scale_x / fs;
	in.sy = m->scale_y / fs;
	in.img_sx = m->scale_x;
	in.img_sy = m->scale_y;
	in.pad_x = a.padx;
	in.pad_y = a.pady;
	in.fs = fs;
	in.grad = pc.kind == PK_PROFILE_SERSIC;
	in.integrator = g_integrator.load();
	PkComp c;
	pk::QagWork ws;
	setup_component(in, c, pk::SerialLanes{&ws});
	out->dbn_dnser = c.dbn;
	out->dlogflux_dre = c.glnf_re;
	out->dlogflux_dnser = c.glnf_n;
	out->dlogflux_daxrat = c.glnf_q;
	out->dlogflux_dbox = c.glnf_b;
	out->rscale = c.rscale;
	out->lumtot = c.lumtot;
	out->ie = c.ie;
	out->bn = c.bn;
	out->acc = c.acc;
	out->rscale_switch = c.rscale_switch;
	out->rscale_max = c.rscale_max;
	out->flux_scale = c.flux_scale;
	out->cos_ang = c.cos_ang;
	out->sin_ang = c.sin_ang;
	out->resolution = (uint32_t)c.res;
	out->max_recursions = (uint32_t)c.max_rec;
	return PK_OK;
}

int pk_evaluate(pk_ctx *ctx, const pk_model *m, double *out, size_t out_capacity, pk_result *res)
{
	if (!ctx) return set_err(PK_EINVAL, "context is NULL");
	if (!m) return set_err(PK_EINVAL, "model is NULL");
	if (!out && !m->dry_run) return set_err(PK_EINVAL, "output buffer is NULL");
	std::lock_guard<std::mutex> lock(ctx->mu);
	PK_HIP(hipSetDevice(ctx->device));
	Analysis a;
	PK_TRY(analyze(m, a));
	if (m->dry_run) {
		// src/model.cpp:263-267
		const size_t n = (size_t)a.dw * a.dh;
		if (n > out_capacity) return set_err(PK_ERANGE, "output buffer holds %zu doubles, %zu needed", out_capacity, n);
		if (out) std::memset(out, 0, n * sizeof(double));
		if (res) {
			std::memset(res, 0, sizeof(*res));
			res->width = a.dw;
			res->height = a.dh;
		}
		return PK_OK;
	}
	if (ctx->precision == PK_FP32) {
		PK_TRY(check_fp32_limits(m));
		return evaluate_impl<float>(ctx, m, a, out, out_capacity, res);
	}
	return evaluate_impl<double>(ctx, m, a, out, out_capacity, res);
}

int pk_evaluate_device_noalloc(pk_ctx *ctx, const pk_model *m, double *out_dev, size_t out_capacity,
                               pk_result *res)
{
	if (!ctx) return set_err(PK_EINVAL, "context is NULL");
	if (!m) return set_err(PK_EINVAL, "model is NULL");
	if (!out_dev) return set_err(PK_EINVAL, "output buffer is NULL");
	std::lock_guard<std::mutex> glock(g_noalloc_mu);
	std::lock_guard<std::mutex> lock(ctx->mu);
	PK_HIP(hipSetDevice(ctx->device));
	void *ws = nullptr;
	auto it = g_scratch_ptr.find(ctx->device);
	if (it == g_scratch_ptr.end()) {
		// first call on this device: the address of its copy of `scratch`
		const hipError_t e = hipGetSymbolAddress(&ws, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !ws)
			return set_err(PK_EDEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", ctx->device,
			               hipGetErrorString(e));
		g_scratch_ptr[ctx->device] = ws;
	} else {
		ws = it->second;
	}
	Analysis a;
	PK_TRY(analyze(m, a));
	if (m->dry_run) {
		// src/model.cpp:263-267 (a fill kernel, not a memset node: capturable)
		const size_t n = (size_t)a.dw * a.dh;
		if (n > out_capacity) return set_err(PK_ERANGE, "output buffer holds %zu doubles, %zu needed", out_capacity, n);
		if (n) hipLaunchKernelGGL(fill_kernel, dim3((uint32_t)std::min<size_t>(cdiv(n, 256), 65535)), dim3(256), 0,
		                          ctx->stream, out_dev, n, 0.0);
		PK_HIP(hipGetLastError());
		PK_HIP(hipStreamSynchronize(ctx->stream));
		if (res) {
			std::memset(res, 0, sizeof(*res));
			res->width = a.dw;
			res->height = a.dh;
		}
		return PK_OK;
	}
	if (ctx->precision == PK_FP32) PK_TRY(check_fp32_limits(m));
	// the context's own workspace is set aside for the call
	struct Swap {
		pk_ctx *c;
		void *dmem, *hmem;
		size_t dcap, hcap;
		~Swap()
		{
			c->dmem = dmem;
			c->dcap = dcap;
			c->hmem = hmem;
			c->hcap = hcap;
			c->fixed_ws = false;
		}
	} swap{ctx, ctx->dmem, ctx->hmem, ctx->dcap, ctx->hcap};
	ctx->dmem = ws;
	ctx->dcap = PK_NOALLOC_SCRATCH_BYTES;
	ctx->hmem = g_noalloc_stage;
	ctx->hcap = PK_NOALLOC_STAGE_BYTES;
	ctx->fixed_ws = true;
	if (ctx->precision == PK_FP32) return evaluate_impl<float>(ctx, m, a, out_dev, out_capacity, res, true);
	return evaluate_impl<double>(ctx, m, a, out_dev, out_capacity, res, true);
}

int pk_evaluate_device(pk_ctx *ctx, const pk_model *m, double *out_dev, size_t out_capacity, pk_result *res)
{
	if (!ctx) return set_err(PK_EINVAL, "context is NULL");
	if (!m) return set_err(PK_EINVAL, "model is NULL");
	if (!out_dev) return set_err(PK_EINVAL, "output buffer is NULL");
	std::lock_guard<std::mutex> lock(ctx->mu);
	PK_HIP(hipSetDevice(ctx->device));
	Analysis a;
	PK_TRY(analyze(m, a));
	if (m->dry_run) {
		// src/model.cpp:263-267
		const size_t n = (size_t)a.dw * a.dh;
		if (n > out_capacity) return set_err(PK_ERANGE, "output buffer holds %zu doubles, %zu needed", out_capacity, n);
		PK_HIP(hipMemsetAsync(out_dev, 0, n * sizeof(double), ctx->stream));
		PK_HIP(hipStreamSynchronize(ctx->stream));
		if (res) {
			std::memset(res, 0, sizeof(*res));
			res->width = a.dw;
			res->height = a.dh;
		}
		return PK_OK;
	}
	if (ctx->precision == PK_FP32) {
		PK_TRY(check_fp32_limits(m));
		return evaluate_impl<float>(ctx, m, a, out_dev, out_capacity, res, true);
	}
	return evaluate_impl<double>(ctx, m, a, out_dev, out_capacity, res, true);
}

int pk_convolver_padding(int type, uint32_t src_w, uint32_t src_h, uint32_t krn_w, uint32_t krn_h, int32_t *lo_x,
                         int32_t *lo_y, int32_t *hi_x, int32_t *hi_y)
{
	int p[4];
	conv_padding(type, src_w, src_h, krn_w, krn_h, p);
	if (lo_x) *lo_x = p[0];
	if (lo_y) *lo_y = p[1];
	if (hi_x) *hi_x = p[2];
	if (hi_y) *hi_y = p[3];
	return PK_OK;
}

}  // extern "C"

namespace {

template <typename T>
int convolve_impl(pk_ctx *ctx, int type, const double *src, uint32_t sw, uint32_t sh, const double *krn, uint32_t kw,
                  uint32_t kh, const uint8_t *mask, int crop, double *out, size_t cap, uint32_t *out_w,
                  uint32_t *out_h, int32_t *offset_x, int32_t *offset_y)
{
	ConvGeom g;
	PK_TRY(conv_geometry(type, sw, sh, kw, kh, crop != 0, g));
	if ((size_t)g.cw * g.ch > cap)
		return set_err(PK_ERANGE, "output buffer holds %zu doubles, %zu needed", cap, (size_t)g.cw * g.ch);
	PkJob jb{};
	jb.dw = sw;
	jb.dh = sh;
	jb.kw = kw;
	jb.kh = kh;
	jb.cw = g.cw;
	jb.ch = g.ch;
	jb.csx = g.csx;
	jb.wrap_y = g.wrap_y;
	jb.csy = g.csy;
	jb.cmx = g.mx;
	jb.cmy = g.my;
	jb.ctiles_x = cdiv(g.cw, CT_W);
	jb.mask_off = mask ? 0 : NONE;
	jb.has_conv = 1;
	jb.has_nc = 0;
	jb.cox = jb.coy = 0;
	jb.crw = g.cw;
	jb.crh = g.ch;
	jb.fs = 1;
	jb.ow = g.cw;
	jb.oh = g.ch;
	jb.otiles_x = cdiv(g.cw, 64);
	jb.omask_off = NONE;
	jb.omask_f = 1;
	const size_t ns = (size_t)sw * sh, nk = (size_t)kw * kh, no = (size_t)g.cw * g.ch;
	Layout L;
	const size_t o_job = L.take(sizeof(PkJob));
	const size_t o_src = L.take(sizeof(T) * ns);
	const size_t o_krn = L.take(sizeof(T) * nk);
	const size_t o_mask = L.take(mask ? ns : 1);
	const size_t in_bytes = L.off;
	const size_t o_conv = L.take(sizeof(T) * no);
	const size_t o_out = L.take(sizeof(double) * no);
	const bool use_fft = type == PK_CONV_FFT;
	FftGeom fg;
	if (use_fft) PK_TRY(fft_geometry(sw, sh, kw, kh, fg));
	const size_t o_k64 = L.take(use_fft ? sizeof(double) * nk : 1);
	const size_t o_fz = L.take(use_fft ? sizeof(double2) * ((size_t)fg.P * fg.Q + 1) : 1);
	const size_t o_ftw = L.take(use_fft ? sizeof(double2) * FFT_TWIDDLES : 1);
	PK_TRY(ensure_dev(ctx, L.off));
	PK_TRY(ensure_host(ctx, std::max(in_bytes, sizeof(double) * no)));
	*hptr<PkJob>(ctx, o_job) = jb;
	T *hs = hptr<T>(ctx, o_src);
	for (size_t t = 0; t < ns; t++) hs[t] = (T)src[t];
	T *hk = hptr<T>(ctx, o_krn);
	psf_conv_layout<T>(krn, kw, kh, hk);
	if (mask) {
		uint8_t *hm = hptr<uint8_t>(ctx, o_mask);
		for (size_t t = 0; t < ns; t++) hm[t] = mask[t] ? 1 : 0;
	}
	hipStream_t s = ctx->stream;
	PK_HIP(hipMemcpyAsync(ctx->dmem, ctx->hmem, in_bytes, hipMemcpyHostToDevice, s));
	const PkJob *djob = dptr<PkJob>(ctx, o_job);
	if (use_fft) {
		// the kernel as given (Convolver::convolve does not normalise it)
		PK_HIP(hipMemcpyAsync(dptr<void>(ctx, o_k64), krn, sizeof(double) * nk, hipMemcpyHostToDevice, s));
		PK_HIP(hipMemcpyAsync(dptr<void>(ctx, o_ftw), fft_twiddles(), sizeof(double2) * FFT_TWIDDLES,
		                      hipMemcpyHostToDevice, s));
		double kmax = 0;
		for (size_t t = 0; t < nk; t++) kmax = std::max(kmax, std::fabs(krn[t]));
		PK_TRY(launch_fft_conv<T>(s, jb, djob, dptr<T>(ctx, o_src), dptr<double>(ctx, o_k64), kmax,
		                          dptr<uint8_t>(ctx, o_mask), dptr<T>(ctx, o_conv), dptr<double2>(ctx, o_fz),
		                          dptr<double2>(ctx, o_ftw)));
	} else {
		const uint32_t nct = jb.ctiles_x * cdiv(g.ch, CT_H);
		PK_TRY(launch_conv<T>(s, djob, 1u, nct, dptr<T>(ctx, o_src), dptr<T>(ctx, o_krn), dptr<uint8_t>(ctx, o_mask),
		                      dptr<T>(ctx, o_conv), kw, kh));
	}
	hipLaunchKernelGGL((compose_kernel<T, double>), dim3(jb.otiles_x * cdiv(g.ch, PX_ROWS)), dim3(256), 0, s, djob, 1u,
	                   dptr<T>(ctx, o_conv), dptr<T>(ctx, o_conv), (const uint8_t *)nullptr,
	                   dptr<double>(ctx, o_out), (const uint32_t *)nullptr);
	PK_HIP(hipGetLastError());
	PK_HIP(hipMemcpyAsync(ctx->hmem, dptr<void>(ctx, o_out), sizeof(double) * no, hipMemcpyDeviceToHost, s));
	PK_HIP(hipStreamSynchronize(s));
	std::memcpy(out, ctx->hmem, sizeof(double) * no);
	if (out_w) *out_w = g.cw;
	if (out_h) *out_h = g.ch;
	if (offset_x) *offset_x = g.offx;
	if (offset_y) *offset_y = g.offy;
	return PK_OK;
}

}  // namespace

extern "C" int pk_convolve(pk_ctx *ctx, int type, const double *src, uint32_t src_w, uint32_t src_h,
                           const double *krn, uint32_t krn_w, uint32_t krn_h, const uint8_t *mask, int crop,
                           double *out, size_t out_capacity, uint32_t *out_w, uint32_t *out_h, int32_t *offset_x,
                           int32_t *offset_y)
{
	if (!ctx || !src || !krn || !out) return set_err(PK_EINVAL, "NULL argument");
	if (type != PK_CONV_BRUTE_OLD && type != PK_CONV_BRUTE && type != PK_CONV_OPENCL && type != PK_CONV_FFT)
		return set_err(PK_EINVAL, "Unsupported convolver type: %d", type);
	if (src_w == 0 || src_h == 0 || krn_w == 0 || krn_h == 0) return set_err(PK_EINVAL, "empty image or kernel");
	std::lock_guard<std::mutex> lock(ctx->mu);
	PK_HIP(hipSetDevice(ctx->device));
	if (ctx->precision == PK_FP32)
		return convolve_impl<float>(ctx, type, src, src_w, src_h, krn, krn_w, krn_h, mask, crop, out, out_capacity,
		                            out_w, out_h, offset_x, offset_y);
	return convolve_impl<double>(ctx, type, src, src_w, src_h, krn, krn_w, krn_h, mask, crop, out, out_capacity, out_w,
	                             out_h, offset_x, offset_y);
}

// ---------------------------------------------------------------------------
// special functions (host copies of the device set-up code)
// ---------------------------------------------------------------------------
extern "C" {

double pk_qgamma(double p, double shape) { return pk::qgamma(p, shape); }
double pk_pgamma(double q, double shape) { return pk::pgamma(q, shape); }
double pk_gammafn(double x) { return pk::gammafn(x); }
double pk_beta(double a, double b) { return pk::beta(a, b); }

double pk_integrate_qags(pk_integration_func_t f, double a, double b, void *params)
{
	pk::QagWork ws;
	return pk::integrate_adaptive([&](double x) { return f(x, params); }, a, b, g_integrator.load(),
	                              pk::SerialLanes{&ws});
}

double pk_integrate_qagi(pk_integration_func_t f, double a, void *params)
{
	pk::QagWork ws;
	return pk::integrate_to_inf([&](double x) { return f(x, params); }, a, g_integrator.load(),
	                            pk::SerialLanes{&ws});
}

int pk_set_integrator(int mode)
{
	if (mode != PK_INTEGRATOR_QUADPACK && mode != PK_INTEGRATOR_EXACT)
		return set_err(PK_EINVAL, "integrator %d: expected PK_INTEGRATOR_QUADPACK or PK_INTEGRATOR_EXACT", mode);
	g_integrator.store(mode);
	return PK_OK;
}

int pk_get_integrator(void) { return g_integrator.load(); }

}  // extern "C"

// ---------------------------------------------------------------------------
// batched, device-resident evaluation
// ---------------------------------------------------------------------------
static pk_tuning default_tuning()
{
	pk_tuning t{};
	t.conv_tiles = PK_CONV_TILES_VALU;
	t.refine_occupancy = 100;
	t.handover_div = R32_TUNE_DEFAULT.handover_div;
	t.claim_pixels = R32_TUNE_DEFAULT.chunk;
	t.claim_cells = R32_TUNE_DEFAULT.chunk_cells;
	t.refine_passes = R32_TUNE_DEFAULT.passes;
	t.claim_spread = 0;
	t.ms_evaluate = 0;
	return t;
}

struct pk_batch {
	pk_ctx *ctx = nullptr;
	uint32_t flags = 0;
	uint32_t n_models = 0, n_comps = 0;
	size_t elem = 8;  // bytes per image element (4: FP32 context)
	std::vector<PkJob> jobs;
	std::vector<SetupIn> setup;  // one per radial/sky component
	std::vector<uint32_t> comp_index;  // user component index -> setup slot (or ~0 for null)
	std::vector<int32_t> comp_kind;
	uint32_t ntiles = 0, nctiles = 0, notiles = 0, nstiles = 0;
	uint32_t min_passes = R32_PASSES_MIN;  // FP32 refinement launches its components need at least
	uint32_t max_kw = 0, max_kh = 0;
	uint32_t adj_kw = 0, adj_kh = 0;  // largest PSF of the adjoint correlations (FFT jobs included)
	bool any_conv = false;
	// device memory
	void *dmem = nullptr;
	size_t o_jobs = 0, o_setup = 0, o_comps = 0, o_cnt = 0, o_psfT = 0, o_mask = 0, o_nc = 0, o_c = 0, o_conv = 0,
	       o_ent = 0, o_acc = 0, o_q = 0, o_out = 0, o_lite = 0;
	size_t o_tj = 0;  // job of every tile: coarse, convolution, composition tiles, then convolution strip tiles
	size_t o_wev = 0; // evaluate_at calls of every coarse-pass wavefront (pk_batch_work_counts)
	size_t o_wsum = 0;
	// convolution tiles are numbered by PSF-size class, one launch per class
	// with the LDS its largest PSF needs (smaller PSFs: more workgroups per CU)
	struct ConvClass {
		uint32_t start, count, max_kw, max_kh;
		bool strips;  // its jobs are in strip-form classes too (PSF sides <= ST_KMAX)
	};
	std::vector<ConvClass> conv_classes;
	// the jobs with PSF sides <= ST_KMAX again, numbered in strip tiles by
	// strip-form class (PSF column steps nj, row blocks c): one launch each
	struct StripClass {
		uint32_t start, count, nj, c, tm;
	};
	std::vector<StripClass> strip_classes;
	// psf profiles: rebinning items (PsfProfile::evaluate), the model data their
	// geometry is recomputed from when the components change, and the
	// normalised FP64 PSFs they read
	struct PsfSrc {
		pk_model model;  // scalar fields only
		Analysis a;
		double sx, sy;
		uint32_t comp;  // user component index
	};
	std::vector<PkPsfItem> psf_items;  // layer-major: layer k holds the k-th psf profile of each model
	std::vector<PsfSrc> psf_src;       // same order
	struct PsfLayer {
		uint32_t start, count, max_pixels;
	};
	std::vector<PsfLayer> psf_layers;
	size_t o_items = 0, o_psfd = 0;
	// models convolved by the FFT convolver: one FFT convolution each, in
	// turn, through one shared transform buffer
	struct FftJob {
		uint32_t job;
		uint64_t psf_off;  // normalised FP64 PSF in o_psfd
		double krn_max;
	};
	std::vector<FftJob> fft_jobs;
	std::vector<uint8_t> is_fft;  // per model
	std::vector<int32_t> origin;  // per model: the offset Model::evaluate(Point &) reports
	size_t o_fz = 0, o_ftw = 0;
	// pixels the largest item of each layer covers (grid width of its launch)
	void psf_update_layers()
	{
		for (auto &ly : psf_layers) {
			ly.max_pixels = 0;
			for (uint32_t k = ly.start; k < ly.start + ly.count; k++) {
				const PkPsfItem &it = psf_items[k];
				if (it.x1 < it.x0 || it.y1 < it.y0) continue;
				ly.max_pixels = std::max(ly.max_pixels, (it.x1 - it.x0 + 1) * (it.y1 - it.y0 + 1));
			}
		}
	}
	// adjoint convolution (PK_BATCH_GRAD): jobs with the whole drawing frame as
	// convolved frame, and the job of each of its tiles
	size_t o_jobs_adj = 0, o_tja = 0;
	uint32_t natiles = 0;
	// the same correlations in strip tiles (when every convolved job's PSF
	// sides are <= ST_KMAX), their job table after the adjoint tiles' in o_tja
	bool adj_strips = false;
	uint32_t nastiles = 0;
	std::vector<StripClass> adj_strip_classes;
	uint32_t qcap = 0;
	std::vector<double> fx_scales;
	uint64_t total_out = 0;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	hipEvent_t ev[7] = {};
	bool ran = false;
	// chi^2 adjoint (PK_BATCH_GRAD)
	size_t o_psfF = 0, o_gd = 0, o_gc = 0, o_data = 0, o_ivar = 0, o_gacc = 0, o_gout = 0, o_chi2 = 0, o_gcnt = 0;
	bool have_data = false, ran_grad = false;
	hipGraph_t graph_grad = nullptr;
	hipGraphExec_t exec_grad = nullptr;
	hipEvent_t ev_grad[3] = {};
	// launch settings (autotuner)
	pk_tuning tune = default_tuning();
	bool tuned = false;
	uint64_t n_pixels = 0;  // drawing pixels of all models
	R32Tune r32() const
	{
		const uint32_t spread_log2 = tune.claim_spread ? (uint32_t)__builtin_ctz(tune.claim_spread) : 0u;
		return R32Tune{tune.handover_div, tune.claim_pixels, tune.claim_cells,
		               std::max(tune.refine_passes, min_passes), spread_log2};
	}
	uint32_t grid(uint32_t full) const { return std::max(1u, (uint32_t)((uint64_t)full * tune.refine_occupancy / 100)); }
	template <typename T> T *d(size_t off) const
	{
		return reinterpret_cast<T *>(static_cast<char *>(dmem) + off);
	}
};

namespace {

// psf profiles of a batch, after the radial profiles: one launch per layer
// (the k-th psf profile of every model), so that the additions into one
// model's image keep the model's component order, as a single evaluation.
template <typename T>
int launch_psf_layers(pk_batch *b, hipStream_t s, T *nc, T *c)
{
	for (const auto &ly : b->psf_layers) {
		if (!ly.max_pixels) continue;
		for (uint32_t k0 = 0; k0 < ly.count; k0 += 65535u) {
			const uint32_t nk = std::min(65535u, ly.count - k0);
			hipLaunchKernelGGL(psf_kernel<T>, dim3(cdiv(ly.max_pixels, 256), nk), dim3(256), 0, s,
			                   b->d<PkPsfItem>(b->o_items) + ly.start + k0, b->d<PkJob>(b->o_jobs),
			                   b->d<double>(b->o_psfd), nc, c);
		}
	}
	PK_HIP(hipGetLastError());
	return PK_OK;
}

// FFT convolutions of a batch, one model after the other through the shared
// transform buffer, into each model's region of the convolution buffer.
template <typename T>
int launch_fft_jobs(pk_batch *b, hipStream_t s, const T *c)
{
	for (const auto &fj : b->fft_jobs)
		PK_TRY(launch_fft_conv<T>(s, b->jobs[fj.job], b->d<PkJob>(b->o_jobs) + fj.job, c,
		                          b->d<double>(b->o_psfd) + fj.psf_off, fj.krn_max, b->d<uint8_t>(b->o_mask),
		                          b->d<T>(b->o_conv), b->d<double2>(b->o_fz), b->d<double2>(b->o_ftw)));
	return PK_OK;
}

int batch_launch(pk_batch *b)
{
	pk_ctx *ctx = b->ctx;
	hipStream_t s = ctx->stream;
	const bool fp32 = ctx->precision == PK_FP32;
	const bool timed = !(b->flags & PK_BATCH_GRAPH);
	const PkJob *djobs = b->d<PkJob>(b->o_jobs);
	PkCounters *dcnt = b->d<PkCounters>(b->o_cnt);
	PkComp *dcomp = b->d<PkComp>(b->o_comps);
	const uint8_t *dmask = b->d<uint8_t>(b->o_mask);
	RefineEntry *dent = b->d<RefineEntry>(b->o_ent);
	const uint32_t nj = b->n_models;
	const uint32_t *tj_coarse = b->d<uint32_t>(b->o_tj), *tj_conv = tj_coarse + b->ntiles,
	               *tj_out = tj_conv + b->nctiles, *tj_strip = tj_out + b->notiles;
	if (timed) PK_HIP(hipEventRecord(b->ev[0], s));
	hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, dcnt);
	if (b->n_comps)
		hipLaunchKernelGGL(setup_kernel, dim3(b->n_comps), dim3(64), 0, s, b->d<SetupIn>(b->o_setup), dcomp,
		                   fp32 ? b->d<CompLite>(b->o_lite) : (CompLite *)nullptr,
		                   b->n_comps);
	PK_HIP(hipGetLastError());
	if (timed) PK_HIP(hipEventRecord(b->ev[1], s));
	if (fp32) {
		float *nc = b->d<float>(b->o_nc), *c = b->d<float>(b->o_c);
		PK_TRY(launch_profiles<float>(s, true, djobs, nj, b->ntiles, dcomp, b->d<CompLite>(b->o_lite), dmask, nc, c,
		                              dent, dcnt,
		                              b->d<unsigned long long>(b->o_acc), b->d<RCell>(b->o_q), b->qcap, UINT32_MAX,
		                              b->grid(ctx->r32_grid_main), b->grid(ctx->r32_grid_last),
		                              b->n_comps > 0, timed ? b->ev[2] : nullptr, tj_coarse, b->r32()));
		PK_TRY(launch_psf_layers<float>(b, s, nc, c));
		if (timed) PK_HIP(hipEventRecord(b->ev[3], s));
		if (b->any_conv) {
			const bool strips = b->tune.conv_tiles == PK_CONV_TILES_STRIPS;
			// the strip classes (disjoint jobs) on two streams, every other
			// class beside the first stream's, so that one launch's last
			// workgroups overlap the next launch's first
			const bool two = strips && b->strip_classes.size() > 1;
			if (two) {
				PK_HIP(hipEventRecord(b->ctx->fork, s));
				PK_HIP(hipStreamWaitEvent(b->ctx->side, b->ctx->fork, 0));
			}
			if (strips)
				for (size_t k = 0; k < b->strip_classes.size(); k++) {
					const auto &sc = b->strip_classes[k];
					const StripLaunch a{djobs, nj, sc.count, sc.start, c, b->d<float>(b->o_psfT), dmask,
					                    b->d<float>(b->o_conv), tj_strip, b->d<float>(b->o_out), nc, dcomp};
					if (!launch_strips(two && (k & 1) ? b->ctx->side : s, sc.nj, sc.c, sc.tm, a))
						return set_err(PK_EINVAL, "no strip convolution for %u x %u + %u PSF steps", sc.nj, sc.c,
						               sc.tm);
				}
			if (two) {
				PK_HIP(hipEventRecord(b->ctx->join, b->ctx->side));
				PK_HIP(hipStreamWaitEvent(s, b->ctx->join, 0));
			}
			for (const auto &cc : b->conv_classes)
				if (!(strips && cc.strips))
					PK_TRY(launch_conv<float>(s, djobs, nj, cc.count, c, b->d<float>(b->o_psfT), dmask,
					                          b->d<float>(b->o_conv), cc.max_kw, cc.max_kh, tj_conv,
					                          b->tune.conv_tiles, b->d<float>(b->o_out), nc, cc.start, dcomp));
			PK_HIP(hipGetLastError());
			PK_TRY(launch_fft_jobs<float>(b, s, c));
		}
		if (timed) PK_HIP(hipEventRecord(b->ev[4], s));
		if (b->notiles)
			hipLaunchKernelGGL((compose_kernel<float, float>), dim3(b->notiles), dim3(256), 0, s, djobs, nj, nc,
			                   b->d<float>(b->o_conv), dmask, b->d<float>(b->o_out), tj_out);
	} else {
		double *nc = b->d<double>(b->o_nc), *c = b->d<double>(b->o_c);
		PK_TRY(launch_profiles<double>(s, false, djobs, nj, b->ntiles, dcomp, nullptr, dmask, nc, c, dent, dcnt, nullptr,
		                               nullptr,
		                               0, 4096, 0, 0, b->n_comps > 0, timed ? b->ev[2] : nullptr, tj_coarse));
		PK_TRY(launch_psf_layers<double>(b, s, nc, c));
		if (timed) PK_HIP(hipEventRecord(b->ev[3], s));
		if (b->any_conv) {
			for (const auto &cc : b->conv_classes)
				PK_TRY(launch_conv<double>(s, djobs, nj, cc.count, c, b->d<double>(b->o_psfT), dmask,
				                           b->d<double>(b->o_conv), cc.max_kw, cc.max_kh, tj_conv, PK_CONV_TILES_VALU,
				                           b->d<double>(b->o_out), nc, cc.start, dcomp));
			PK_HIP(hipGetLastError());
			PK_TRY(launch_fft_jobs<double>(b, s, c));
		}
		if (timed) PK_HIP(hipEventRecord(b->ev[4], s));
		if (b->notiles)
			hipLaunchKernelGGL((compose_kernel<double, double>), dim3(b->notiles), dim3(256), 0, s, djobs, nj, nc,
			                   b->d<double>(b->o_conv), dmask, b->d<double>(b->o_out), tj_out);
	}
	PK_HIP(hipGetLastError());
	if (timed) PK_HIP(hipEventRecord(b->ev[5], s));
	return PK_OK;
}

// The adjoint part of pk_batch_evaluate_grad, after batch_launch: residuals
// and chi^2 -> adjoint drawing images -> PSF correlation -> coarse and
// refinement adjoints -> gradient slots.
template <typename T>
int grad_launch_t(pk_batch *b)
{
	pk_ctx *ctx = b->ctx;
	hipStream_t s = ctx->stream;
	const bool timed = !(b->flags & PK_BATCH_GRAPH);
	const PkJob *djobs = b->d<PkJob>(b->o_jobs);
	PkComp *dcomp = b->d<PkComp>(b->o_comps);
	PkCounters *gcnt = b->d<PkCounters>(b->o_gcnt);
	double *gacc = b->d<double>(b->o_gacc), *chi2 = b->d<double>(b->o_chi2);
	const uint32_t nj = b->n_models;
	T *Gd = b->d<T>(b->o_gd), *Gc = b->d<T>(b->o_gc);
	if (timed) PK_HIP(hipEventRecord(b->ev_grad[0], s));
	const uint32_t nacc = b->n_comps * GA_SLOTS;
	hipLaunchKernelGGL(reset_grad_kernel, dim3(std::max(1u, std::min(256u, cdiv(std::max(nacc, nj), 256)))),
	                   dim3(256), 0, s, gcnt, (const PkCounters *)b->d<PkCounters>(b->o_cnt), gacc, nacc, chi2, nj);
	const uint32_t *tj_coarse = b->d<uint32_t>(b->o_tj);
	hipLaunchKernelGGL(resid_kernel<T>, dim3(b->ntiles), dim3(256), 0, s, djobs, nj, (const T *)b->d<T>(b->o_out),
	                   (const T *)b->d<T>(b->o_data), (const T *)b->d<T>(b->o_ivar), Gd, chi2, tj_coarse,
	                   (const uint8_t *)b->d<uint8_t>(b->o_mask));
	PK_HIP(hipGetLastError());
	if (b->any_conv) {
		if (sizeof(T) == 4 && b->adj_strips && b->tune.conv_tiles == PK_CONV_TILES_STRIPS) {
			for (const auto &sc : b->adj_strip_classes) {
				const StripLaunch a{b->d<PkJob>(b->o_jobs_adj), nj, sc.count, sc.start,
				                    reinterpret_cast<const float *>(Gd), b->d<float>(b->o_psfF),
				                    b->d<uint8_t>(b->o_mask), reinterpret_cast<float *>(Gc),
				                    b->d<uint32_t>(b->o_tja) + b->natiles, nullptr, nullptr, nullptr};
				if (!launch_strips(s, sc.nj, sc.c, sc.tm, a))
					return set_err(PK_EINVAL, "no strip correlation for %u x %u + %u PSF steps", sc.nj, sc.c, sc.tm);
			}
		} else {
			PK_TRY(launch_conv<T>(s, b->d<PkJob>(b->o_jobs_adj), nj, b->natiles, Gd, b->d<T>(b->o_psfF),
			                      b->d<uint8_t>(b->o_mask), Gc, b->adj_kw, b->adj_kh, b->d<uint32_t>(b->o_tja),
			                      b->tune.conv_tiles));
		}
	}
	const bool fp32 = sizeof(T) == 4;
	hipLaunchKernelGGL((coarse_grad_kernel<T, fp32>), dim3(b->ntiles), dim3(256), 0, s, djobs, nj,
	                   (const PkComp *)dcomp, (const uint8_t *)b->d<uint8_t>(b->o_mask), (const T *)Gd,
	                   (const T *)Gc, gacc, tj_coarse);
	PK_HIP(hipGetLastError());
	const RefineEntry *dent = b->d<RefineEntry>(b->o_ent);
	if (fp32) {
		const CompLite *dlite = b->d<CompLite>(b->o_lite);
		RCell *q0 = b->d<RCell>(b->o_q), *q1 = q0 + b->qcap;
		const uint32_t gm = b->grid(ctx->r32g_grid_main), gl = b->grid(ctx->r32g_grid_last);
		const R32Tune hd = b->r32();
		// per-entry adjoint weights, in the fixed-point accumulators of the
		// forward pass (free once it has finished; 2 doubles per entry)
		double *gent = b->d<double>(b->o_acc);
		hipLaunchKernelGGL(grad_entry_weights_kernel<T>, dim3(1024), dim3(256), 0, s, djobs, dent,
		                   (const PkCounters *)b->d<PkCounters>(b->o_cnt), (const T *)Gd, (const T *)Gc, gent);
		PK_HIP(hipGetLastError());
		if (timed) PK_HIP(hipEventRecord(b->ev_grad[2], s));
		RCell *pq = q0 + 2 * (size_t)b->qcap;
		const int np = (int)std::min<uint32_t>(std::max(hd.passes, R32_PASSES_MIN), R32_PASSES);
		for (int p = 0; p < np - 1; p++) {
			RCell *qin = (p & 1) ? q0 : q1, *qout = (p & 1) ? q1 : q0;
			if (p == 0)
				hipLaunchKernelGGL((refine32_kernel<0, false, true>), dim3(gm), dim3(64), 0, s, djobs, dcomp, dlite,
				                   dent, gcnt, (const RCell *)nullptr, qout, b->qcap, 0,
				                   (unsigned long long *)nullptr, hd, pq, (const double *)gent, gacc);
			else
				hipLaunchKernelGGL((refine32_kernel<1, false, true>), dim3(gm), dim3(64), 0, s, djobs, dcomp, dlite,
				                   dent, gcnt, (const RCell *)qin, qout, b->qcap, p, (unsigned long long *)nullptr,
				                   hd, pq, (const double *)gent, gacc);
			hipLaunchKernelGGL(r32_resolve_kernel<true>, dim3(R32_RESOLVE_GRID), dim3(256), 0, s, dcomp, dlite, gcnt,
			                   (const RCell *)pq, qout, b->qcap, p, (unsigned long long *)nullptr, (const double *)gent,
			                   gacc);
		}
		hipLaunchKernelGGL((refine32_kernel<1, true, true>), dim3(gl), dim3(64), 0, s, djobs, dcomp, dlite, dent, gcnt,
		                   (const RCell *)((np & 1) ? q1 : q0), (RCell *)nullptr, b->qcap, np - 1,
		                   (unsigned long long *)nullptr, hd, (RCell *)nullptr, (const double *)gent, gacc);
	} else {
		if (timed) PK_HIP(hipEventRecord(b->ev_grad[2], s));
		hipLaunchKernelGGL((refine_kernel<double, false, true>), dim3(4096), dim3(64), 0, s, djobs,
		                   (const PkComp *)dcomp, dent, gcnt, (double *)nullptr, (double *)nullptr,
		                   (const double *)Gd, (const double *)Gc, gacc);
	}
	PK_HIP(hipGetLastError());
	hipLaunchKernelGGL(grad_finalize_kernel, dim3(cdiv(std::max(1u, b->n_comps), 64)), dim3(64), 0, s,
	                   (const PkComp *)dcomp, (const double *)gacc, b->d<double>(b->o_gout), b->n_comps);
	PK_HIP(hipGetLastError());
	if (timed) PK_HIP(hipEventRecord(b->ev_grad[1], s));
	return PK_OK;
}

int grad_launch(pk_batch *b)
{
	PK_TRY(batch_launch(b));
	return b->elem == 4 ? grad_launch_t<float>(b) : grad_launch_t<double>(b);
}

// ---- autotuner ------------------------------------------------------------------
std::string tune_key(const pk_batch *b)
{
	char k[512];
	std::snprintf(k, sizeof k, "%s fp%d grad%d models%u pixels%llu psf%ux%u comps%u", b->ctx->arch.c_str(),
	              b->elem == 4 ? 32 : 64, (b->flags & PK_BATCH_GRAD) ? 1 : 0, b->n_models,
	              (unsigned long long)b->n_pixels, b->max_kw, b->max_kh, b->n_comps);
	return k;
}

void drop_graphs(pk_batch *b)
{
	if (b->exec) (void)hipGraphExecDestroy(b->exec);
	if (b->graph) (void)hipGraphDestroy(b->graph);
	if (b->exec_grad) (void)hipGraphExecDestroy(b->exec_grad);
	if (b->graph_grad) (void)hipGraphDestroy(b->graph_grad);
	b->exec = nullptr;
	b->graph = nullptr;
	b->exec_grad = nullptr;
	b->graph_grad = nullptr;
}

// One eager evaluation with settings t; stage[] = device ms of set-up,
// coarse, refinement, convolution, composition, total.  +inf stage times
// when the refinement queue overflowed with these settings.
// One eager timed evaluation with settings t: the stage times (set-up, coarse,
// refinement, convolution, composition, total).  A gradient batch with data
// runs its adjoint as well, and the total includes it.
int tune_run(pk_batch *b, const pk_tuning &t, float stage[6])
{
	b->tune = t;
	const uint32_t flags = b->flags;
	const bool adjoint = (flags & PK_BATCH_GRAD) && b->have_data;
	b->flags &= ~(uint32_t)PK_BATCH_GRAPH;  // eager launches record the stage events
	int st = batch_launch(b);
	if (st == PK_OK && adjoint) st = b->elem == 4 ? grad_launch_t<float>(b) : grad_launch_t<double>(b);
	b->flags = flags;
	PK_TRY(st);
	PK_HIP(hipStreamSynchronize(b->ctx->stream));
	for (int k = 0; k < 5; k++) PK_HIP(hipEventElapsedTime(&stage[k], b->ev[k], b->ev[k + 1]));
	PK_HIP(hipEventElapsedTime(&stage[5], b->ev[0], b->ev[5]));
	PkCounters cnt{};
	PK_HIP(hipMemcpy(&cnt, b->d<void>(b->o_cnt), sizeof(cnt), hipMemcpyDeviceToHost));
	bool overflow = cnt.overflow != 0;
	if (adjoint) {
		float ms = 0;
		PK_HIP(hipEventElapsedTime(&ms, b->ev_grad[0], b->ev_grad[1]));
		stage[5] += ms;
		PK_HIP(hipMemcpy(&cnt, b->d<void>(b->o_gcnt), sizeof(cnt), hipMemcpyDeviceToHost));
		overflow = overflow || cnt.overflow != 0;
	}
	if (overflow)
		for (int k = 0; k < 6; k++) stage[k] = INFINITY;
	return PK_OK;
}

// Coordinate search: for each setting in turn, every candidate value is run
// five times (the fastest run counts) with the others at their best so far, timed
// on the stage that setting acts on.  The last run uses the chosen settings,
// so the batch holds its images.
int autotune(pk_batch *b)
{
	const bool fp32 = b->elem == 4;
	float st[6];
	pk_tuning best = default_tuning();
	PK_TRY(tune_run(b, best, st));  // first launches load the code objects
	auto measure = [&](const pk_tuning &t, int stage, float &ms) -> int {
		ms = INFINITY;
		for (int rep = 0; rep < 5; rep++) {
			float a[6];
			PK_TRY(tune_run(b, t, a));
			ms = std::min(ms, a[stage]);
		}
		return PK_OK;
	};
	auto search = [&](uint32_t pk_tuning::*field, std::initializer_list<uint32_t> values, int stage) -> int {
		float best_ms = INFINITY;
		uint32_t best_v = best.*field;
		for (uint32_t v : values) {
			pk_tuning t = best;
			t.*field = v;
			float ms;
			PK_TRY(measure(t, stage, ms));
			if (ms < best_ms) {
				best_ms = ms;
				best_v = v;
			}
		}
		best.*field = best_v;
		return PK_OK;
	};
	// a gradient batch with data is timed on forward + adjoint (the settings
	// act on both), any other on the stage a setting acts on
	const bool whole = (b->flags & PK_BATCH_GRAD) && b->have_data;
	const int s_conv = whole ? 5 : 3, s_ref = whole ? 5 : 2;
	if (fp32 && b->any_conv) {
		const bool band = conv_mfma_lds_bytes((int)b->max_kw, (int)b->max_kh) <= LDS_LIMIT;
		if (band && b->nstiles)
			PK_TRY(search(&pk_tuning::conv_tiles, {PK_CONV_TILES_VALU, PK_CONV_TILES_MFMA, PK_CONV_TILES_STRIPS},
			              s_conv));
		else if (band)
			PK_TRY(search(&pk_tuning::conv_tiles, {PK_CONV_TILES_VALU, PK_CONV_TILES_MFMA}, s_conv));
		else if (b->nstiles)
			PK_TRY(search(&pk_tuning::conv_tiles, {PK_CONV_TILES_VALU, PK_CONV_TILES_STRIPS}, s_conv));
	}
	// the refinement settings interact (claim sizes and spreading): two sweeps
	for (int sweep = 0; fp32 && b->n_comps && sweep < 2; sweep++) {
		PK_TRY(search(&pk_tuning::refine_occupancy, {50, 75, 100}, s_ref));
		PK_TRY(search(&pk_tuning::claim_pixels, {32, 64, 128, 256, 512}, s_ref));
		PK_TRY(search(&pk_tuning::claim_cells, {16, 64, 128, 256, 512}, s_ref));
		PK_TRY(search(&pk_tuning::claim_spread, {0, 4, 16, 64}, s_ref));
		PK_TRY(search(&pk_tuning::refine_passes, {4, 5, 6, 7}, s_ref));
	}
	PK_TRY(tune_run(b, best, st));
	best.ms_evaluate = st[5];
	b->tune = best;
	b->tuned = true;
	drop_graphs(b);
	std::lock_guard<std::mutex> lock(g_tune_mu);
	g_tune_cache[tune_key(b)] = best;
	g_tune_dirty = true;
	return PK_OK;
}

// Settings for the first evaluation of a PK_BATCH_AUTOTUNE batch: the cached
// ones for this device and workload, else a tuning run.
int ensure_tuned(pk_batch *b)
{
	if (b->tuned || !(b->flags & PK_BATCH_AUTOTUNE)) return PK_OK;
	{
		std::lock_guard<std::mutex> lock(g_tune_mu);
		auto it = g_tune_cache.find(tune_key(b));
		if (it != g_tune_cache.end()) {
			b->tune = it->second;
			b->tuned = true;
			return PK_OK;
		}
	}
	return autotune(b);
}

// Rows of the strip tiles of an image of ch rows: the fewest tiles of at most
// ST_H rows, of equal height (a short last tile would end a launch on it).
uint32_t strip_rows(uint32_t ch)
{
	const uint32_t n = std::max(1u, cdiv(ch, (uint32_t)ST_H));
	return std::max(1u, cdiv(ch, n));
}

// A job the strip form convolves: direct (not FFT) convolution, PSF sides <= ST_KMAX.
bool strip_job(const pk_batch *b, const PkJob &jb)
{
	return jb.has_conv && !b->is_fft[&jb - b->jobs.data()] && jb.kw <= (uint32_t)ST_KMAX &&
	       jb.kh <= (uint32_t)ST_KMAX;
}

}  // namespace

extern "C" {

int pk_batch_destroy(pk_batch *b)
{
	if (!b) return PK_OK;
	(void)hipSetDevice(b->ctx->device);
	(void)hipStreamSynchronize(b->ctx->stream);
	drop_graphs(b);
	for (auto &e : b->ev_grad)
		if (e) (void)hipEventDestroy(e);
	for (auto &e : b->ev)
		if (e) (void)hipEventDestroy(e);
	if (b->dmem) (void)hipFree(b->dmem);
	delete b;
	return PK_OK;
}

int pk_batch_create(pk_ctx *ctx, const pk_model *models, uint32_t n_models, uint32_t flags, pk_batch **out)
{
	if (!ctx || !out || (!models && n_models)) return set_err(PK_EINVAL, "NULL argument");
	if (n_models == 0) return set_err(PK_EINVAL, "empty batch");
	if (flags & ~(uint32_t)(PK_BATCH_GRAPH | PK_BATCH_GRAD | PK_BATCH_AUTOTUNE))
		return set_err(PK_EINVAL, "unknown batch flags 0x%x", flags);
	std::lock_guard<std::mutex> lock(ctx->mu);
	PK_HIP(hipSetDevice(ctx->device));
	const bool fp32 = ctx->precision == PK_FP32;
	const size_t elem = fp32 ? 4 : 8;
	pk_batch *b = new pk_batch();
	b->ctx = ctx;
	b->flags = flags;
	b->n_models = n_models;
	b->elem = elem;
	std::vector<double> psf_all;
	std::vector<uint8_t> mask_all;
	// normalised FP64 PSFs (Model::set_psf) of models with psf profiles or
	// the FFT convolver, one per model
	std::vector<double> psfd_all;
	uint32_t psfd_model = ~0u;
	uint64_t psfd_off = 0;
	auto model_psfd = [&](uint32_t mi, const pk_model *m) -> uint64_t {
		if (psfd_model != mi) {
			psfd_model = mi;
			psfd_off = psfd_all.size();
			const size_t pn = (size_t)m->psf_width * m->psf_height;
			double sum = 0;
			for (size_t t = 0; t < pn; t++) sum += m->psf[t];
			for (size_t t = 0; t < pn; t++) psfd_all.push_back(sum > 0 ? m->psf[t] / sum : m->psf[t]);
		}
		return psfd_off;
	};
	size_t fft_points = 0;  // complex points of the largest FFT convolution
	uint64_t img_off = 0, out_off = 0;
	auto fail = [&](int code) {
		pk_batch_destroy(b);
		return code;
	};
	for (uint32_t mi = 0; mi < n_models; mi++) {
		const pk_model *m = &models[mi];
		Analysis a;
		int st = analyze(m, a);
		if (st) return fail(set_err(st, "model %u: %s", mi, g_err.c_str()));
		if (m->dry_run) {
			// src/model.cpp:263-267: a zero image of the drawing dimensions; the
			// pixel pass writes the zeros (no components), the composition
			// returns them
			for (uint32_t c = 0; c < m->n_components; c++) {
				b->comp_kind.push_back(m->components[c].kind);
				b->comp_index.push_back(~0u);
			}
			PkJob jb{};
			jb.dw = a.dw;
			jb.dh = a.dh;
			jb.comp_off = (uint32_t)b->setup.size();
			jb.img_off = img_off;
			jb.mask_off = NONE;
			jb.sx = m->scale_x / m->finesampling;
			jb.sy = m->scale_y / m->finesampling;
			jb.tile_off = b->ntiles;
			jb.tiles_x = cdiv(a.dw, 64);
			b->ntiles += jb.tiles_x * cdiv(a.dh, PX_ROWS);
			jb.has_nc = 1;
			jb.psf_off = psf_all.size();
			jb.conv_off = img_off;
			jb.cw = a.dw;
			jb.ch = a.dh;
			jb.crw = a.dw;
			jb.crh = a.dh;
			jb.fs = 1;
			jb.ow = a.dw;
			jb.oh = a.dh;
			jb.out_off = out_off;
			jb.omask_off = NONE;
			jb.omask_f = 1;
			jb.fx_scale = 1.0;
			b->jobs.push_back(jb);
			b->is_fft.push_back(0);
			b->origin.push_back(0);
			b->origin.push_back(0);
			img_off += (uint64_t)a.dw * a.dh;
			out_off += (uint64_t)a.dw * a.dh;
			continue;
		}
		if (fp32 && check_fp32_limits(m)) return fail(set_err(PK_EINVAL, "model %u: %s", mi, g_err.c_str()));
		const bool fft = a.conv_required && m->convolver == PK_CONV_FFT;
		// the FFT convolver's cropped result without wrapped rows is the
		// same-size linear convolution, whose adjoint is the direct correlation
		if (fft && (flags & PK_BATCH_GRAD) &&
		    !(m->crop && a.dh + m->psf_height - 1 <= 2 * std::max(a.dw, m->psf_width)))
			return fail(set_err(PK_EINVAL,
			                    "model %u: gradients with the FFT convolver need crop and a drawing height up to "
			                    "twice its width",
			                    mi));
		// calculation mask in the drawing frame (and the user's mask for the
		// returned image when it was adjusted), as Model::evaluate
		std::vector<uint8_t> dmask;
		bool use_mask = false, adjusted = false;
		if (has_mask(m)) {
			const int ms = calc_mask(m, a, dmask, adjusted);
			if (ms) return fail(set_err(ms, "model %u: %s", mi, g_err.c_str()));
			use_mask = true;
		}
		if (flags & PK_BATCH_GRAD) {
			for (uint32_t c = 0; c < m->n_components; c++) {
				const pk_component &pc = m->components[c];
				if (pc.kind != PK_PROFILE_SERSIC && pc.kind != PK_PROFILE_SKY && pc.kind != PK_PROFILE_NULL)
					return fail(set_err(PK_EINVAL, "model %u: gradients are available for sersic and sky profiles", mi));
				if (pc.kind == PK_PROFILE_SERSIC && pc.rescale_flux)
					return fail(set_err(PK_EINVAL, "model %u: gradients do not support rescale_flux", mi));
			}
		}
		const uint32_t fs = m->finesampling;
		const double sx = m->scale_x / fs, sy = m->scale_y / fs;
		PkJob jb{};
		jb.dw = a.dw;
		jb.dh = a.dh;
		jb.comp_off = (uint32_t)b->setup.size();
		jb.img_off = img_off;
		jb.mask_off = NONE;
		if (use_mask) {
			jb.mask_off = mask_all.size();
			mask_all.insert(mask_all.end(), dmask.begin(), dmask.end());
		}
		jb.sx = sx;
		jb.sy = sy;
		jb.tile_off = b->ntiles;
		jb.tiles_x = cdiv(a.dw, 64);
		b->ntiles += jb.tiles_x * cdiv(a.dh, PX_ROWS);
		bool any_nc = false;
		for (uint32_t c = 0; c < m->n_components; c++) {
			const pk_component &pc = m->components[c];
			b->comp_kind.push_back(pc.kind);
			b->min_passes = std::max(b->min_passes, refine_passes_needed(pc));
			if (pc.kind == PK_PROFILE_NULL) {
				b->comp_index.push_back(~0u);
				continue;
			}
			if (pc.kind == PK_PROFILE_PSF) {
				PkPsfItem it = psf_item(m, pc, a, sx, sy);
				it.job = mi;
				it.psf_off = model_psfd(mi, m);
				b->psf_items.push_back(it);
				pk_batch::PsfSrc src{};
				src.model = *m;
				src.model.components = nullptr;
				src.model.psf = nullptr;
				src.model.mask = nullptr;
				src.a = a;
				src.sx = sx;
				src.sy = sy;
				src.comp = (uint32_t)b->comp_kind.size() - 1;
				b->psf_src.push_back(src);
				b->comp_index.push_back(~0u);
				if (!pc.convolve) any_nc = true;
				continue;
			}
			SetupIn in{};
			in.c = pc;
			in.magzero = m->magzero;
			in.sx = sx;
			in.sy = sy;
			in.img_sx = m->scale_x;
			in.img_sy = m->scale_y;
			in.pad_x = a.padx;
			in.pad_y = a.pady;
			in.fp32_mode = fp32 ? 1 : 0;
			in.fs = fs;
			in.grad = (flags & PK_BATCH_GRAD) ? 1 : 0;
			in.integrator = g_integrator.load();
			b->comp_index.push_back((uint32_t)b->setup.size());
			b->setup.push_back(in);
			if (!pc.convolve) any_nc = true;
		}
		jb.ncomp = (uint32_t)b->setup.size() - jb.comp_off;
		if (jb.ncomp > 32) return fail(set_err(PK_EINVAL, "model %u: at most 32 radial/sky profiles per model", mi));
		jb.has_conv = a.conv_required ? 1 : 0;
		jb.has_nc = (any_nc || !a.conv_required) ? 1 : 0;
		jb.kw = a.conv_required ? m->psf_width : 0;
		jb.kh = a.conv_required ? m->psf_height : 0;
		jb.psf_off = psf_all.size();
		// crop of the PSF padding (src/model.cpp:289-306): origin (padx, pady)
		// of the drawing frame
		const bool cropped = a.padx > 0 && a.pady > 0;
		const int pcx = cropped ? (int)a.padx : 0, pcy = cropped ? (int)a.pady : 0;
		jb.crw = a.dw - 2 * (uint32_t)pcx;
		jb.crh = a.dh - 2 * (uint32_t)pcy;
		jb.conv_off = img_off;
		jb.ctile_off = 0;  // numbered by PSF-size class below
		int32_t ox = 0, oy = 0;
		if (fft) {
			// the FFT convolver's frames, as a single evaluation; its result
			// goes to a region of its own after the drawing-size ones
			int gx = 0, gy = 0;
			const int gs = frame_geometry(m, a, jb, gx, gy);
			if (gs) return fail(set_err(gs, "model %u: %s", mi, g_err.c_str()));
			FftGeom fg;
			const int fs_ = fft_geometry(a.dw, a.dh, m->psf_width, m->psf_height, fg);
			if (fs_) return fail(set_err(fs_, "model %u: %s", mi, g_err.c_str()));
			fft_points = std::max(fft_points, (size_t)fg.P * fg.Q);
			ox = gx;
			oy = gy;
			jb.ctiles_x = 0;
			jb.conv_off = (uint64_t)jb.cw * jb.ch;  // size for now; placed after the loop
			b->any_conv = true;
			if (flags & PK_BATCH_GRAD) {
				// the adjoint correlates with the PSF directly
				jb.psf_off = psf_all.size();
				const size_t pn = (size_t)m->psf_width * m->psf_height;
				double sum = 0;
				for (size_t t = 0; t < pn; t++) sum += m->psf[t];
				for (size_t t = 0; t < pn; t++) psf_all.push_back(sum > 0 ? m->psf[t] / sum : m->psf[t]);
				b->adj_kw = std::max(b->adj_kw, m->psf_width);
				b->adj_kh = std::max(b->adj_kh, m->psf_height);
			}
			pk_batch::FftJob fj{mi, model_psfd(mi, m), 0};
			for (size_t t = 0; t < (size_t)m->psf_width * m->psf_height; t++)
				fj.krn_max = std::max(fj.krn_max, std::fabs(psfd_all[fj.psf_off + t]));
			b->fft_jobs.push_back(fj);
		} else if (a.conv_required) {
			const size_t pn = (size_t)m->psf_width * m->psf_height;
			double sum = 0;
			for (size_t t = 0; t < pn; t++) sum += m->psf[t];
			for (size_t t = 0; t < pn; t++) psf_all.push_back(sum > 0 ? m->psf[t] / sum : m->psf[t]);
			b->max_kw = std::max(b->max_kw, m->psf_width);
			b->max_kh = std::max(b->max_kh, m->psf_height);
			b->adj_kw = std::max(b->adj_kw, m->psf_width);
			b->adj_kh = std::max(b->adj_kh, m->psf_height);
			b->any_conv = true;
			// only the pixels the crop keeps are convolved: the convolved frame
			// is the crop region, its pixel (X, Y) the same-size convolution at
			// drawing pixel (X + pcx, Y + pcy)
			jb.cw = jb.crw;
			jb.ch = jb.crh;
			jb.csx = pcx;
			jb.csy = pcy;
			jb.mox = -pcx;
			jb.moy = -pcy;
			jb.cmx = -pcx;  // calculation-mask pixel of convolved (X, Y): (X + pcx, Y + pcy)
			jb.cmy = -pcy;
			jb.cox = 0;
			jb.coy = 0;
			jb.ctiles_x = cdiv(jb.cw, CT_W);
		} else {
			jb.ctiles_x = 0;
			jb.cw = a.dw;
			jb.ch = a.dh;
			jb.cox = pcx;
			jb.coy = pcy;
		}
		jb.fs = 1;
		uint32_t ow = jb.crw, oh = jb.crh;
		if (fs > 1 && !m->return_finesampled) {
			jb.fs = fs;
			ow = cdiv(ow, fs);
			oh = cdiv(oh, fs);
		}
		jb.ow = ow;
		jb.oh = oh;
		jb.out_off = out_off;
		// without downsampling (and without the user's mask to apply again) the
		// convolved frame is the returned image: the convolution writes it
		// (composition tiles are set below)
		jb.fuse_out = (a.conv_required && !fft && jb.fs == 1 && !adjusted) ? 1 : 0;
		jb.omask_off = NONE;
		jb.omask_f = 1;
		if (adjusted) {
			jb.omask_off = mask_all.size();
			jb.omask_w = m->mask_width;
			jb.omask_h = m->mask_height;
			jb.omask_f = (ow == m->mask_width) ? 1 : fs;
			for (size_t t = 0; t < (size_t)m->mask_width * m->mask_height; t++)
				mask_all.push_back(m->mask[t] ? 1 : 0);
		}
		jb.fx_scale = fixed_point_scale(m);
		for (uint32_t k = jb.comp_off; k < jb.comp_off + jb.ncomp; k++) b->setup[k].fx_scale = jb.fx_scale;
		b->jobs.push_back(jb);
		b->is_fft.push_back(fft ? 1 : 0);
		b->origin.push_back(ox);
		b->origin.push_back(oy);
		img_off += (uint64_t)a.dw * a.dh;
		out_off += (uint64_t)ow * oh;
	}
	// FFT results after the drawing-size region of the convolution buffer
	uint64_t conv_total = img_off;
	for (const auto &fj : b->fft_jobs) {
		PkJob &jb = b->jobs[fj.job];
		const uint64_t n = jb.conv_off;
		jb.conv_off = conv_total;
		conv_total += n;
	}
	b->n_comps = (uint32_t)b->setup.size();
	b->n_pixels = img_off;
	b->total_out = out_off;
	{
		uint32_t rows = 0;
		if (b->any_conv && conv_rows_per_launch(b->max_kw, b->max_kh, elem, rows)) return fail(PK_EINVAL);
		// the convolution writes returned images only when it runs in one
		// launch; the composition covers every other job
		const bool one_launch = !b->any_conv || rows >= b->max_kh;
		std::vector<uint8_t> nc_psf(b->jobs.size(), 0);  // job has an unconvolved psf profile
		for (const auto &it : b->psf_items)
			if (!it.to_conv) nc_psf[it.job] = 1;
		for (auto &jb : b->jobs) {
			if (!one_launch) jb.fuse_out = 0;
			// FP32: a job whose unconvolved part is only sky, unmasked, whose
			// convolution writes the returned image, takes that part as one
			// value (the unconvolved image is never stored)
			bool sky_only = elem == 4 && jb.has_nc && jb.fuse_out && jb.mask_off == NONE &&
			                !nc_psf[&jb - b->jobs.data()];
			for (uint32_t k = jb.comp_off; sky_only && k < jb.comp_off + jb.ncomp; k++)
				if (!b->setup[k].c.convolve && b->setup[k].c.kind != PK_PROFILE_SKY) sky_only = false;
			jb.nc_sky = sky_only ? 1 : 0;
			jb.otile_off = b->notiles;
			jb.otiles_x = jb.fuse_out ? 0 : cdiv(jb.ow, 64);
			b->notiles += jb.otiles_x * cdiv(jb.oh, PX_ROWS);
		}
		// convolution tiles grouped by PSF size (largest side <= 17, 21, ..., 41,
		// ST_KMAX, more)
		const uint32_t class_lim[] = {17, 21, 25, 29, 33, 37, 41, (uint32_t)ST_KMAX, UINT32_MAX};
		const uint32_t n_class = sizeof(class_lim) / sizeof(class_lim[0]);
		for (uint32_t lim_lo = 0, ci = 0; ci < n_class; lim_lo = class_lim[ci], ci++) {
			pk_batch::ConvClass cc{b->nctiles, 0, 0, 0, class_lim[ci] <= (uint32_t)ST_KMAX};
			for (auto &jb : b->jobs) {
				const uint32_t side = std::max(jb.kw, jb.kh);
				if (!jb.has_conv || b->is_fft[&jb - b->jobs.data()] || side <= lim_lo || side > class_lim[ci]) continue;
				jb.ctile_off = b->nctiles;
				b->nctiles += jb.ctiles_x * cdiv(jb.ch, CT_H);
				cc.max_kw = std::max(cc.max_kw, jb.kw);
				cc.max_kh = std::max(cc.max_kh, jb.kh);
			}
			cc.count = b->nctiles - cc.start;
			if (cc.count) b->conv_classes.push_back(cc);
		}
		for (auto &jb : b->jobs)
			if (!jb.has_conv || b->is_fft[&jb - b->jobs.data()]) jb.ctile_off = b->nctiles;
		// strip tiles of the tallest images first, so that a launch ends on
		// short tiles (workgroups are started in tile order)
		std::vector<uint32_t> by_height(b->jobs.size());
		for (uint32_t k = 0; k < by_height.size(); k++) by_height[k] = k;
		std::stable_sort(by_height.begin(), by_height.end(),
		                 [&](uint32_t x, uint32_t y) { return b->jobs[x].ch > b->jobs[y].ch; });
		for (uint32_t c = 1; c <= (ST_KMAX + 15) / 16; c++)
			for (uint32_t tm = 0; tm <= (uint32_t)ST_TAIL; tm++)
			for (uint32_t nj = 1; nj <= (ST_KMAX + 3) / 4; nj++) {
				pk_batch::StripClass sc{b->nstiles, 0, nj, c, tm};
				for (uint32_t k : by_height) {
					PkJob &jb = b->jobs[k];
					uint32_t jc, jtm;
					strip_row_form(jb.kh, jc, jtm);
					if (!strip_job(b, jb) || cdiv(jb.kw, 4) != nj || jc != c || jtm != tm) continue;
					jb.stile_off = b->nstiles;
					jb.st_rows = strip_rows(jb.ch);
					b->nstiles += jb.ctiles_x * cdiv(jb.ch, jb.st_rows);
				}
				sc.count = b->nstiles - sc.start;
				if (sc.count) b->strip_classes.push_back(sc);
			}
		for (auto &jb : b->jobs)
			if (!strip_job(b, jb)) jb.stile_off = b->nstiles;
	}
	if (img_off >= (1ull << 32)) return fail(set_err(PK_EINVAL, "batch too large (%llu pixels)", (unsigned long long)img_off));

	// job of every tile of the three tilings (one scalar load per workgroup)
	std::vector<uint32_t> tj((size_t)b->ntiles + b->nctiles + b->notiles + b->nstiles);
	for (uint32_t jn = 0; jn < n_models; jn++) {
		const PkJob &jb = b->jobs[jn];
		const uint32_t nt = (jn + 1 < n_models ? b->jobs[jn + 1].tile_off : b->ntiles) - jb.tile_off;
		const uint32_t nc = jb.has_conv ? jb.ctiles_x * cdiv(jb.ch, CT_H) : 0u;  // (tiles numbered by class)
		const uint32_t no = (jn + 1 < n_models ? b->jobs[jn + 1].otile_off : b->notiles) - jb.otile_off;
		std::fill_n(tj.begin() + jb.tile_off, nt, jn);
		std::fill_n(tj.begin() + b->ntiles + jb.ctile_off, nc, jn);
		std::fill_n(tj.begin() + b->ntiles + b->nctiles + jb.otile_off, no, jn);
		const uint32_t ns = strip_job(b, jb) ? jb.ctiles_x * cdiv(jb.ch, jb.st_rows) : 0u;
		std::fill_n(tj.begin() + b->ntiles + b->nctiles + b->notiles + jb.stile_off, ns, jn);
	}
	// adjoint of the convolution (PK_BATCH_GRAD): the correlation of the
	// adjoint image with the PSF covers the whole drawing frame
	std::vector<PkJob> adj;
	std::vector<uint32_t> tja;
	if (flags & PK_BATCH_GRAD) {
		adj = b->jobs;
		for (uint32_t jn = 0; jn < n_models; jn++) {
			PkJob &jb = adj[jn];
			jb.ctile_off = b->natiles;
			jb.ctiles_x = 0;
			jb.mask_off = NONE;  // the adjoint image is zero at masked pixels already
			jb.conv_off = jb.img_off;  // the correlation lands in the drawing-frame layout
			jb.fuse_out = 0;
			jb.nc_sky = 0;  // no unconvolved part is added to the correlation
			if (!jb.has_conv) continue;
			jb.cw = jb.dw;
			jb.ch = jb.dh;
			// forward: out(X) = sum_k src(X + k - h) K(kw-1-k), h = kw / 2, so
			// G_src(x) = sum_k G(x + k - (kw-1-h)) K(k): the unflipped PSF with
			// the source shifted by 2h - kw + 1 (0 for odd, 1 for even sides)
			jb.csx = 2 * (int32_t)(jb.kw / 2) - (int32_t)jb.kw + 1;
			jb.csy = 2 * (int32_t)(jb.kh / 2) - (int32_t)jb.kh + 1;
			jb.ctiles_x = cdiv(jb.dw, CT_W);
			const uint32_t n = jb.ctiles_x * cdiv(jb.dh, CT_H);
			tja.insert(tja.end(), n, jn);
			b->natiles += n;
		}
		b->adj_strips = b->any_conv;
		for (const auto &jb : adj)
			if (jb.has_conv && (jb.kw > (uint32_t)ST_KMAX || jb.kh > (uint32_t)ST_KMAX)) b->adj_strips = false;
		if (b->adj_strips) {
			// strip tiles by class, the tallest images first (as the forward ones)
			std::vector<uint32_t> by_height(adj.size());
			for (uint32_t k = 0; k < by_height.size(); k++) by_height[k] = k;
			std::stable_sort(by_height.begin(), by_height.end(),
			                 [&](uint32_t x, uint32_t y) { return adj[x].ch > adj[y].ch; });
			for (uint32_t c = 1; c <= (ST_KMAX + 15) / 16; c++)
				for (uint32_t tm = 0; tm <= (uint32_t)ST_TAIL; tm++)
				for (uint32_t nj = 1; nj <= (ST_KMAX + 3) / 4; nj++) {
					pk_batch::StripClass sc{b->nastiles, 0, nj, c, tm};
					for (uint32_t k : by_height) {
						PkJob &jb = adj[k];
						uint32_t jc, jtm;
						strip_row_form(jb.kh, jc, jtm);
						if (!jb.has_conv || cdiv(jb.kw, 4) != nj || jc != c || jtm != tm) continue;
						jb.stile_off = b->nastiles;
						jb.st_rows = strip_rows(jb.ch);
						const uint32_t n = jb.ctiles_x * cdiv(jb.ch, jb.st_rows);
						tja.insert(tja.end(), n, k);
						b->nastiles += n;
					}
					sc.count = b->nastiles - sc.start;
					if (sc.count) b->adj_strip_classes.push_back(sc);
				}
		}
	}
	// psf profiles in layers (k-th psf profile of each model), items in
	// layer-major order
	if (!b->psf_items.empty()) {
		std::vector<uint32_t> layer(b->psf_items.size()), seen(n_models, 0);
		uint32_t n_layers = 0;
		for (size_t k = 0; k < b->psf_items.size(); k++) {
			layer[k] = seen[b->psf_items[k].job]++;
			n_layers = std::max(n_layers, layer[k] + 1);
		}
		std::vector<size_t> order(b->psf_items.size());
		for (size_t k = 0; k < order.size(); k++) order[k] = k;
		std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return layer[x] < layer[y]; });
		std::vector<PkPsfItem> items;
		std::vector<pk_batch::PsfSrc> src;
		for (size_t k : order) {
			items.push_back(b->psf_items[k]);
			src.push_back(b->psf_src[k]);
		}
		b->psf_items.swap(items);
		b->psf_src.swap(src);
		b->psf_layers.assign(n_layers, pk_batch::PsfLayer{0, 0, 0});
		uint32_t start = 0;
		for (uint32_t l = 0; l < n_layers; l++) {
			b->psf_layers[l].start = start;
			for (size_t k = start; k < b->psf_items.size() && layer[order[k]] == l; k++) b->psf_layers[l].count++;
			start += b->psf_layers[l].count;
		}
		b->psf_update_layers();
	}
	Layout L;
	b->o_jobs = L.take(sizeof(PkJob) * b->jobs.size());
	b->o_setup = L.take(sizeof(SetupIn) * std::max<size_t>(1, b->setup.size()));
	b->o_psfT = L.take(elem * std::max<size_t>(1, psf_all.size()));
	b->o_tj = L.take(sizeof(uint32_t) * std::max<size_t>(1, tj.size()));
	b->o_items = L.take(sizeof(PkPsfItem) * std::max<size_t>(1, b->psf_items.size()));
	b->o_psfd = L.take(sizeof(double) * std::max<size_t>(1, psfd_all.size()));
	const size_t in_bytes = L.off;
	b->o_mask = L.take(std::max<size_t>(1, mask_all.size()));
	b->o_comps = L.take(sizeof(PkComp) * std::max<size_t>(1, b->setup.size()));
	b->o_lite = L.take(sizeof(CompLite) * std::max<size_t>(1, b->setup.size()));
	b->o_cnt = L.take(sizeof(PkCounters));
	b->o_wev = L.take(sizeof(uint32_t) * 4 * std::max<size_t>(1, b->ntiles));
	b->o_wsum = L.take(sizeof(unsigned long long));
	b->o_nc = L.take(elem * img_off);
	b->o_c = L.take(elem * (b->any_conv ? img_off : 1));
	bool conv_buffer = false;  // some convolved job leaves its composition to compose_kernel
	for (const auto &jb : b->jobs) conv_buffer = conv_buffer || (jb.has_conv && !jb.fuse_out);
	b->o_conv = L.take(elem * (conv_buffer ? conv_total : 1));
	b->o_fz = L.take(fft_points ? sizeof(double2) * (fft_points + 1) : 1);
	b->o_ftw = L.take(fft_points ? sizeof(double2) * FFT_TWIDDLES : 1);
	b->o_ent = L.take(sizeof(RefineEntry) * img_off);
	b->o_acc = L.take(fp32 ? 2 * sizeof(unsigned long long) * img_off : 8);
	b->qcap = fp32 ? r32_qcap_batch(b->setup.size()) : 0;
	b->o_q = L.take(3 * sizeof(RCell) * std::max<uint32_t>(1, b->qcap));
	b->o_out = L.take(elem * out_off);
	if (flags & PK_BATCH_GRAD) {
		b->o_psfF = L.take(elem * std::max<size_t>(1, psf_all.size()));
		b->o_gd = L.take(elem * img_off);
		b->o_gc = L.take(elem * (b->any_conv ? img_off : 1));
		b->o_data = L.take(elem * out_off);
		b->o_ivar = L.take(elem * out_off);
		b->o_gacc = L.take(sizeof(double) * GA_SLOTS * std::max<size_t>(1, b->setup.size()));
		b->o_gout = L.take(sizeof(double) * PK_GRAD_SLOTS * std::max<size_t>(1, b->setup.size()));
		b->o_chi2 = L.take(sizeof(double) * n_models);
		b->o_gcnt = L.take(sizeof(PkCounters));
		b->o_jobs_adj = L.take(sizeof(PkJob) * adj.size());
		b->o_tja = L.take(sizeof(uint32_t) * std::max<size_t>(1, tja.size()));
	}
	if (hipMalloc(&b->dmem, L.off) != hipSuccess) {
		(void)hipGetLastError();
		return fail(set_err(PK_ENOMEM, "device allocation of %zu bytes failed", L.off));
	}
	std::vector<unsigned char> stage(in_bytes, 0);
	std::memcpy(stage.data() + b->o_jobs, b->jobs.data(), sizeof(PkJob) * b->jobs.size());
	if (!tj.empty()) std::memcpy(stage.data() + b->o_tj, tj.data(), sizeof(uint32_t) * tj.size());
	if (!b->setup.empty()) std::memcpy(stage.data() + b->o_setup, b->setup.data(), sizeof(SetupIn) * b->setup.size());
	if (!b->psf_items.empty())
		std::memcpy(stage.data() + b->o_items, b->psf_items.data(), sizeof(PkPsfItem) * b->psf_items.size());
	if (!psfd_all.empty()) std::memcpy(stage.data() + b->o_psfd, psfd_all.data(), sizeof(double) * psfd_all.size());
	for (const auto &jb : b->jobs) {
		if (!jb.has_conv || b->is_fft[&jb - b->jobs.data()]) continue;
		if (fp32)
			psf_conv_layout<float>(psf_all.data() + jb.psf_off, jb.kw, jb.kh,
			                       reinterpret_cast<float *>(stage.data() + b->o_psfT) + jb.psf_off);
		else
			psf_conv_layout<double>(psf_all.data() + jb.psf_off, jb.kw, jb.kh,
			                        reinterpret_cast<double *>(stage.data() + b->o_psfT) + jb.psf_off);
	}
	PK_HIP(hipMemcpy(b->dmem, stage.data(), in_bytes, hipMemcpyHostToDevice));
	if (!mask_all.empty())
		PK_HIP(hipMemcpy(b->d<void>(b->o_mask), mask_all.data(), mask_all.size(), hipMemcpyHostToDevice));
	if (fft_points)
		PK_HIP(hipMemcpy(b->d<void>(b->o_ftw), fft_twiddles(), sizeof(double2) * FFT_TWIDDLES, hipMemcpyHostToDevice));
	if (flags & PK_BATCH_GRAD) {
		// PSF layout of the adjoint convolution: kt[k * kh + l] = krn(k, l)
		// (the flipped PSF in the layout conv_kernel reads)
		std::vector<unsigned char> f(elem * std::max<size_t>(1, psf_all.size()), 0);
		for (const auto &jb : b->jobs) {
			if (!jb.has_conv) continue;
			const double *k = psf_all.data() + jb.psf_off;
			for (uint32_t a = 0; a < jb.kw; a++)
				for (uint32_t l = 0; l < jb.kh; l++) {
					const double v = k[a + (size_t)l * jb.kw];
					const size_t o = jb.psf_off + (size_t)a * jb.kh + l;
					if (fp32) reinterpret_cast<float *>(f.data())[o] = (float)v;
					else reinterpret_cast<double *>(f.data())[o] = v;
				}
		}
		PK_HIP(hipMemcpy(b->d<void>(b->o_psfF), f.data(), f.size(), hipMemcpyHostToDevice));
		PK_HIP(hipMemcpy(b->d<void>(b->o_jobs_adj), adj.data(), sizeof(PkJob) * adj.size(), hipMemcpyHostToDevice));
		if (!tja.empty())
			PK_HIP(hipMemcpy(b->d<void>(b->o_tja), tja.data(), sizeof(uint32_t) * tja.size(), hipMemcpyHostToDevice));
		for (auto &e : b->ev_grad)
			if (hipEventCreate(&e) != hipSuccess) return fail(set_err(PK_EDEVICE, "hipEventCreate failed"));
	}
	for (auto &e : b->ev)
		if (hipEventCreate(&e) != hipSuccess) return fail(set_err(PK_EDEVICE, "hipEventCreate failed"));
	*out = b;
	return PK_OK;
}

int pk_batch_autotune(pk_batch *b)
{
	if (!b) return set_err(PK_EINVAL, "batch is NULL");
	std::lock_guard<std::mutex> lock(b->ctx->mu);
	PK_HIP(hipSetDevice(b->ctx->device));
	PK_TRY(autotune(b));
	b->ran = true;
	return PK_OK;
}

int pk_batch_get_tuning(const pk_batch *b, pk_tuning *t)
{
	if (!b || !t) return set_err(PK_EINVAL, "NULL argument");
	*t = b->tune;
	return PK_OK;
}

int pk_batch_set_tuning(pk_batch *b, const pk_tuning *t)
{
	if (!b || !t) return set_err(PK_EINVAL, "NULL argument");
	if (!tuning_valid(*t))
		return set_err(PK_EINVAL,
		               "invalid tuning (conv_tiles %u, occupancy %u, handover %u, claims %u/%u, passes %u, spread %u)",
		               t->conv_tiles, t->refine_occupancy, t->handover_div, t->claim_pixels, t->claim_cells,
		               t->refine_passes, t->claim_spread);
	std::lock_guard<std::mutex> lock(b->ctx->mu);
	PK_HIP(hipSetDevice(b->ctx->device));
	PK_HIP(hipStreamSynchronize(b->ctx->stream));
	drop_graphs(b);
	b->tune = *t;
	b->tuned = true;
	return PK_OK;
}

int pk_batch_set_data(pk_batch *b, const double *data, const double *sigma, size_t n)
{
	if (!b || !data) return set_err(PK_EINVAL, "NULL argument");
	if (!(b->flags & PK_BATCH_GRAD)) return set_err(PK_EINVAL, "the batch was not created with PK_BATCH_GRAD");
	if (n != b->total_out) return set_err(PK_EINVAL, "expected %llu data values, got %zu",
	                                      (unsigned long long)b->total_out, n);
	std::vector<unsigned char> d(b->elem * n), iv(b->elem * n);
	for (size_t t = 0; t < n; t++) {
		const double sg = sigma ? sigma[t] : 1.0;
		if (!(sg > 0) || !std::isfinite(sg)) return set_err(PK_EINVAL, "sigma[%zu] = %g, must be > 0", t, sg);
		const double ivv = 1 / (sg * sg);
		if (b->elem == 4) {
			reinterpret_cast<float *>(d.data())[t] = (float)data[t];
			reinterpret_cast<float *>(iv.data())[t] = (float)ivv;
		} else {
			reinterpret_cast<double *>(d.data())[t] = data[t];
			reinterpret_cast<double *>(iv.data())[t] = ivv;
		}
	}
	std::lock_guard<std::mutex> lock(b->ctx->mu);
	PK_HIP(hipSetDevice(b->ctx->device));
	PK_HIP(hipMemcpyAsync(b->d<void>(b->o_data), d.data(), d.size(), hipMemcpyHostToDevice, b->ctx->stream));
	PK_HIP(hipMemcpyAsync(b->d<void>(b->o_ivar), iv.data(), iv.size(), hipMemcpyHostToDevice, b->ctx->stream));
	PK_HIP(hipStreamSynchronize(b->ctx->stream));
	b->have_data = true;
	return PK_OK;
}

int pk_batch_evaluate_grad(pk_batch *b)
{
	if (!b) return set_err(PK_EINVAL, "batch is NULL");
	if (!(b->flags & PK_BATCH_GRAD)) return set_err(PK_EINVAL, "the batch was not created with PK_BATCH_GRAD");
	if (!b->have_data) return set_err(PK_EINVAL, "no data: call pk_batch_set_data first");
	std::lock_guard<std::mutex> lock(b->ctx->mu);
	PK_HIP(hipSetDevice(b->ctx->device));
	PK_TRY(ensure_tuned(b));
	b->ran = true;
	b->ran_grad = true;
	if (!(b->flags & PK_BATCH_GRAPH)) return grad_launch(b);
	if (!b->exec_grad) {
		PK_HIP(hipStreamBeginCapture(b->ctx->stream, hipStreamCaptureModeThreadLocal));
		int st = grad_launch(b);
		hipGraph_t g = nullptr;
		hipError_t e = hipStreamEndCapture(b->ctx->stream, &g);
		if (st) {
			if (g) (void)hipGraphDestroy(g);
			return st;
		}
		if (e != hipSuccess) return set_err(PK_EDEVICE, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
		b->graph_grad = g;
		PK_HIP(hipGraphInstantiate(&b->exec_grad, g, nullptr, nullptr, 0));
	}
	PK_HIP(hipGraphLaunch(b->exec_grad, b->ctx->stream));
	return PK_OK;
}

int pk_batch_get_chi2_grad(pk_batch *b, double *chi2, double *grad)
{
	if (!b) return set_err(PK_EINVAL, "batch is NULL");
	if (!b->ran_grad) return set_err(PK_EINVAL, "no gradient evaluated yet");
	std::lock_guard<std::mutex> lock(b->ctx->mu);
	PK_HIP(hipSetDevice(b->ctx->device));
	PK_HIP(hipStreamSynchronize(b->ctx->stream));
	// the forward pass's counters first (its refinement pattern is the one
	// the adjoint follows), then the adjoint's own
	PkCounters fwd{}, cnt{};
	PK_HIP(hipMemcpy(&fwd, b->d<void>(b->o_cnt), sizeof(fwd), hipMemcpyDeviceToHost));
	PK_HIP(hipMemcpy(&cnt, b->d<void>(b->o_gcnt), sizeof(cnt), hipMemcpyDeviceToHost));
	if (fwd.overflow || cnt.overflow) return set_err(PK_ERANGE, "sub-pixel refinement work queue capacity exceeded");
	if (fwd.n_bad) return set_err(PK_ERANGE, "sub-pixel refinement produced %u non-finite values", fwd.n_bad);
	if (cnt.n_bad) return set_err(PK_ERANGE, "the refinement adjoint produced %u non-finite values", cnt.n_bad);
	if (chi2) PK_HIP(hipMemcpy(chi2, b->d<void>(b->o_chi2), sizeof(double) * b->n_models, hipMemcpyDeviceToHost));
	if (grad) {
		std::vector<double> g((size_t)PK_GRAD_SLOTS * std::max<size_t>(1, b->n_comps));
		PK_HIP(hipMemcpy(g.data(), b->d<void>(b->o_gout), sizeof(double) * g.size(), hipMemcpyDeviceToHost));
		for (size_t i = 0; i < b->comp_index.size(); i++) {
			double *o = grad + i * PK_GRAD_SLOTS;
			const uint32_t slot = b->comp_index[i];
			for (int k = 0; k < PK_GRAD_SLOTS; k++) o[k] = slot == ~0u ? 0.0 : g[(size_t)slot * PK_GRAD_SLOTS + k];
		}
	}
	return PK_OK;
}

int pk_batch_set_components(pk_batch *b, const pk_component *comps, uint32_t n)
{
	if (!b || !comps) return set_err(PK_EINVAL, "NULL argument");
	if (n != b->comp_index.size()) return set_err(PK_EINVAL, "expected %zu components, got %u", b->comp_index.size(), n);
	for (uint32_t i = 0; i < n; i++)
		if (comps[i].kind != b->comp_kind[i]) return set_err(PK_EINVAL, "component %u changed its kind", i);
	for (const auto &src : b->psf_src) {
		const pk_component &pc = comps[src.comp];
		const auto &old = b->psf_items[&src - b->psf_src.data()];
		if ((pc.convolve != 0) != (old.to_conv != 0))
			return set_err(PK_EINVAL, "component %u changed its convolve flag", src.comp);
		PK_TRY(validate_component(pc, true));
	}
	// every component is validated before any is copied, so a rejected
	// update leaves the batch as it was
	for (uint32_t i = 0; i < n; i++) {
		if (b->comp_index[i] == ~0u) continue;
		const SetupIn &in = b->setup[b->comp_index[i]];
		if ((comps[i].convolve != 0) != (in.c.convolve != 0))
			return set_err(PK_EINVAL, "component %u changed its convolve flag", i);
		PK_TRY(validate_component(comps[i], true));
		if (b->elem == 4 && !comps[i].adjust && (comps[i].resolution == 0 || comps[i].resolution > 255))
			return set_err(PK_EINVAL, "resolution %u: the FP32 path supports 1..255", comps[i].resolution);
	}
	for (uint32_t i = 0; i < n; i++)
		if (b->comp_index[i] != ~0u) b->setup[b->comp_index[i]].c = comps[i];
	uint32_t passes = R32_PASSES_MIN;
	for (uint32_t i = 0; i < n; i++) passes = std::max(passes, refine_passes_needed(comps[i]));
	// fixed-point scales follow the new magnitudes
	for (auto &jb : b->jobs) {
		double f = 0;
		for (uint32_t k = jb.comp_off; k < jb.comp_off + jb.ncomp; k++) {
			const SetupIn &in = b->setup[k];
			if (in.c.kind == PK_PROFILE_SKY) continue;
			f += std::pow(10, -0.4 * (in.c.mag - in.magzero));
		}
		jb.fx_scale = (f > 0 && std::isfinite(f)) ? std::ldexp(1.0, 60) / f : 1.0;
		for (uint32_t k = jb.comp_off; k < jb.comp_off + jb.ncomp; k++) b->setup[k].fx_scale = jb.fx_scale;
	}
	std::lock_guard<std::mutex> lock(b->ctx->mu);
	PK_HIP(hipSetDevice(b->ctx->device));
	// deeper recursion than before may need more refinement launches: the
	// captured graphs are dropped (recaptured at the next evaluation)
	if (passes > b->min_passes) {
		PK_HIP(hipStreamSynchronize(b->ctx->stream));
		drop_graphs(b);
		b->min_passes = passes;
	}
	PK_HIP(hipMemcpyAsync(b->d<void>(b->o_jobs), b->jobs.data(), sizeof(PkJob) * b->jobs.size(),
	                      hipMemcpyHostToDevice, b->ctx->stream));
	PK_HIP(hipMemcpyAsync(b->d<void>(b->o_setup), b->setup.data(), sizeof(SetupIn) * b->setup.size(),
	                      hipMemcpyHostToDevice, b->ctx->stream));
	if (!b->psf_items.empty()) {
		// new positions and magnitudes of the psf profiles; a layer whose
		// largest item changed its size needs new launches
		std::vector<uint32_t> before;
		for (const auto &ly : b->psf_layers) before.push_back(ly.max_pixels);
		for (size_t k = 0; k < b->psf_items.size(); k++) {
			const auto &src = b->psf_src[k];
			PkPsfItem it = psf_item(&src.model, comps[src.comp], src.a, src.sx, src.sy);
			it.job = b->psf_items[k].job;
			it.psf_off = b->psf_items[k].psf_off;
			b->psf_items[k] = it;
		}
		b->psf_update_layers();
		for (size_t l = 0; l < before.size(); l++)
			if (before[l] != b->psf_layers[l].max_pixels) {
				PK_HIP(hipStreamSynchronize(b->ctx->stream));
				drop_graphs(b);
				break;
			}
		PK_HIP(hipMemcpyAsync(b->d<void>(b->o_items), b->psf_items.data(), sizeof(PkPsfItem) * b->psf_items.size(),
		                      hipMemcpyHostToDevice, b->ctx->stream));
	}
	PK_HIP(hipStreamSynchronize(b->ctx->stream));
	return PK_OK;
}

int pk_batch_evaluate(pk_batch *b)
{
	if (!b) return set_err(PK_EINVAL, "batch is NULL");
	std::lock_guard<std::mutex> lock(b->ctx->mu);
	PK_HIP(hipSetDevice(b->ctx->device));
	PK_TRY(ensure_tuned(b));
	b->ran = true;
	if (!(b->flags & PK_BATCH_GRAPH)) return batch_launch(b);
	if (!b->exec) {
		PK_HIP(hipStreamBeginCapture(b->ctx->stream, hipStreamCaptureModeThreadLocal));
		int st = batch_launch(b);
		hipGraph_t g = nullptr;
		hipError_t e = hipStreamEndCapture(b->ctx->stream, &g);
		if (st) {
			if (g) (void)hipGraphDestroy(g);
			return st;
		}
		if (e != hipSuccess) return set_err(PK_EDEVICE, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
		b->graph = g;
		PK_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
	}
	PK_HIP(hipGraphLaunch(b->exec, b->ctx->stream));
	return PK_OK;
}

int pk_batch_sync(pk_batch *b, pk_batch_info *info)
{
	if (!b) return set_err(PK_EINVAL, "batch is NULL");
	std::lock_guard<std::mutex> lock(b->ctx->mu);
	PK_HIP(hipSetDevice(b->ctx->device));
	PK_HIP(hipStreamSynchronize(b->ctx->stream));
	PkCounters cnt{};
	PK_HIP(hipMemcpy(&cnt, b->d<void>(b->o_cnt), sizeof(cnt), hipMemcpyDeviceToHost));
	if (info) {
		std::memset(info, 0, sizeof(*info));
		info->n_refine_entries = cnt.n_entries;
		info->n_subcells = cnt.n_subcells;
		if (b->ran_grad) {
			PkCounters g{};
			PK_HIP(hipMemcpy(&g, b->d<void>(b->o_gcnt), sizeof(g), hipMemcpyDeviceToHost));
			info->n_grad_subcells = g.n_subcells;
		}
		info->conv_tiles = b->elem == 4 ? b->tune.conv_tiles : (uint32_t)PK_CONV_TILES_VALU;
		info->n_conv_launches = 0;
		for (const auto &cc : b->conv_classes)
			if (!(info->conv_tiles == PK_CONV_TILES_STRIPS && cc.strips)) info->n_conv_launches++;
		if (info->conv_tiles == PK_CONV_TILES_STRIPS)
			info->n_conv_launches += (uint32_t)b->strip_classes.size();
		if (b->ran && !(b->flags & PK_BATCH_GRAPH)) {
			float t[6] = {};
			(void)hipEventElapsedTime(&t[0], b->ev[0], b->ev[1]);
			(void)hipEventElapsedTime(&t[1], b->ev[1], b->ev[2]);
			(void)hipEventElapsedTime(&t[2], b->ev[2], b->ev[3]);
			(void)hipEventElapsedTime(&t[3], b->ev[3], b->ev[4]);
			(void)hipEventElapsedTime(&t[4], b->ev[4], b->ev[5]);
			(void)hipEventElapsedTime(&t[5], b->ev[0], b->ev[5]);
			info->ms_setup = t[0];
			info->ms_coarse = t[1];
			info->ms_refine = t[2];
			info->ms_conv = t[3];
			info->ms_compose = t[4];
			info->ms_total = t[5];
			if (b->ran_grad) {
				float tg = 0, tr = 0;
				(void)hipEventElapsedTime(&tg, b->ev_grad[0], b->ev_grad[1]);
				(void)hipEventElapsedTime(&tr, b->ev_grad[2], b->ev_grad[1]);
				info->ms_grad = tg;
				info->ms_grad_refine = tr;
			}
		}
	}
	if (cnt.overflow) return set_err(PK_ERANGE, "sub-pixel refinement work queue capacity exceeded");
	if (cnt.n_bad) return set_err(PK_ERANGE, "sub-pixel refinement produced %u non-finite values", cnt.n_bad);
	return PK_OK;
}

int pk_batch_work_counts(pk_batch *b, pk_work_counts *out)
{
	if (!b || !out) return set_err(PK_EINVAL, "NULL argument");
	std::memset(out, 0, sizeof(*out));
	for (const auto &jb : b->jobs) {
		uint32_t nrad = 0;
		for (uint32_t k = jb.comp_off; k < jb.comp_off + jb.ncomp; k++)
			if (b->setup[k].c.kind != PK_PROFILE_SKY) nrad++;
		out->pixel_component_pairs += (uint64_t)jb.dw * jb.dh * nrad;
	}
	if (!b->ntiles) return PK_OK;
	std::lock_guard<std::mutex> lock(b->ctx->mu);
	PK_HIP(hipSetDevice(b->ctx->device));
	hipStream_t s = b->ctx->stream;
	// the component constants of the last evaluation (set-up ran), the
	// classification of the pixel pass with its evaluations counted
	const PkJob *djobs = b->d<PkJob>(b->o_jobs);
	const uint32_t *tj = b->d<uint32_t>(b->o_tj);
	uint32_t *wev = b->d<uint32_t>(b->o_wev);
	if (b->elem == 4)
		hipLaunchKernelGGL((coarse32_kernel<float, true>), dim3(b->ntiles), dim3(256), 0, s, djobs, b->n_models,
		                   (const PkComp *)b->d<PkComp>(b->o_comps), (const uint8_t *)b->d<uint8_t>(b->o_mask),
		                   (float *)nullptr, (float *)nullptr, (RefineEntry *)nullptr, (PkCounters *)nullptr,
		                   (unsigned long long *)nullptr, tj, wev);
	else
		hipLaunchKernelGGL((coarse_kernel<double, true>), dim3(b->ntiles), dim3(256), 0, s, djobs, b->n_models,
		                   (const PkComp *)b->d<PkComp>(b->o_comps), (const uint8_t *)b->d<uint8_t>(b->o_mask),
		                   (double *)nullptr, (double *)nullptr, (RefineEntry *)nullptr, (PkCounters *)nullptr,
		                   (unsigned long long *)nullptr, tj, wev);
	PK_HIP(hipGetLastError());
	unsigned long long *dsum = b->d<unsigned long long>(b->o_wsum);
	PK_HIP(hipMemsetAsync(dsum, 0, sizeof(unsigned long long), s));
	hipLaunchKernelGGL(sum_wave_evals_kernel, dim3(256), dim3(256), 0, s, (const uint32_t *)wev,
	                   (uint64_t)4 * b->ntiles, dsum);
	PK_HIP(hipGetLastError());
	unsigned long long tot = 0;
	PK_HIP(hipMemcpyAsync(&tot, dsum, sizeof(tot), hipMemcpyDeviceToHost, s));
	PK_HIP(hipStreamSynchronize(s));
	out->coarse_evaluations = tot;
	PkCounters cnt{};
	PK_HIP(hipMemcpy(&cnt, b->d<void>(b->o_cnt), sizeof(cnt), hipMemcpyDeviceToHost));
	out->refine_evaluations = cnt.n_subcells;
	out->refined_pixels = cnt.n_entries;
	return PK_OK;
}

int pk_batch_image_dims(const pk_batch *b, uint32_t i, uint32_t *w, uint32_t *h)
{
	if (!b || i >= b->n_models) return set_err(PK_EINVAL, "bad batch or image index");
	if (w) *w = b->jobs[i].ow;
	if (h) *h = b->jobs[i].oh;
	return PK_OK;
}

int pk_batch_get_image(pk_batch *b, uint32_t i, double *out, size_t cap)
{
	if (!b || !out || i >= b->n_models) return set_err(PK_EINVAL, "bad argument");
	const PkJob &jb = b->jobs[i];
	const size_t n = (size_t)jb.ow * jb.oh;
	if (n > cap) return set_err(PK_ERANGE, "output buffer holds %zu doubles, %zu needed", cap, n);
	std::lock_guard<std::mutex> lock(b->ctx->mu);
	PK_HIP(hipSetDevice(b->ctx->device));
	PK_HIP(hipStreamSynchronize(b->ctx->stream));
	if (b->elem == 8) {
		PK_HIP(hipMemcpy(out, b->d<double>(b->o_out) + jb.out_off, n * 8, hipMemcpyDeviceToHost));
	} else {
		std::vector<float> tmp(n);
		PK_HIP(hipMemcpy(tmp.data(), b->d<float>(b->o_out) + jb.out_off, n * 4, hipMemcpyDeviceToHost));
		for (size_t t = 0; t < n; t++) out[t] = tmp[t];
	}
	return PK_OK;
}

int pk_batch_device_images(const pk_batch *b, void **base, size_t *elem_size)
{
	if (!b) return set_err(PK_EINVAL, "batch is NULL");
	if (base) *base = b->d<void>(b->o_out);
	if (elem_size) *elem_size = b->elem;
	return PK_OK;
}

int pk_batch_image_offset(const pk_batch *b, uint32_t i, uint64_t *offset)
{
	if (!b || i >= b->n_models) return set_err(PK_EINVAL, "bad batch or image index");
	if (offset) *offset = b->jobs[i].out_off;
	return PK_OK;
}

int pk_batch_image_origin(const pk_batch *b, uint32_t i, int32_t *x, int32_t *y)
{
	if (!b || i >= b->n_models) return set_err(PK_EINVAL, "bad batch or image index");
	if (x) *x = b->origin[2 * (size_t)i];
	if (y) *y = b->origin[2 * (size_t)i + 1];
	return PK_OK;
}

int pk_evaluate_batch(pk_ctx *ctx, const pk_model *models, uint32_t n_models, double *const *outs,
                      const size_t *out_capacity, pk_result *results)
{
	if (!ctx || !models || !outs || !out_capacity) return set_err(PK_EINVAL, "NULL argument");
	pk_batch *b = nullptr;
	PK_TRY(pk_batch_create(ctx, models, n_models, 0, &b));
	int st = pk_batch_evaluate(b);
	pk_batch_info info{};
	if (st == PK_OK) st = pk_batch_sync(b, &info);
	for (uint32_t i = 0; st == PK_OK && i < n_models; i++) {
		if (!outs[i]) {
			st = set_err(PK_EINVAL, "output buffer %u is NULL", i);
			break;
		}
		st = pk_batch_get_image(b, i, outs[i], out_capacity[i]);
		if (st == PK_OK && results) {
			pk_result &r = results[i];
			std::memset(&r, 0, sizeof(r));
			r.width = b->jobs[i].ow;
			r.height = b->jobs[i].oh;
			r.offset_x = b->origin[2 * (size_t)i];
			r.offset_y = b->origin[2 * (size_t)i + 1];
		}
	}
	const std::string err = g_err;
	pk_batch_destroy(b);
	if (st != PK_OK) g_err = err;
	return st;
}

int pk_chi2_grad(pk_ctx *ctx, const pk_model *models, uint32_t n_models, const double *data, const double *sigma,
                 size_t n_values, double *chi2, double *grad)
{
	if (!ctx || !models || !data) return set_err(PK_EINVAL, "NULL argument");
	pk_batch *b = nullptr;
	PK_TRY(pk_batch_create(ctx, models, n_models, PK_BATCH_GRAD, &b));
	int st = pk_batch_set_data(b, data, sigma, n_values);
	if (st == PK_OK) st = pk_batch_evaluate_grad(b);
	if (st == PK_OK) st = pk_batch_get_chi2_grad(b, chi2, grad);
	const std::string err = g_err;
	pk_batch_destroy(b);
	if (st != PK_OK) g_err = err;
	return st;
}

}  // extern "C"
