/*
 * profit_hip.h -- C-ABI of the MI355X (gfx950) implementation of libprofit's
 * image-generation hot path: radial-profile evaluation with adaptive
 * sub-pixel refinement, sky/psf/null profiles, and PSF convolution
 * (BRUTE_OLD / BRUTE / FFT), behind libprofit's Model / Profile / Convolver
 * semantics.
 *
 * Conventions
 *  - Every entry point returns a status code (PK_OK == 0) and never throws.
 *    The message of the last failure on the calling thread is available from
 *    pk_last_error().  Status codes map 1:1 onto libprofit's exceptions
 *    (include/profit/exceptions.h:41-105): PK_EINVAL -> invalid_parameter,
 *    PK_EUNKNOWN -> unknown_parameter, PK_EFFT -> fft_error, PK_EDEVICE takes
 *    the role of opencl_error.
 *  - Images are row-major with idx = i + j * width
 *    (include/profit/image.h:522-530).  Host buffers are caller owned.
 *  - Masks are one byte per pixel, non-zero meaning "evaluate this pixel"
 *    (include/profit/image.h Mask = surface<bool>).
 *  - A context binds one device and one HIP stream (a batch's convolution
 *    also forks onto a second stream of the context and joins back within the
 *    call); calls on one context are serialised internally, separate contexts
 *    may run concurrently.
 */
#ifndef PROFIT_HIP_H
#define PROFIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PK_VERSION_MAJOR 1
#define PK_VERSION_MINOR 9
#define PK_VERSION_PATCH 3

/* ---- status codes ------------------------------------------------------ */
enum {
	PK_OK = 0,
	PK_EINVAL = 1,    /* profit::invalid_parameter */
	PK_EUNKNOWN = 2,  /* profit::unknown_parameter */
	PK_EDEVICE = 3,   /* HIP runtime / device failure (role of opencl_error) */
	PK_EFFT = 4,      /* profit::fft_error */
	PK_ENOMEM = 5,    /* device or host allocation failed; never truncates */
	PK_ERANGE = 6     /* output buffer too small / work queue capacity exceeded */
};

/* ---- precision of the evaluation (context property) -------------------- */
enum {
	PK_FP64 = 0,  /* everything in FP64: per-pixel parity with libprofit <= 1e-12 */
	PK_FP32 = 1   /* FP32 bulk evaluation + convolution, FP64 constants/decisions */
};

/* ---- profile kinds (src/model.cpp:103-135 names) ------------------------ */
enum {
	PK_PROFILE_NULL = 0,
	PK_PROFILE_SKY = 1,
	PK_PROFILE_SERSIC = 2,
	PK_PROFILE_MOFFAT = 3,
	PK_PROFILE_FERRER = 4,
	PK_PROFILE_CORESERSIC = 5,
	PK_PROFILE_KING = 6,
	PK_PROFILE_BROKENEXP = 7,
	PK_PROFILE_PSF = 8
};

/* ---- convolver types (include/profit/convolve.h:44-57) ------------------ */
enum {
	PK_CONV_BRUTE_OLD = 0,
	PK_CONV_BRUTE = 1,
	PK_CONV_OPENCL = 2,   /* accepted as an alias of BRUTE (device direct convolution) */
	PK_CONV_FFT = 3
};

/*
 * Index of the profile-specific parameters inside pk_component.p[].
 * Radial parameters shared by all radial profiles are named fields
 * (src/radial.cpp:724-751).
 */
enum {
	/* sersic (src/sersic.cpp:352-360) */
	PK_SERSIC_RE = 0, PK_SERSIC_NSER = 1,
	/* moffat (src/moffat.cpp) */
	PK_MOFFAT_FWHM = 0, PK_MOFFAT_CON = 1,
	/* ferrer (src/ferrer.cpp) */
	PK_FERRER_ROUT = 0, PK_FERRER_A = 1, PK_FERRER_B = 2,
	/* coresersic (src/coresersic.cpp) */
	PK_CORESERSIC_RE = 0, PK_CORESERSIC_RB = 1, PK_CORESERSIC_NSER = 2,
	PK_CORESERSIC_A = 3, PK_CORESERSIC_B = 4,
	/* king (src/king.cpp) */
	PK_KING_RC = 0, PK_KING_RT = 1, PK_KING_A = 2,
	/* brokenexp (src/brokenexponential.cpp) */
	PK_BROKENEXP_H1 = 0, PK_BROKENEXP_H2 = 1, PK_BROKENEXP_RB = 2,
	PK_BROKENEXP_A = 3,
	/* sky (src/sky.cpp): requested bg */
	PK_SKY_BG = 0
};

#define PK_MAX_PARAMS 6

/*
 * One profile of a model: the parameters libprofit registers through
 * Profile::register_parameter (src/profile.cpp:69-83, src/radial.cpp:738-751).
 * Defaults are those of the libprofit constructors; pk_component_defaults()
 * fills them in.
 */
typedef struct pk_component {
	int32_t kind;          /* PK_PROFILE_* */
	int32_t convolve;      /* Profile "convolve" (src/profile.cpp:44) */
	/* radial / psf common parameters */
	double xcen, ycen, mag, ang, axrat, box;
	int32_t rough, adjust, rescale_flux, force_cpu;
	double acc, rscale_switch, rscale_max;
	uint32_t resolution, max_recursions;
	double p[PK_MAX_PARAMS];
} pk_component;

/*
 * A model: the state libprofit keeps in profit::Model
 * (include/profit/model.h:60-408) plus the selected convolver.
 */
typedef struct pk_model {
	uint32_t width, height;          /* requested dimensions */
	double scale_x, scale_y;         /* image pixel scale */
	double magzero;
	uint32_t finesampling;           /* >= 1 */
	int32_t crop;                    /* Model::set_crop, default 1 */
	int32_t return_finesampled;      /* default 1 */
	int32_t adjust_mask;             /* default 1 */
	int32_t dry_run;                 /* default 0 */
	const uint8_t *mask;             /* NULL = no mask */
	uint32_t mask_width, mask_height;
	const double *psf;               /* as given to Model::set_psf (normalised inside) */
	uint32_t psf_width, psf_height;
	double psf_scale_x, psf_scale_y;
	int32_t convolver;               /* PK_CONV_* ; default BRUTE (src/model.cpp:144-150) */
	int32_t reuse_krn_fft;
	uint32_t n_components;
	const pk_component *components;
} pk_model;

/* Result metadata of an evaluation. */
typedef struct pk_result {
	uint32_t width, height;    /* dimensions of the returned image */
	int32_t offset_x, offset_y;/* Model::evaluate(offset_out) (src/model.cpp:259-320) */
	uint64_t n_refined_pixels; /* pixel/component pairs that entered sub-pixel refinement */
	uint64_t n_subcells;       /* evaluate_at calls made by the refinement */
	double ms_setup, ms_profiles, ms_convolution, ms_total; /* device timings */
} pk_result;

typedef struct pk_ctx pk_ctx;

/* ---- library ----------------------------------------------------------- */
/* profit::version() (include/profit/library.h:38-60) */
const char *pk_version(void);
/* profit::init() / profit::finish() (include/profit/library.h:67-80):
 * device check plus loading / saving the autotuner cache (below). */
int pk_init(void);
void pk_finish(void);
/* Message of the last error raised on this thread ("" if none). */
const char *pk_last_error(void);
/* Number of visible HIP devices (replaces get_opencl_info, include/profit/opencl.h:124-168). */
int pk_device_count(int *count);

/* ---- context: replaces OpenCLEnvPtr / get_opencl_environment ------------
 * (include/profit/opencl.h:79-168, consulted by src/radial.cpp:269-273). */
int pk_ctx_create(int device, int precision, pk_ctx **ctx_out);
int pk_ctx_destroy(pk_ctx *ctx);
int pk_ctx_precision(const pk_ctx *ctx);
/* The context's device: its architecture name and compute-unit count
 * ("gfx950/256"); in the role of OpenCLEnv::get_device_name. */
const char *pk_ctx_device_name(const pk_ctx *ctx);

/* Fills a component with libprofit's constructor defaults for `kind`
 * (src/radial.cpp:724-737, src/sersic.cpp:352-360, src/moffat.cpp,
 * src/ferrer.cpp:116-124, src/king.cpp, src/brokenexponential.cpp,
 * src/coresersic.cpp, src/sky.cpp, src/psf.cpp). */
int pk_component_defaults(int kind, pk_component *c);
/* Fills a model with profit::Model's defaults (src/model.cpp:54-73). */
int pk_model_defaults(uint32_t width, uint32_t height, pk_model *m);

/* Validates a model exactly as Model::analyze_inputs + Profile::validate do
 * (src/model.cpp:186-219, src/radial.cpp:204-214 and the per-profile validate
 * methods) and reports its drawing dimensions (requested dimensions times
 * finesampling plus the PSF padding, src/model.cpp:183). */
int pk_model_output_dims(const pk_model *m, uint32_t *width, uint32_t *height);

/* Per-component constants after libprofit's set-up
 * (RadialProfile::initial_calculations, src/radial.cpp:137-199, with the
 * profile overrides): normalisation and the adjusted sub-pixel integration
 * parameters that libprofit writes back into the profile.  Host-side only. */
typedef struct pk_constants {
	double rscale, lumtot, ie, bn, acc, rscale_switch, rscale_max, flux_scale, cos_ang, sin_ang;
	uint32_t resolution, max_recursions;
	/* Sersic only, for the chi^2 gradient: d bn / d nser and the derivatives
	 * of log(flux_scale) with respect to re, nser, axrat and box */
	double dbn_dnser, dlogflux_dre, dlogflux_dnser, dlogflux_daxrat, dlogflux_dbox;
} pk_constants;
int pk_component_constants(const pk_model *m, uint32_t index, pk_constants *out);

/* Model::evaluate(Point &offset_out) (src/model.cpp:259-320).
 * Writes the model image (FP64, row-major) into `out`, which must hold at
 * least `out_capacity` doubles.  `res` (may be NULL) receives dims, offset
 * and statistics. */
int pk_evaluate(pk_ctx *ctx, const pk_model *m, double *out, size_t out_capacity,
                pk_result *res);
/* pk_evaluate with the returned image (FP64) written to device memory of the
 * context's device (out_dev, out_capacity doubles); returns when it is
 * there. */
int pk_evaluate_device(pk_ctx *ctx, const pk_model *m, double *out_dev, size_t out_capacity, pk_result *res);

/* pk_evaluate_device that allocates nothing: no device memory, no pinned
 * host memory (the context's own workspace is not used either).  No
 * counterpart in the reference: for callers that cannot allocate -- code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Its workspace is a fixed array in device memory of
 * PK_NOALLOC_SCRATCH_BYTES per device (held from the library's load until the
 * process ends) and a host staging area of PK_NOALLOC_STAGE_BYTES; a model
 * needing more fails with PK_ERANGE.  Roughly: 28 bytes per drawing pixel
 * in FP64 (44 with a PSF: up to about 1700 x 1700 drawing pixels), or 64 MiB
 * of refinement queues + 48 bytes per drawing pixel in FP32 (about
 * 1150 x 1150), plus the FFT buffer of an FFT-convolved model, must fit the
 * former; the mask, PSF and model description the latter.  Calls
 * are serialised by a process-wide lock; the workspace keeps the last call's
 * contents.  Same results as pk_evaluate_device. */
#define PK_NOALLOC_SCRATCH_BYTES (128u << 20)
#define PK_NOALLOC_STAGE_BYTES (16u << 20)
int pk_evaluate_device_noalloc(pk_ctx *ctx, const pk_model *m, double *out_dev, size_t out_capacity,
                               pk_result *res);

/* Sub-pixel integrations started by the last pk_evaluate / pk_evaluate_device
 * / pk_evaluate_device_noalloc on this context, per component of its model
 * (component order; 0 for sky, psf and null profiles): the pixels whose
 * subsample_pixel recursion began (level 0 of libprofit's
 * Model::get_profile_integrations, include/profit/model.h:116). */
int pk_ctx_last_integrations(const pk_ctx *ctx, uint64_t *counts, uint32_t n_components);

/* Convolver::convolve(src, krn, mask, crop, offset_out)
 * (include/profit/convolve.h:98-99, src/convolve.cpp:55-79) for a convolver of
 * type `type`; the kernel is used as given (no normalisation). */
int pk_convolve(pk_ctx *ctx, int type, const double *src, uint32_t src_w, uint32_t src_h,
                const double *krn, uint32_t krn_w, uint32_t krn_h,
                const uint8_t *mask, int crop, double *out, size_t out_capacity,
                uint32_t *out_w, uint32_t *out_h, int32_t *offset_x, int32_t *offset_y);

/* Convolver::padding (src/convolve.cpp:50-53, 263-268). */
int pk_convolver_padding(int type, uint32_t src_w, uint32_t src_h, uint32_t krn_w, uint32_t krn_h,
                         int32_t *lo_x, int32_t *lo_y, int32_t *hi_x, int32_t *hi_y);

/* ---- batched, device-resident evaluation --------------------------------
 * No counterpart in libprofit, which evaluates one Model per call: a batch
 * holds the work list of many models (each with its own dimensions,
 * profiles and PSF) resident in device memory and produces all their images
 * in one sequence of launches (optionally captured in a hipGraph), with the
 * per-component set-up (bn, normalisation integrals, adjusted subsampling
 * constants) computed in FP64 on the device.  Supported per model: radial,
 * sky, psf and null profiles, BRUTE/BRUTE_OLD/FFT convolution, cropped or
 * not, calculation masks, finesampling, dry runs (a zero image of the
 * drawing dimensions).  PK_BATCH_GRAD batches take sersic,
 * sky and null profiles, the BRUTE convolvers, and the FFT convolver when it
 * crops and wraps no rows (drawing height + PSF height - 1 <= 2 max(drawing
 * width, PSF width)).  Output images
 * are FP32 in a PK_FP32 context and FP64 in a PK_FP64 one. */
typedef struct pk_batch pk_batch;

enum {
	PK_BATCH_GRAPH = 1,    /* capture the launch sequence once, replay it */
	PK_BATCH_GRAD = 2,     /* also allocate what pk_batch_evaluate_grad needs */
	PK_BATCH_AUTOTUNE = 4  /* tune the launch settings at the first evaluation
	                          (pk_batch_autotune, or the cached result for the
	                          same device and workload) */
};

/* Launch settings of a batch.  None of them changes which pixels are refined
 * or how; the refinement settings do not change the result at all (its sums
 * are exact fixed-point), the matrix-core convolution forms change the FP32
 * rounding of convolved pixels only.
 * Defaults: VALU tiles, 100, 4, 256, 256, 5, 0. */
enum {
	PK_CONV_TILES_VALU = 0,   /* FP32 FMA tiles */
	PK_CONV_TILES_MFMA = 1,   /* matrix-core tiles (banded PSF rows) */
	PK_CONV_TILES_STRIPS = 2  /* matrix-core strips (PSF rows as the matrix rows; PSF sides <= 48,
	                             larger ones use the banded tiles; the adjoint's correlation too when
	                             a gradient batch holds one) */
};
typedef struct pk_tuning {
	uint32_t conv_tiles;       /* FP32 direct convolution: PK_CONV_TILES_* (FP64 always uses VALU tiles) */
	uint32_t refine_occupancy; /* FP32 refinement wavefronts, percent (1..100) of the most that fit the GPU */
	uint32_t handover_div;     /* no effect (kept for the layout): each refinement pass refines one
	                              level and hands all failing sub-cells to the next (>= 1) */
	uint32_t claim_pixels;     /* refinement pixels one wavefront claims at once, at most (>= 1) */
	uint32_t claim_cells;      /* cells of the previous level one wavefront claims at once, at most (>= 1) */
	uint32_t refine_passes;    /* FP32 refinement launches, 4..7: one per level for levels 0 .. n - 2 (a level's
	                              failing sub-cells are the next launch's work), the last one every deeper
	                              level depth-first; a batch raises it where its components' max_recursions
	                              and resolution need more level launches for the last one's cell stack */
	uint32_t claim_spread;     /* 0: refinement pixels are claimed in the order the pixel pass listed
	                              them; a power of two n <= 4096: in runs of n taken from 64 places of
	                              the list in turn (spreads deep galaxy centres over wavefronts) */
	double ms_evaluate;        /* device ms of one evaluation with these settings, when measured (else 0);
	                              a PK_BATCH_GRAD batch tuned with its data: forward + adjoint */
} pk_tuning;

typedef struct pk_batch_info {
	uint64_t n_refine_entries;  /* pixels that entered sub-pixel refinement */
	uint64_t n_subcells;        /* evaluate_at calls made by the refinement */
	/* per-stage device time of the last evaluation (not available in graph mode: 0) */
	double ms_setup, ms_coarse, ms_refine, ms_conv, ms_compose, ms_total;
	double ms_grad;             /* adjoint part of the last pk_batch_evaluate_grad */
	uint32_t n_conv_launches;   /* convolution launches per evaluation (one per PSF-size class) */
	uint32_t conv_tiles;        /* PK_CONV_TILES_* they used */
	double ms_grad_refine;      /* refinement adjoint of the last pk_batch_evaluate_grad (eager mode) */
	uint64_t n_grad_subcells;   /* evaluate_at calls (value and partials) of that refinement adjoint */
} pk_batch_info;

int pk_batch_create(pk_ctx *ctx, const pk_model *models, uint32_t n_models, uint32_t flags, pk_batch **out);
/* Replaces the parameters of all components (the models' components in
 * order, same kinds as at creation); e.g. one step of a fit. */
int pk_batch_set_components(pk_batch *b, const pk_component *components, uint32_t n_components);
/* Enqueues one evaluation of every model on the context's stream. */
int pk_batch_evaluate(pk_batch *b);
/* Waits for the last evaluation; fails with PK_ERANGE if a refinement stack
 * overflowed.  `info` may be NULL. */
int pk_batch_sync(pk_batch *b, pk_batch_info *info);
/* Work of one evaluation, for rooflines: (pixel, radial component) pairs the
 * pixel pass classifies, evaluate_at calls it makes outside the refinement
 * (counted by a separate classification pass launched here, with the last
 * evaluation's component constants), and the refinement's evaluate_at calls
 * and refined pixel/component pairs of the last evaluation.  No counterpart
 * in the reference. */
typedef struct pk_work_counts {
	uint64_t pixel_component_pairs, coarse_evaluations, refine_evaluations, refined_pixels;
} pk_work_counts;
int pk_batch_work_counts(pk_batch *b, pk_work_counts *out);
int pk_batch_image_dims(const pk_batch *b, uint32_t i, uint32_t *width, uint32_t *height);
/* Copies image i (converted to FP64) to host memory. */
int pk_batch_get_image(pk_batch *b, uint32_t i, double *out, size_t out_capacity);
/* Device address of all output images (packed, model order), their element
 * size in bytes (4 or 8) and the element offset of image i. */
int pk_batch_device_images(const pk_batch *b, void **base, size_t *elem_size);
int pk_batch_image_offset(const pk_batch *b, uint32_t i, uint64_t *offset);
/* The offset Model::evaluate(Point &offset_out) reports for image i
 * (src/model.cpp:259-318): non-zero only for an uncropped FFT convolution. */
int pk_batch_image_origin(const pk_batch *b, uint32_t i, int32_t *x, int32_t *y);
int pk_batch_destroy(pk_batch *b);

/* ---- autotuner -------------------------------------------------------------
 * No counterpart in libprofit (its only tuned state is FFTW wisdom, imported
 * by init() and exported by finish(), src/library.cpp:96-165).  The tuner
 * times eager evaluations of the batch over candidate settings, one setting
 * at a time (convolution tiles, then hand-on divisor, occupancy, claim sizes,
 * poll interval, claim spread), keeps the fastest and leaves the images of one more
 * evaluation with it.  Results are cached per device and workload shape
 * (precision, gradient flag, model count, pixel count, largest PSF, component
 * count); pk_init loads the cache from $PROFIT_HOME/hip_tuning (PROFIT_HOME
 * defaults to $HOME/.profit, as in src/utils.cpp:471-495) and pk_finish
 * writes it back; pk_clear_cache empties it and removes the file
 * (profit::clear_cache, src/library.cpp:251-270). */
int pk_batch_autotune(pk_batch *b);
int pk_batch_get_tuning(const pk_batch *b, pk_tuning *t);
/* Replaces the settings (validated; a captured graph is captured again). */
int pk_batch_set_tuning(pk_batch *b, const pk_tuning *t);
int pk_clear_cache(void);

/* ---- chi^2 and its gradient ----------------------------------------------
 * No counterpart in libprofit (fits with it use finite differences, e.g. in
 * ProFit).  For every model of a PK_BATCH_GRAD batch:
 *   chi2 = sum_p ((data_p - model_p) / sigma_p)^2 over its returned image,
 * and d chi2 / d parameter for every component, by a hand-written adjoint of
 * the evaluation: returned image -> crop/downsample -> PSF correlation ->
 * directly evaluated pixels and the sub-pixel refinement.  The refinement
 * pattern (which sub-cells are refined, the adjusted resolution,
 * rscale_switch, rscale_max and acc) is that of the evaluation at the given
 * parameters and is held fixed: the gradient is that of the piecewise-smooth
 * image with this pattern.  Gradients are available for sersic and sky
 * profiles (null profiles: zeros), any PSF dimensions.  With a
 * calculation mask, the pixels it leaves out (and the returned pixels the
 * adjusted mask zeroes) carry no gradient; chi2 still counts them, with
 * model 0 (give them a large sigma to leave them out of the fit).
 * Gradient slots per component (PK_GRAD_SLOTS doubles, unused slots 0):
 *   sersic: xcen, ycen, mag, re, nser, ang (degrees), axrat, box
 *   sky:    bg */
#define PK_GRAD_SLOTS 8
enum {
	PK_GRAD_XCEN = 0, PK_GRAD_YCEN = 1, PK_GRAD_MAG = 2, PK_GRAD_RE = 3, PK_GRAD_NSER = 4,
	PK_GRAD_ANG = 5, PK_GRAD_AXRAT = 6, PK_GRAD_BOX = 7,
	PK_GRAD_SKY_BG = 0
};
/* Observed images and their per-pixel sigma (> 0), host arrays of n doubles
 * packed like the batch's returned images (pk_batch_image_offset); sigma
 * NULL means 1 everywhere.  Kept in device memory until replaced. */
int pk_batch_set_data(pk_batch *b, const double *data, const double *sigma, size_t n);
/* Enqueues one evaluation of every model followed by chi2 and its gradient
 * (in a PK_BATCH_GRAPH batch: one more captured graph). */
int pk_batch_evaluate_grad(pk_batch *b);
/* Waits and copies chi2[n_models] and grad[n_components * PK_GRAD_SLOTS]
 * (components in the order given at creation). */
int pk_batch_get_chi2_grad(pk_batch *b, double *chi2, double *grad);

/* ---- one-shot forms --------------------------------------------------------
 * No counterpart in libprofit (a caller loops over Model::evaluate).  Each
 * call creates a batch, runs it once and destroys it (device allocation per
 * call); repeated calls on the same models should keep a pk_batch instead. */
/* Evaluates every model; image i (FP64, returned dims) goes to outs[i], which
 * holds out_capacity[i] doubles; results[i] (may be NULL) gets its dims and
 * offset. */
int pk_evaluate_batch(pk_ctx *ctx, const pk_model *models, uint32_t n_models, double *const *outs,
                      const size_t *out_capacity, pk_result *results);
/* chi2[n_models] and grad[n_components * PK_GRAD_SLOTS] of the models against
 * data and sigma (n_values doubles each, packed like the batch's returned
 * images; sigma NULL = 1), as pk_batch_set_data + pk_batch_evaluate_grad +
 * pk_batch_get_chi2_grad. */
int pk_chi2_grad(pk_ctx *ctx, const pk_model *models, uint32_t n_models, const double *data, const double *sigma,
                 size_t n_values, double *chi2, double *grad);

/* ---- FP64 special functions of the math back-end ------------------------
 * include/profit/utils.h:52-113 (GSL/Rmath in libprofit, own implementations
 * here; the same code runs on the device during per-component set-up). */
double pk_qgamma(double p, double shape);
double pk_pgamma(double q, double shape);
double pk_gammafn(double x);
double pk_beta(double a, double b);
typedef double (*pk_integration_func_t)(double x, void *params);
double pk_integrate_qags(pk_integration_func_t f, double a, double b, void *params);
double pk_integrate_qagi(pk_integration_func_t f, double a, void *params);

/* Integrator of the luminosity integrals of the king, brokenexp and
 * coresersic profiles and of pk_integrate_qags / pk_integrate_qagi
 * (process-wide, read when a model's or batch's set-up inputs are built).
 *   PK_INTEGRATOR_QUADPACK (default): libprofit's -- QUADPACK qags (21-point
 *     Gauss-Kronrod) on [a, b] and qagiu (15-point rule on f(a + (1-t)/t)/t^2
 *     over (0, 1]) with Wynn epsilon extrapolation, epsabs = epsrel = 1e-4,
 *     at most 100 intervals, as GSL runs them (src/utils.cpp:155-186).  The
 *     result carries the quadrature error libprofit's has (up to ~1e-4).
 *   PK_INTEGRATOR_EXACT: double-exponential quadrature (tanh-sinh on [a, b],
 *     exp-sinh on [a, inf)) to a relative 1e-14: the integral itself.
 * No counterpart in the reference (it always uses the first). */
#define PK_INTEGRATOR_QUADPACK 0
#define PK_INTEGRATOR_EXACT 1
int pk_set_integrator(int mode);
int pk_get_integrator(void);

#ifdef __cplusplus
}
#endif

#endif /* PROFIT_HIP_H */
