"""ctypes mirror of the plain-data types of include/profit_hip.h.

This module only describes memory layouts and constants; it loads no
library.  The structures follow the state libprofit keeps in profit::Model
(include/profit/model.h:60-408) and in its profiles (src/radial.cpp:724-751,
src/sersic.cpp:352-360 and the other profile constructors).
"""
import ctypes as C

PK_OK, PK_EINVAL, PK_EUNKNOWN, PK_EDEVICE, PK_EFFT, PK_ENOMEM, PK_ERANGE = range(7)
PK_FP64, PK_FP32 = 0, 1
PK_INTEGRATOR_QUADPACK, PK_INTEGRATOR_EXACT = 0, 1

PROFILE_KINDS = {
    "null": 0,
    "sky": 1,
    "sersic": 2,
    "moffat": 3,
    "ferrer": 4,
    "ferrers": 4,  # alias accepted by Model::add_profile (src/model.cpp:117)
    "coresersic": 5,
    "king": 6,
    "brokenexp": 7,
    "psf": 8,
}

CONVOLVERS = {"brute-old": 0, "brute": 1, "opencl": 2, "fft": 3}
BRUTE_OLD, BRUTE, OPENCL, FFT = 0, 1, 2, 3

# profile-specific parameter slots of pk_component.p[]
PROFILE_PARAMS = {
    "null": [],
    "sky": ["bg"],
    "sersic": ["re", "nser"],
    "moffat": ["fwhm", "con"],
    "ferrer": ["rout", "a", "b"],
    "coresersic": ["re", "rb", "nser", "a", "b"],
    "king": ["rc", "rt", "a"],
    "brokenexp": ["h1", "h2", "rb", "a"],
    "psf": [],
}

MAX_PARAMS = 6


class Component(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("convolve", C.c_int32),
        ("xcen", C.c_double),
        ("ycen", C.c_double),
        ("mag", C.c_double),
        ("ang", C.c_double),
        ("axrat", C.c_double),
        ("box", C.c_double),
        ("rough", C.c_int32),
        ("adjust", C.c_int32),
        ("rescale_flux", C.c_int32),
        ("force_cpu", C.c_int32),
        ("acc", C.c_double),
        ("rscale_switch", C.c_double),
        ("rscale_max", C.c_double),
        ("resolution", C.c_uint32),
        ("max_recursions", C.c_uint32),
        ("p", C.c_double * MAX_PARAMS),
    ]


class Model(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("scale_x", C.c_double),
        ("scale_y", C.c_double),
        ("magzero", C.c_double),
        ("finesampling", C.c_uint32),
        ("crop", C.c_int32),
        ("return_finesampled", C.c_int32),
        ("adjust_mask", C.c_int32),
        ("dry_run", C.c_int32),
        ("mask", C.POINTER(C.c_uint8)),
        ("mask_width", C.c_uint32),
        ("mask_height", C.c_uint32),
        ("psf", C.POINTER(C.c_double)),
        ("psf_width", C.c_uint32),
        ("psf_height", C.c_uint32),
        ("psf_scale_x", C.c_double),
        ("psf_scale_y", C.c_double),
        ("convolver", C.c_int32),
        ("reuse_krn_fft", C.c_int32),
        ("n_components", C.c_uint32),
        ("components", C.POINTER(Component)),
    ]


class Result(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("offset_x", C.c_int32),
        ("offset_y", C.c_int32),
        ("n_refined_pixels", C.c_uint64),
        ("n_subcells", C.c_uint64),
        ("ms_setup", C.c_double),
        ("ms_profiles", C.c_double),
        ("ms_convolution", C.c_double),
        ("ms_total", C.c_double),
    ]


class Constants(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("rscale", "lumtot", "ie", "bn", "acc", "rscale_switch", "rscale_max",
                                          "flux_scale", "cos_ang", "sin_ang")] + [
        ("resolution", C.c_uint32), ("max_recursions", C.c_uint32)] + [
        (n, C.c_double) for n in ("dbn_dnser", "dlogflux_dre", "dlogflux_dnser", "dlogflux_daxrat",
                                  "dlogflux_dbox")]


class BatchInfo(C.Structure):
    _fields_ = [
        ("n_refine_entries", C.c_uint64),
        ("n_subcells", C.c_uint64),
        ("ms_setup", C.c_double),
        ("ms_coarse", C.c_double),
        ("ms_refine", C.c_double),
        ("ms_conv", C.c_double),
        ("ms_compose", C.c_double),
        ("ms_total", C.c_double),
        ("ms_grad", C.c_double),
        ("n_conv_launches", C.c_uint32),
        ("conv_tiles", C.c_uint32),
        ("ms_grad_refine", C.c_double),
        ("n_grad_subcells", C.c_uint64),
    ]


class WorkCounts(C.Structure):
    """pk_work_counts: work of one batch evaluation (rooflines)."""
    _fields_ = [(n, C.c_uint64) for n in ("pixel_component_pairs", "coarse_evaluations", "refine_evaluations",
                                          "refined_pixels")]


PK_BATCH_GRAPH = 1
PK_BATCH_GRAD = 2
PK_BATCH_AUTOTUNE = 4

PK_CONV_TILES_VALU = 0
PK_CONV_TILES_MFMA = 1
PK_CONV_TILES_STRIPS = 2


class Tuning(C.Structure):
    """pk_tuning: launch settings of a batch (autotuner)."""
    _fields_ = [
        ("conv_tiles", C.c_uint32),
        ("refine_occupancy", C.c_uint32),
        ("handover_div", C.c_uint32),
        ("claim_pixels", C.c_uint32),
        ("claim_cells", C.c_uint32),
        ("refine_passes", C.c_uint32),
        ("claim_spread", C.c_uint32),
        ("ms_evaluate", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}

# gradient slots of a component (pk_batch_get_chi2_grad)
GRAD_SLOTS = 8
GRAD_PARAMS = {
    "sersic": ["xcen", "ycen", "mag", "re", "nser", "ang", "axrat", "box"],
    "sky": ["bg"],
    "null": [],
}


RADIAL_DEFAULTS = dict(
    xcen=0.0, ycen=0.0, mag=15.0, ang=0.0, axrat=1.0, box=0.0,
    rough=False, adjust=True, rescale_flux=False, force_cpu=False,
    acc=0.1, rscale_switch=1.0, rscale_max=0.0, resolution=9, max_recursions=2,
)

PROFILE_DEFAULTS = {
    "sersic": dict(re=1.0, nser=1.0),
    "moffat": dict(fwhm=3.0, con=2.0),
    "ferrer": dict(rout=3.0, a=1.0, b=1.0, acc=1.0),
    "coresersic": dict(re=1.0, rb=1.0, nser=4.0, a=1.0, b=1.0),
    "king": dict(rc=1.0, rt=3.0, a=2.0),
    "brokenexp": dict(h1=1.0, h2=1.0, rb=1.0, a=1.0),
    "sky": dict(bg=0.0),
    "psf": dict(xcen=0.0, ycen=0.0, mag=0.0),
    "null": dict(),
}


def make_component(name, **params):
    """Builds a Component for profile `name` with libprofit's constructor
    defaults, overridden by `params` (names as in Profile::parameter)."""
    name = "ferrer" if name == "ferrers" else name
    if name not in PROFILE_PARAMS:
        raise ValueError("Unknown profile name: " + name)
    c = Component()
    c.kind = PROFILE_KINDS[name]
    vals = dict(RADIAL_DEFAULTS)
    vals["convolve"] = False
    vals.update(PROFILE_DEFAULTS[name])
    vals.update(params)
    slots = PROFILE_PARAMS[name]
    for k, v in vals.items():
        if k in slots:
            c.p[slots.index(k)] = float(v)
        elif k in ("xcen", "ycen", "mag", "ang", "axrat", "box", "acc", "rscale_switch", "rscale_max"):
            setattr(c, k, float(v))
        elif k in ("rough", "adjust", "rescale_flux", "force_cpu", "convolve"):
            setattr(c, k, int(bool(v)))
        elif k in ("resolution", "max_recursions"):
            setattr(c, k, int(v))
        else:
            raise KeyError("Profile %s doesn't support parameter %s" % (name, k))
    return c


class ModelSpec:
    """Owns the buffers a pk_model points to (components, psf, mask)."""

    def __init__(self, width, height, components=(), psf=None, mask=None, scale=(1.0, 1.0),
                 psf_scale=(1.0, 1.0), magzero=0.0, finesampling=1, crop=True,
                 return_finesampled=True, adjust_mask=True, dry_run=False,
                 convolver=BRUTE, reuse_krn_fft=False):
        import numpy as np
        self.m = Model()
        m = self.m
        m.width, m.height = int(width), int(height)
        m.scale_x, m.scale_y = float(scale[0]), float(scale[1])
        m.magzero = float(magzero)
        m.finesampling = int(finesampling)
        m.crop = int(bool(crop))
        m.return_finesampled = int(bool(return_finesampled))
        m.adjust_mask = int(bool(adjust_mask))
        m.dry_run = int(bool(dry_run))
        m.psf_scale_x, m.psf_scale_y = float(psf_scale[0]), float(psf_scale[1])
        m.convolver = int(CONVOLVERS[convolver] if isinstance(convolver, str) else convolver)
        m.reuse_krn_fft = int(bool(reuse_krn_fft))
        comps = list(components)
        self._comps = (Component * max(1, len(comps)))(*comps)
        m.n_components = len(comps)
        m.components = C.cast(self._comps, C.POINTER(Component))
        if psf is not None:
            self._psf = np.ascontiguousarray(psf, dtype=np.float64)
            m.psf_height, m.psf_width = self._psf.shape
            m.psf = self._psf.ctypes.data_as(C.POINTER(C.c_double))
        else:
            self._psf = None
        if mask is not None:
            self._mask = np.ascontiguousarray(mask, dtype=np.uint8)
            m.mask_height, m.mask_width = self._mask.shape
            m.mask = self._mask.ctypes.data_as(C.POINTER(C.c_uint8))
        else:
            self._mask = None

    @property
    def ref(self):
        return C.byref(self.m)
