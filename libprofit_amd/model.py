"""Python mirror of libprofit's Model / Profile / Convolver interface over the
C-ABI (include/profit_hip.h).

Names, argument meaning and errors follow the C++ classes:
  Model            include/profit/model.h:60-408, src/model.cpp
  Profile          include/profit/profile.h:108-197, src/profile.cpp:69-163
  create_convolver include/profit/convolve.h:194-206, src/convolve.cpp:521-584
The computation always runs in the HIP library; there is no CPU path.
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import InvalidParameter, UnknownParameter, check, load

BOOL_PARAMS = {"rough", "adjust", "force_cpu", "convolve", "rescale_flux"}
UINT_PARAMS = {"resolution", "max_recursions"}
RADIAL_DOUBLE = {"xcen", "ycen", "mag", "ang", "axrat", "box", "acc", "rscale_switch", "rscale_max"}


class Context:
    """A device + precision binding (takes the role of OpenCLEnvPtr,
    include/profit/opencl.h:79-168)."""

    def __init__(self, device=0, precision="fp64"):
        self.lib = load()
        self.precision = abi.PK_FP32 if precision in ("fp32", "float", abi.PK_FP32) else abi.PK_FP64
        h = C.c_void_p()
        check(self.lib.pk_ctx_create(int(device), self.precision, C.byref(h)))
        self.handle = h
        import weakref
        self._batches = weakref.WeakSet()

    def close(self):
        # a batch must not outlive its context
        for b in list(getattr(self, "_batches", ())):
            b.close()
        if self.handle:
            self.lib.pk_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init():
    """profit::init() (include/profit/library.h:67-70): checks for a device
    and loads the autotuner cache."""
    check(load().pk_init())
    return True


def finish():
    """profit::finish() (include/profit/library.h:78-80): writes the
    autotuner cache back."""
    load().pk_finish()


def clear_cache():
    """profit::clear_cache() (include/profit/library.h, src/library.cpp:251-270):
    forgets the autotuner cache and removes its file."""
    check(load().pk_clear_cache())


INTEGRATORS = {"quadpack": abi.PK_INTEGRATOR_QUADPACK, "exact": abi.PK_INTEGRATOR_EXACT}


def set_integrator(mode):
    """Integrator of the king / brokenexp / coresersic luminosity integrals:
    "quadpack" (libprofit's qags / qagiu at 1e-4, the default) or "exact"
    (double-exponential quadrature to 1e-14).  Read when a model's or a
    batch's set-up inputs are built."""
    check(load().pk_set_integrator(INTEGRATORS[mode]))


def get_integrator():
    v = load().pk_get_integrator()
    return {k for k, m in INTEGRATORS.items() if m == v}.pop()


_default_ctx = {}


def default_context(precision="fp64"):
    key = "fp32" if precision in ("fp32", "float", abi.PK_FP32) else "fp64"
    if key not in _default_ctx:
        _default_ctx[key] = Context(0, key)
    return _default_ctx[key]


class Convolver:
    """Convolver handle (include/profit/convolve.h:65-128)."""

    def __init__(self, type_, reuse_krn_fft=False, context=None):
        self.type = type_
        self.reuse_krn_fft = reuse_krn_fft
        self.context = context

    def convolve(self, src, krn, mask=None, crop=True):
        """Convolver::convolve -> (image, offset)."""
        ctx = self.context or default_context()
        src = np.ascontiguousarray(src, dtype=np.float64)
        krn = np.ascontiguousarray(krn, dtype=np.float64)
        H, W = src.shape
        kh, kw = krn.shape
        mx = max(W, kw)
        cap = max(4 * mx * mx, W * H, 1)
        out = np.zeros(cap, dtype=np.float64)
        mptr = None
        if mask is not None:
            mask = np.ascontiguousarray(mask, dtype=np.uint8)
            mptr = mask.ctypes.data
        ow, oh, ox, oy = C.c_uint32(), C.c_uint32(), C.c_int32(), C.c_int32()
        check(load().pk_convolve(ctx.handle, self.type, src.ctypes.data, W, H, krn.ctypes.data, kw, kh, mptr,
                                 int(bool(crop)), out.ctypes.data, cap, C.byref(ow), C.byref(oh), C.byref(ox),
                                 C.byref(oy)))
        return out[: ow.value * oh.value].reshape(oh.value, ow.value).copy(), (ox.value, oy.value)

    def padding(self, src_dims, krn_dims):
        v = [C.c_int32() for _ in range(4)]
        check(load().pk_convolver_padding(self.type, src_dims[0], src_dims[1], krn_dims[0], krn_dims[1],
                                          *[C.byref(x) for x in v]))
        return (v[0].value, v[1].value), (v[2].value, v[3].value)


def create_convolver(type_, reuse_krn_fft=False, context=None, **_prefs):
    """create_convolver(ConvolverType | str, prefs) (src/convolve.cpp:521-584)."""
    if isinstance(type_, str):
        if type_ not in abi.CONVOLVERS:
            raise InvalidParameter(abi.PK_EINVAL, "Convolver of type %s is not supported" % type_)
        type_ = abi.CONVOLVERS[type_]
    if type_ not in (0, 1, 2, 3):
        raise InvalidParameter(abi.PK_EINVAL, "Unsupported convolver type: %d" % type_)
    return Convolver(type_, reuse_krn_fft, context)


class Profile:
    """A profile of a Model with libprofit's named-parameter interface."""

    def __init__(self, model, name):
        self.model = model
        self.name = name
        kind = "ferrer" if name == "ferrers" else name
        self.kind = kind
        self.values = dict(abi.RADIAL_DEFAULTS)
        self.values["convolve"] = False
        self.values.update(abi.PROFILE_DEFAULTS[kind])
        if kind in ("sky", "null", "psf"):
            keep = {"convolve"} | set(abi.PROFILE_PARAMS[kind]) | (
                {"xcen", "ycen", "mag"} if kind == "psf" else set())
            self.values = {k: v for k, v in self.values.items() if k in keep}

    def _registry(self):
        double = set(abi.PROFILE_PARAMS[self.kind])
        if self.kind == "psf":
            double |= {"xcen", "ycen", "mag"}
        if self.kind not in ("sky", "null", "psf"):
            double |= RADIAL_DOUBLE
            return ({"rough", "adjust", "force_cpu", "convolve"} | ({"rescale_flux"} if self.kind == "sersic" else set()),
                    set(UINT_PARAMS), double)
        return {"convolve"}, set(), double

    def parameter(self, name, value=None):
        """Profile::parameter(name, bool|unsigned|double) and
        Profile::parameter("name=value") (src/profile.cpp:126-163)."""
        bools, uints, doubles = self._registry()
        if value is None:
            parts = name.split("=")
            if len(parts) != 2:
                raise InvalidParameter(abi.PK_EINVAL, "missing = in parameter: " + name)
            pname, sval = parts[0].strip(), parts[1].strip()
            try:
                if pname in bools:
                    self.values[pname] = bool(int(sval, 10))
                    return
                if pname in uints:
                    self.values[pname] = int(sval, 10)
                    return
                if pname in doubles:
                    self.values[pname] = float(sval)
                    return
            except ValueError:
                tname = "bool" if pname in bools else "unsigned int" if pname in uints else "double"
                raise InvalidParameter(abi.PK_EINVAL,
                                       "Parameter %s in profile %s is of type %s, but given value cannot be parsed as %s: %s"
                                       % (pname, self.name, tname, tname, sval))
            raise UnknownParameter(abi.PK_EUNKNOWN, "Profile %s doesn't support parameter %s, or parameter has invalid value: %s"
                                   % (self.name, pname, sval))
        if isinstance(value, (bool, np.bool_)):
            reg, tname = bools, "bool"
        elif isinstance(value, (int, np.integer)) and name in uints:
            reg, tname = uints, "unsigned int"
        else:
            reg, tname = doubles, "double"
        if name not in reg:
            raise InvalidParameter(abi.PK_EINVAL, "Unknown %s parameter in profile %s: %s" % (tname, self.name, name))
        self.values[name] = value

    def component(self):
        return abi.make_component(self.kind, **self.values)

    def get_name(self):
        return self.name

    def do_convolve(self):
        return bool(self.values.get("convolve", False))


class Model:
    """profit::Model (include/profit/model.h:60-408)."""

    def __init__(self, width=0, height=0, context=None):
        self.width, self.height = int(width), int(height)
        self.finesampling = 1
        self.scale = (1.0, 1.0)
        self.magzero = 0.0
        self.psf = None
        self.psf_scale = (1.0, 1.0)
        self.mask = None
        self.convolver = None
        self.adjust_mask = True
        self.crop = True
        self.dry_run = False
        self.return_finesampled = True
        self.omp_threads = 0
        self.profiles = []
        self.context = context
        self.last_result = None

    def add_profile(self, name):
        if name not in abi.PROFILE_KINDS:
            raise InvalidParameter(abi.PK_EINVAL, "Unknown profile name: " + name)
        p = Profile(self, name)
        self.profiles.append(p)
        return p

    def has_profiles(self):
        return len(self.profiles) > 0

    def set_dimensions(self, dims):
        self.width, self.height = int(dims[0]), int(dims[1])

    def set_finesampling(self, fs):
        if fs == 0:
            raise ValueError("finesampling must be > 0")
        self.finesampling = int(fs)

    def set_psf(self, psf):
        self.psf = np.array(psf, dtype=np.float64, copy=True)

    def set_image_pixel_scale(self, scale):
        self.scale = (float(scale[0]), float(scale[1]))

    def get_image_pixel_scale(self):
        return self.scale

    def set_psf_pixel_scale(self, scale):
        self.psf_scale = (float(scale[0]), float(scale[1]))

    def set_magzero(self, magzero):
        self.magzero = float(magzero)

    def set_mask(self, mask):
        self.mask = None if mask is None else np.array(mask, dtype=np.uint8)

    def set_adjust_mask(self, v):
        self.adjust_mask = bool(v)

    def set_convolver(self, conv):
        self.convolver = conv

    def set_crop(self, v):
        self.crop = bool(v)

    def set_dry_run(self, v):
        self.dry_run = bool(v)

    def set_return_finesampled(self, v):
        self.return_finesampled = bool(v)

    def set_omp_threads(self, n):
        self.omp_threads = int(n)  # accepted and ignored on the device path

    def spec(self):
        conv = self.convolver.type if self.convolver is not None else abi.BRUTE
        reuse = self.convolver.reuse_krn_fft if self.convolver is not None else False
        return abi.ModelSpec(self.width, self.height, [p.component() for p in self.profiles], psf=self.psf,
                             mask=self.mask, scale=self.scale, psf_scale=self.psf_scale, magzero=self.magzero,
                             finesampling=self.finesampling, crop=self.crop,
                             return_finesampled=self.return_finesampled, adjust_mask=self.adjust_mask,
                             dry_run=self.dry_run, convolver=conv, reuse_krn_fft=reuse)

    def evaluate(self, return_offset=False):
        """Model::evaluate(offset_out) (src/model.cpp:259-320)."""
        ctx = self.context or default_context()
        spec = self.spec()
        img, off, res = evaluate_spec(ctx, spec)
        self.last_result = res
        return (img, off) if return_offset else img


def evaluate_spec(ctx, spec):
    """Runs pk_evaluate on an abi.ModelSpec -> (image, offset, pk_result)."""
    L = load()
    m = spec.m
    fs = max(1, m.finesampling)
    dw = m.width * fs + (m.psf_width or 0)
    dh = m.height * fs + (m.psf_height or 0)
    mx = max(dw, dh, m.psf_width or 0)
    cap = max(4 * mx * mx, dw * dh, 1)
    out = np.zeros(cap, dtype=np.float64)
    res = abi.Result()
    check(L.pk_evaluate(ctx.handle, spec.ref, out.ctypes.data, cap, C.byref(res)))
    img = out[: res.width * res.height].reshape(res.height, res.width).copy()
    return img, (res.offset_x, res.offset_y), res


def evaluate_batch(ctx, specs):
    """pk_evaluate_batch: every model of `specs` evaluated once (a batch created
    and destroyed inside the call) -> [(image, offset)]."""
    L = load()
    specs = list(specs)
    arr = (abi.Model * len(specs))(*[s.m for s in specs])
    outs, caps = [], []
    for s in specs:
        m = s.m
        fs = max(1, m.finesampling)
        dw = m.width * fs + 2 * (m.psf_width or 0)
        dh = m.height * fs + 2 * (m.psf_height or 0)
        mx = max(dw, dh, m.psf_width or 0, m.psf_height or 0)
        outs.append(np.zeros(max(4 * mx * mx, dw * dh, 1)))
        caps.append(outs[-1].size)
    ptrs = (C.c_void_p * len(specs))(*[o.ctypes.data for o in outs])
    capa = (C.c_size_t * len(specs))(*caps)
    res = (abi.Result * len(specs))()
    check(L.pk_evaluate_batch(ctx.handle, arr, len(specs), ptrs, capa, res))
    return [(o[: r.width * r.height].reshape(r.height, r.width).copy(), (r.offset_x, r.offset_y))
            for o, r in zip(outs, res)]


def chi2_grad_once(ctx, specs, images, sigmas=None):
    """pk_chi2_grad: chi^2 of every model against its data and the gradient of
    every component (a batch created and destroyed inside the call) ->
    (chi2[n_models], grad[n_components, GRAD_SLOTS])."""
    L = load()
    specs = list(specs)
    arr = (abi.Model * len(specs))(*[s.m for s in specs])
    data = np.ascontiguousarray(np.concatenate([np.asarray(d, dtype=np.float64).ravel() for d in images]))
    sig = None
    if sigmas is not None:
        sig = np.ascontiguousarray(np.concatenate(
            [np.broadcast_to(np.asarray(sg, dtype=np.float64), np.shape(d)).ravel() for d, sg in zip(images, sigmas)]))
    ncomp = sum(int(s.m.n_components) for s in specs)
    chi2 = np.zeros(len(specs))
    grad = np.zeros((max(1, ncomp), abi.GRAD_SLOTS))
    check(L.pk_chi2_grad(ctx.handle, arr, len(specs), data.ctypes.data, None if sig is None else sig.ctypes.data,
                         data.size, chi2.ctypes.data, grad.ctypes.data))
    return chi2, grad[:ncomp]


class Batch:
    """Device-resident batch of models (pk_batch_* of include/profit_hip.h):
    all images produced by one launch sequence, optionally a hipGraph; with
    grad=True also chi^2 against set data and its gradient."""

    def __init__(self, ctx, specs, graph=True, grad=False, autotune=False):
        self.ctx = ctx
        self.L = load()
        self.specs = list(specs)  # keeps the host buffers alive
        arr = (abi.Model * len(self.specs))(*[s.m for s in self.specs])
        self._arr = arr
        h = C.c_void_p()
        flags = ((abi.PK_BATCH_GRAPH if graph else 0) | (abi.PK_BATCH_GRAD if grad else 0)
                 | (abi.PK_BATCH_AUTOTUNE if autotune else 0))
        check(self.L.pk_batch_create(ctx.handle, arr, len(self.specs), flags, C.byref(h)))
        self.handle = h
        self.n = len(self.specs)
        self.n_components = sum(int(s.m.n_components) for s in self.specs)
        ctx._batches.add(self)

    def offsets(self):
        """Element offset of every returned image in the packed output."""
        out = []
        for i in range(self.n):
            o = C.c_uint64()
            check(self.L.pk_batch_image_offset(self.handle, i, C.byref(o)))
            out.append(o.value)
        return out

    def origin(self, i):
        """(x, y) offset of image i as Model.evaluate reports it (non-zero
        only for an uncropped FFT convolution)."""
        x, y = C.c_int32(), C.c_int32()
        check(self.L.pk_batch_image_origin(self.handle, i, C.byref(x), C.byref(y)))
        return x.value, y.value

    def set_data(self, images, sigmas=None):
        """Observed images (one array per model, shaped like the returned
        image) and their sigma (arrays, scalars or None = 1)."""
        data = np.concatenate([np.asarray(d, dtype=np.float64).ravel() for d in images])
        sig = None
        if sigmas is not None:
            parts = []
            for d, sg in zip(images, sigmas):
                parts.append(np.broadcast_to(np.asarray(sg, dtype=np.float64), np.shape(d)).ravel())
            sig = np.ascontiguousarray(np.concatenate(parts))
        data = np.ascontiguousarray(data)
        check(self.L.pk_batch_set_data(self.handle, data.ctypes.data, None if sig is None else sig.ctypes.data,
                                       data.size))

    def evaluate_grad(self):
        check(self.L.pk_batch_evaluate_grad(self.handle))

    def chi2_grad(self):
        """(chi2[n_models], grad[n_components, GRAD_SLOTS]) of the last evaluate_grad."""
        chi2 = np.zeros(self.n)
        grad = np.zeros((max(1, self.n_components), abi.GRAD_SLOTS))
        check(self.L.pk_batch_get_chi2_grad(self.handle, chi2.ctypes.data, grad.ctypes.data))
        return chi2, grad[: self.n_components]

    def evaluate(self):
        check(self.L.pk_batch_evaluate(self.handle))

    def autotune(self):
        """Times candidate launch settings, keeps the fastest (pk_batch_autotune)."""
        check(self.L.pk_batch_autotune(self.handle))
        return self.tuning

    @property
    def tuning(self):
        t = abi.Tuning()
        check(self.L.pk_batch_get_tuning(self.handle, C.byref(t)))
        return t

    @tuning.setter
    def tuning(self, t):
        if isinstance(t, dict):
            cur = self.tuning
            for k, v in t.items():
                setattr(cur, k, v)
            t = cur
        check(self.L.pk_batch_set_tuning(self.handle, C.byref(t)))

    def sync(self):
        info = abi.BatchInfo()
        check(self.L.pk_batch_sync(self.handle, C.byref(info)))
        return info

    def work_counts(self):
        """pk_batch_work_counts: work of the last evaluation."""
        wc = abi.WorkCounts()
        check(self.L.pk_batch_work_counts(self.handle, C.byref(wc)))
        return wc

    def set_components(self, comps):
        arr = (abi.Component * len(comps))(*comps)
        check(self.L.pk_batch_set_components(self.handle, arr, len(comps)))

    def dims(self, i):
        w, h = C.c_uint32(), C.c_uint32()
        check(self.L.pk_batch_image_dims(self.handle, i, C.byref(w), C.byref(h)))
        return w.value, h.value

    def image(self, i):
        w, h = self.dims(i)
        out = np.zeros(w * h, dtype=np.float64)
        check(self.L.pk_batch_get_image(self.handle, i, out.ctypes.data, out.size))
        return out.reshape(h, w)

    def close(self):
        if self.handle:
            self.L.pk_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
