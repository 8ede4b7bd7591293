"""CPU-side checks of the C-ABI library: it loads, exports every entry point
include/profit_hip.h declares, and its host logic (defaults, validation,
geometry) behaves like libprofit's.  No device work is launched."""
import ctypes as C

import numpy as np
import pytest

from libprofit_amd import _lib, abi


@pytest.fixture(scope="module")
def L():
    return _lib.load()


def test_exports_every_declared_symbol(L):
    names = _lib.exported_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n


def test_version(L):
    assert L.pk_version().decode() == "1.9.3"


def test_component_defaults(L):
    # constructor defaults (src/radial.cpp:724-737, per-profile constructors)
    for name, kind in abi.PROFILE_KINDS.items():
        c = abi.Component()
        assert L.pk_component_defaults(kind, C.byref(c)) == 0
        ref = abi.make_component(name)
        for f, _ in abi.Component._fields_:
            if f == "p":
                assert list(c.p) == list(ref.p), name
            elif name != "psf" or f not in ("mag",):
                assert getattr(c, f) == getattr(ref, f), (name, f)
    c = abi.Component()
    assert L.pk_component_defaults(99, C.byref(c)) == abi.PK_EINVAL


def test_model_defaults(L):
    m = abi.Model()
    assert L.pk_model_defaults(7, 9, C.byref(m)) == 0
    assert (m.width, m.height, m.scale_x, m.finesampling, m.crop, m.return_finesampled, m.adjust_mask,
            m.convolver) == (7, 9, 1.0, 1, 1, 1, 1, abi.BRUTE)


def dims(L, spec):
    w, h = C.c_uint32(), C.c_uint32()
    st = L.pk_model_output_dims(spec.ref, C.byref(w), C.byref(h))
    return st, (w.value, h.value)


def test_output_dims_and_validation_match_oracle(L, oracle):
    cases = [
        abi.ModelSpec(100, 80, [abi.make_component("sersic")]),
        abi.ModelSpec(100, 80, [abi.make_component("sersic", convolve=True)], psf=np.ones((25, 25))),
        abi.ModelSpec(100, 80, [abi.make_component("sersic", convolve=True)], psf=np.ones((4, 6)), finesampling=3),
        abi.ModelSpec(50, 50, [abi.make_component("sersic", convolve=True)], psf=np.ones((5, 5)),
                      mask=np.pad(np.ones((30, 30), np.uint8), 10)),
        abi.ModelSpec(50, 50, [abi.make_component("sersic", convolve=True)], psf=np.ones((5, 5)),
                      mask=np.pad(np.ones((48, 48), np.uint8), 1)),
        abi.ModelSpec(0, 10, [abi.make_component("sersic")]),
        abi.ModelSpec(10, 10, [abi.make_component("sersic", axrat=2.0)]),
        abi.ModelSpec(10, 10, [abi.make_component("king", rt=-1.0)]),
        abi.ModelSpec(10, 10, [abi.make_component("sersic", convolve=True)]),
        abi.ModelSpec(10, 10, [abi.make_component("sersic")], mask=np.ones((4, 4))),
        abi.ModelSpec(10, 10, [abi.make_component("psf")]),
    ]
    for spec in cases:
        st, d = dims(L, spec)
        try:
            od = oracle.output_dims(spec)
            ost, omsg = 0, ""
        except oracle.OracleError as e:
            ost, omsg, od = e.code, str(e), None
        assert st == ost
        if st == 0:
            assert d == od
        else:
            assert L.pk_last_error().decode() == omsg


def test_convolver_padding(L, oracle):
    for t in (abi.BRUTE_OLD, abi.BRUTE, abi.FFT):
        for (sw, sh, kw, kh) in ((20, 20, 2, 2), (40, 31, 7, 9), (101, 100, 25, 24)):
            v = [C.c_int32() for _ in range(4)]
            assert L.pk_convolver_padding(t, sw, sh, kw, kh, *[C.byref(x) for x in v]) == 0
            o = [C.c_int() for _ in range(4)]
            oracle.lib().or_convolver_padding(t, sw, sh, kw, kh, *[C.byref(x) for x in o])
            assert [x.value for x in v] == [x.value for x in o]


def test_python_mirror_parameters():
    from libprofit_amd.model import InvalidParameter, Model, UnknownParameter
    m = Model(10, 10)
    p = m.add_profile("sersic")
    p.parameter("xcen", 5.0)
    p.parameter("nser=4")
    p.parameter("resolution", 8)
    p.parameter("rough", True)
    p.parameter("rescale_flux=1")
    assert p.values["xcen"] == 5.0 and p.values["nser"] == 4.0 and p.values["resolution"] == 8
    with pytest.raises(UnknownParameter):
        p.parameter("foo=1")
    with pytest.raises(InvalidParameter):
        p.parameter("xcen", True)  # not a bool parameter
    with pytest.raises(InvalidParameter):
        p.parameter("nser=abc")
    with pytest.raises(InvalidParameter):
        p.parameter("nser")
    with pytest.raises(InvalidParameter):
        m.add_profile("nope")
    m.add_profile("ferrers")  # alias (src/model.cpp:117)


def test_clear_cache_removes_the_tuning_file(L, tmp_path, monkeypatch):
    # profit::clear_cache (src/library.cpp:251-270) removes the cache under $PROFIT_HOME
    monkeypatch.setenv("PROFIT_HOME", str(tmp_path))
    f = tmp_path / "hip_tuning"
    f.write_text("gfx950/256 fp32 grad0 models1 pixels100 psf0x0 comps1\t0 100 4 64 16 16 0 1.000000\n")
    assert L.pk_clear_cache() == 0
    assert not f.exists()
    assert L.pk_clear_cache() == 0  # nothing to remove
