"""The C++ host API (include/profit/*.h, libprofit_amd/_lib/libprofit.so) and
profit-cli: host-side checks on the CPU, device results against the FP64
oracle on the GPU."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from libprofit_amd import abi
from parity_util import TOL_FP32, TOL_FP64, comp, moffat_psf, rel_errors, sersic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "_build", "host_api_check")
CLI = os.path.join(ROOT, "libprofit_amd", "_lib", "profit-cli")


def _need(path):
    if not os.path.exists(path):
        pytest.fail("%s is not built: run __graft_entry__.build()" % path)


def test_host_api_cpu_checks():
    _need(CHECK)
    r = subprocess.run([CHECK, "cpu"], capture_output=True, text=True, timeout=60)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "checks passed" in r.stdout


def test_cli_help_version_and_errors():
    _need(CLI)
    r = subprocess.run([CLI, "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "-p [count,]name" in r.stdout
    r = subprocess.run([CLI, "-V"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "libprofit version 1.9.3" in r.stdout
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1  # no profiles
    r = subprocess.run([CLI, "-p", "foo"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Unknown profile name: foo" in r.stderr
    r = subprocess.run([CLI, "-p", "sersic", "-P", "3:3:1,2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Not enough values provided for PSF" in r.stderr


def _load(path):
    raw = open(path, "rb").read()
    w, h = np.frombuffer(raw[:8], dtype=np.uint32)
    ox, oy = np.frombuffer(raw[8:16], dtype=np.int32)
    return np.frombuffer(raw[16:], dtype=np.float64).reshape(int(h), int(w)), (int(ox), int(oy))


@pytest.mark.gpu
def test_host_api_gpu_matches_oracle(oracle):
    _need(CHECK)
    print("CASE host api gpu", flush=True)
    d = tempfile.mkdtemp(prefix="pk_host_")
    r = subprocess.run([CHECK, "gpu", d], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr

    def check(name, spec, tol, conv=None):
        got, goff = _load(os.path.join(d, name + ".bin"))
        ref, roff = oracle.evaluate(spec)
        assert got.shape == ref.shape, (name, got.shape, ref.shape)
        assert goff == roff, (name, goff, roff)
        pix, flux = rel_errors(got, ref)
        print(name, pix, flux, flush=True)
        assert pix <= tol, (name, pix)

    check("c1", abi.ModelSpec(100, 100, [sersic(xcen=50.3, ycen=49.7, mag=15, re=10, nser=4, ang=30, axrat=0.6,
                                                box=0)]), TOL_FP64)
    psf_ref, _ = oracle.evaluate(moffat_psf(15, 3.0, 2.5))
    psf, _ = _load(os.path.join(d, "psf15.bin"))
    assert np.max(np.abs(psf - psf_ref)) <= 1e-12 * psf_ref.max()
    for conv in ("brute", "brute-old", "fft"):
        for crop in (True, False):
            spec = abi.ModelSpec(96, 80, [sersic(xcen=47.1, ycen=41.4, re=7, nser=2, axrat=0.6, ang=10,
                                                 convolve=True), comp("sky", bg=1e-9)], psf=psf, convolver=conv,
                                 crop=crop)
            check("conv_%s%s" % (conv, "" if crop else "_nocrop"), spec, TOL_FP64 if conv != "fft" else 1e-9)
    mask = np.ones((40, 50), dtype=np.uint8)
    mask[0::3, 0::2] = 0
    check("mask_fs2", abi.ModelSpec(50, 40, [sersic(xcen=20.2, ycen=22.7, re=6, nser=3, axrat=0.5, ang=70)],
                                    mask=mask, finesampling=2, return_finesampled=False), TOL_FP64)
    src = np.fmod(np.arange(37 * 41) * 0.37, 1.0).reshape(41, 37)
    krn = np.fmod(np.arange(9 * 6) * 0.61, 1.0).reshape(6, 9)
    got, _ = _load(os.path.join(d, "convolver_brute.bin"))
    ref, _ = oracle.convolve(src, krn, conv="brute")
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    got, goff = _load(os.path.join(d, "convolver_fft_nocrop.bin"))
    ref, roff = oracle.convolve(src, krn, crop=False, conv="fft")
    assert got.shape == ref.shape and goff == roff
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))
    check("c1_fp32", abi.ModelSpec(100, 100, [sersic(xcen=50.3, ycen=49.7, mag=15, re=10, nser=4, ang=30,
                                                     axrat=0.6)]), TOL_FP32)


@pytest.mark.gpu
def test_cli_matches_oracle(oracle):
    """profit-cli -t output and its timing line (src/profit-cli.cpp:389-412)."""
    _need(CLI)
    print("CASE cli", flush=True)
    args = [CLI, "-w", "60", "-H", "50", "-p", "sersic:xcen=30.3:ycen=24.1:mag=15:re=6:nser=2.5:ang=20:axrat=0.7",
            "-p", "sky:bg=1e-8", "-P", "3:3:1,2,1,2,4,2,1,2,1", "-p", "moffat:xcen=20:ycen=20:fwhm=4:con=3:convolve=1",
            "-T", "brute", "-i", "3", "-t"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout[:400], r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].startswith("Created convolver in")
    assert lines[1].startswith("Ran 3 iterations in") and "[ms] per iteration)" in lines[1]
    got = np.array([[float(v) for v in ln.split()] for ln in lines[2:] if ln.strip()])
    psf = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]], dtype=float)
    spec = abi.ModelSpec(60, 50, [sersic(xcen=30.3, ycen=24.1, mag=15, re=6, nser=2.5, ang=20, axrat=0.7),
                                  comp("sky", bg=1e-8),
                                  comp("moffat", xcen=20, ycen=20, fwhm=4, con=3, convolve=True)], psf=psf)
    ref, _ = oracle.evaluate(spec)
    pix, _ = rel_errors(got, ref)
    assert pix <= 1e-13, pix  # the text output carries 17 significant digits


REF_CHECK = os.path.join(ROOT, "tests", "cpp", "_build", "ref_api_check")


@pytest.mark.gpu
def test_reference_api_caller_matches_oracle(oracle):
    """tests/cpp/ref_api_check, written against libprofit's public API only
    (ProfileStats::total, RadialProfileStats, OpenCLEnvPtr via
    get_opencl_environment / Model::set_opencl_env, the 7-argument
    ConvolverCreationPreferences, has_simd_instruction_set,
    opencl_version_*, get_opencl_info, get_profile_integrations, the
    profit-cli run loop), compiled against this library's headers: its
    image against the oracle, its level-0 integrations against the oracle's
    refined pixels of each profile."""
    _need(REF_CHECK)
    print("CASE reference api caller", flush=True)
    d = tempfile.mkdtemp(prefix="pk_ref_")
    out = os.path.join(d, "img.bin")
    r = subprocess.run([REF_CHECK, out, "3"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "Ran 3 iterations in" in r.stdout
    assert "Stats for profile sersic" in r.stdout and "Stats for profile moffat" in r.stdout
    psf, _ = oracle.evaluate(moffat_psf(15, 3.0, 2.5))
    s = comp("sersic", xcen=40.3, ycen=34.7, mag=15, re=6, nser=2.5, ang=30, axrat=0.6, box=0.1, convolve=True)
    m = comp("moffat", xcen=30.2, ycen=41.6, mag=16, fwhm=4.0, con=3.0, convolve=True)
    spec = abi.ModelSpec(80, 70, [s, m, comp("sky", bg=1e-9)], psf=psf)
    got, goff = _load(out)
    ref, roff = oracle.evaluate(spec)
    assert got.shape == ref.shape and goff == roff
    pix, flux = rel_errors(got, ref)
    assert pix <= TOL_FP64 and flux <= TOL_FP64, (pix, flux)
    # level-0 integrations per profile: the pixels the oracle refines for
    # that profile alone (same PSF padding)
    for name, c in (("sersic", s), ("moffat", m)):
        oracle.evaluate(abi.ModelSpec(80, 70, [c], psf=psf))
        refined, _ = oracle.stats()
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("Integrations per recursion level for profile "
                                                                     + name)]
        assert line, name
        lines = r.stdout.splitlines()
        lv0 = lines[lines.index(line[0]) + 1]
        assert lv0.strip() == "Level 0: %d integrations" % refined, (name, lv0, refined)
