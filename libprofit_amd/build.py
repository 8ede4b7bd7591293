"""Builds libprofit_amd/_lib/libprofit_hip.so (gfx950) with hipcc.

Usage: python -m libprofit_amd.build  (or libprofit_amd.build.build()).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libprofit_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["pk_api.hip"]
# the strip-form convolution kernels, compiled once per part of the PSF
# column steps (pk_strips.h: ST_PARTS) so that the parts build in parallel
STRIP_SOURCE = "pk_strips.hip"
STRIP_PARTS = 6
HEADERS = ["pk_math.h", "pk_profiles.h", "pk_kernels.h", "pk_device.h", "pk_strips.h"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # products and sums are not fused unless written as fma(): FP64 results
    # then round like libprofit's CPU build
    "-ffp-contract=off",
    # no packed-FP32 instructions are generated
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False, csrc=CSRC, out=LIB, defines=()):
    """Compiles `csrc` (default: the package sources) into `out`: one object
    per source and strip part, compiled in parallel, then linked."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    obj_dir = out + ".obj"
    os.makedirs(obj_dir, exist_ok=True)
    base = [HIPCC] + FLAGS + ["-D" + d for d in defines] + ["-I", os.path.join(ROOT, "include"), "-I", csrc]
    jobs = [(os.path.join(csrc, s), os.path.join(obj_dir, s + ".o"), []) for s in SOURCES]
    jobs += [(os.path.join(csrc, STRIP_SOURCE), os.path.join(obj_dir, "strips%d.o" % p), ["-DPK_STRIP_PART=%d" % p])
             for p in range(STRIP_PARTS)]
    cmds = [base + extra + ["-c", src, "-o", obj] for src, obj, extra in jobs]
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd))
    workers = max(1, min(len(cmds), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(workers) as pool:
        results = list(pool.map(lambda cmd: subprocess.run(cmd).returncode, cmds))
    for cmd, rc in zip(cmds, results):
        if rc:
            raise subprocess.CalledProcessError(rc, cmd)
    tmp = out + ".tmp"
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + [obj for _, obj, _ in jobs] + ["-o", tmp]
    if verbose:
        print(" ".join(link))
    subprocess.check_call(link)
    os.replace(tmp, out)
    return out


HOST_LIB = os.path.join(OUT_DIR, "libprofit.so")
CLI = os.path.join(OUT_DIR, "profit-cli")
HOST_SOURCES = ["profit.cpp", "fits.cpp"]
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wno-unused-function"]


def build_host(verbose=False):
    """The C++ host API (include/profit/*.h) over libprofit_hip.so, and
    profit-cli on top of it."""
    host = os.path.join(HERE, "host")
    inc = ["-I", os.path.join(ROOT, "include")]
    link = ["-L", OUT_DIR, "-l:libprofit_hip.so", "-Wl,-rpath,$ORIGIN"]
    cmds = [
        [CXX] + HOST_FLAGS + inc + ["-shared"] + [os.path.join(host, s) for s in HOST_SOURCES] + link +
        ["-o", HOST_LIB + ".tmp"],
        [CXX] + HOST_FLAGS + inc + [os.path.join(host, "profit_cli.cpp"), "-L", OUT_DIR, "-l:libprofit.so",
                                    "-l:libprofit_hip.so", "-Wl,-rpath,$ORIGIN", "-o", CLI + ".tmp"],
    ]
    for out, cmd in zip((HOST_LIB, CLI), cmds):
        if verbose:
            print(" ".join(cmd))
        subprocess.check_call(cmd)
        os.replace(out + ".tmp", out)
    return HOST_LIB


def build_host_check(verbose=False):
    """The test programs of the C++ host API: tests/cpp/host_api_check and
    tests/cpp/ref_api_check (a caller written against libprofit's API)."""
    out_dir = os.path.join(ROOT, "tests", "cpp", "_build")
    os.makedirs(out_dir, exist_ok=True)
    outs = []
    for name in ("host_api_check", "ref_api_check"):
        src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
        out = os.path.join(out_dir, name)
        cmd = [CXX] + HOST_FLAGS + ["-I", os.path.join(ROOT, "include"), src, "-L", OUT_DIR, "-l:libprofit.so",
                                    "-l:libprofit_hip.so", "-Wl,-rpath," + OUT_DIR,
                                    "-Wl,-rpath,$ORIGIN/../../../libprofit_amd/_lib", "-o", out + ".tmp"]
        if verbose:
            print(" ".join(cmd))
        subprocess.check_call(cmd)
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_all(verbose=False):
    build(verbose)
    build_host(verbose)
    build_host_check(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print(LIB)
