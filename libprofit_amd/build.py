"""Builds libprofit_amd/_lib/libprofit_hip.so (gfx950) with hipcc.

Usage: python -m libprofit_amd.build  (or libprofit_amd.build.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libprofit_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["pk_api.hip"]
HEADERS = ["pk_math.h", "pk_profiles.h", "pk_kernels.h", "pk_device.h"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # products and sums are not fused unless written as fma(): FP64 results
    # then round like libprofit's CPU build
    "-ffp-contract=off",
    # no packed-FP32 instructions are generated
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False, csrc=CSRC, out=LIB, defines=()):
    """Compiles `csrc` (default: the package sources) into `out`."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-D" + d for d in defines] + ["-I", os.path.join(ROOT, "include"), "-I", csrc]
    cmd += [os.path.join(csrc, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
