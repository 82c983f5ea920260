"""Builds librforest_hip.so in-tree for gfx950 with hipcc.

The library is the product: kernels (csrc/rf_kernels.hip), host
orchestration and the C ABI (csrc/rf_engine.cpp), planner and tables
(csrc/rf_plan.cpp).  GMP comes from /opt/conda (gmp.h is not installed
system-wide on this image).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "librforest_hip.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("rf_kernels.hip", "rf_engine.cpp", "rf_plan.cpp")]
GMP_PREFIX = os.environ.get("RF_GMP_PREFIX", "/opt/conda")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # exactness of the fp64 modular arithmetic relies on explicit FMAs only
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + HIPCC_FLAGS + [
        "-I" + os.path.join(GMP_PREFIX, "include"),
        "-o", LIB + ".tmp",
    ] + SOURCES + [
        # gmp.h comes from GMP_PREFIX; the library is linked by soname only
        # (no run path into GMP_PREFIX, whose older libstdc++ would shadow
        # the one libamdhip64 needs), so the system libgmp.so.10 serves it.
        "-L" + os.path.join(GMP_PREFIX, "lib"),
        "-l:libgmp.so.10",
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building librforest_hip.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
