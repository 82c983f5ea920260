"""Builds librforest_hip.so in-tree for gfx950 with hipcc.

The library is the product: kernels (csrc/rf_ntt.hip transforms,
csrc/rf_crt.hip CRT and carries, csrc/rf_util.hip data kernels,
csrc/rf_fftops.hip transform-domain arithmetic, csrc/rf_noalloc.hip the
fixed scratch of the no-allocation entry point), host orchestration and the
C ABI (csrc/rf_engine.cpp one-shot calls, batches, operands, trees;
csrc/rf_serve.cpp the request server; csrc/rf_fft.cpp the general transform
API; csrc/rf_mod.cpp remainders; csrc/rf_hostpar.cpp the host copy pool),
planner and tables (csrc/rf_plan.cpp).  GMP comes from /opt/conda (gmp.h is not installed
system-wide on this image).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "librforest_hip.so")
SOURCES = [os.path.join(HERE, "csrc", f)
           for f in ("rf_ntt.hip", "rf_crt.hip", "rf_util.hip", "rf_fftops.hip", "rf_noalloc.hip", "rf_engine.cpp",
                     "rf_serve.cpp", "rf_plan.cpp", "rf_fft.cpp", "rf_mod.cpp", "rf_hostpar.cpp")]
OBJDIR = os.path.join(ROOT, "_build")
GMP_PREFIX = os.environ.get("RF_GMP_PREFIX", "/opt/conda")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # exactness of the fp64 modular arithmetic relies on explicit FMAs only
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


# Per-source additions: the CRT kernel for 22 and 24 primes unrolls its
# mixed-radix conversion only above LLVM's default pragma-unroll size limit;
# otherwise its Garner digits are indexed at run time and live in scratch.
# The CRT kernels are scheduled for memory clauses (their coefficient loads
# grouped): the same registers, no spills at any prime count, and k_crt<11>
# 0.985 ms against 0.999-1.011 per 1024 products (alternating runs on one
# MI355X).  The transform kernels keep the default schedule: max-ilp spills
# k_fwd<13, 4> and memory clauses slow it by 2-3 %.
SOURCE_FLAGS = {"rf_crt.hip": ["-mllvm", "-pragma-unroll-threshold=200000",
                               "-mllvm", "-amdgpu-sched-strategy=max-memory-clause"]}


def build(verbose=False):
    """Compiles every source for gfx950 in parallel, then links the library."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJDIR, exist_ok=True)
    inc = ["-I" + os.path.join(GMP_PREFIX, "include")]
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [hipcc] + HIPCC_FLAGS + SOURCE_FLAGS.get(os.path.basename(src), []) + inc + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = False
    for src, pr in procs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed compiling librforest_hip.so sources")
    link = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs + [
        # gmp.h comes from GMP_PREFIX; the library is linked by soname only
        # (no run path into GMP_PREFIX, whose older libstdc++ would shadow
        # the one libamdhip64 needs), so the system libgmp.so.10 serves it.
        "-L" + os.path.join(GMP_PREFIX, "lib"),
        "-l:libgmp.so.10",
    ]
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("linking librforest_hip.so failed")
    os.replace(LIB + ".tmp", LIB)
    return LIB


C_CALLER_SRC = os.path.join(ROOT, "tests", "c", "test_dropin.c")
C_CALLER_BIN = os.path.join(ROOT, "tests", "c", "test_dropin")


def c_caller_cmd(out, extra=()):
    """gcc command line for the C caller of the drop-in (tests/c/test_dropin.c):
    C11 against include/rforest_hip.h, linked to the library and GMP."""
    return (["gcc", "-std=gnu11", "-O2", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
             "-I" + os.path.join(GMP_PREFIX, "include")] + list(extra) +
            [C_CALLER_SRC, "-o", out, "-L" + HERE, "-lrforest_hip", "-l:libgmp.so.10",
             "-Wl,-rpath,$ORIGIN/../../rforest_amd"])


def build_c_caller():
    """Builds tests/c/test_dropin (needs the library built first)."""
    res = subprocess.run(c_caller_cmd(C_CALLER_BIN + ".tmp"), capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("building the C caller test program failed")
    os.replace(C_CALLER_BIN + ".tmp", C_CALLER_BIN)
    return C_CALLER_BIN


if __name__ == "__main__":
    print(build(verbose=True))
