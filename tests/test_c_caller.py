"""The drop-in called from C: tests/c/test_dropin.c, compiled with gcc as C11
against include/rforest_hip.h (and, where the reference's sources are
present, beside the reference's own src/include/hwmpz.h, whose declarations
of the same functions must agree), linked to librforest_hip.so, and run as
a GPU test: seeded signed products of several shapes through
mpz_rmatrix_mult_fft and cu_mpz_rmatrix_mult_fft against schoolbook GMP.
The reference's harness of the same kind is tests/test_gpu_matmul.c:18-89."""
import os
import subprocess

import pytest

from rforest_amd import build as rb

REF_INCLUDE = "/root/reference/src/include"


def test_header_compiles_as_c11_and_links(tmp_path):
    out = str(tmp_path / "caller")
    res = subprocess.run(rb.c_caller_cmd(out), capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="the reference's headers are not present here")
def test_header_agrees_with_the_reference_header():
    """Both headers in one C11 translation unit: conflicting declarations of
    mpz_rmatrix_mult_fft, cu_mpz_rmatrix_mult_fft, hw_mpz_setup/clear,
    mpz_rmatrix_gpu_fft_suitable or the memory helpers would not compile."""
    cmd = rb.c_caller_cmd("/dev/null", ["-DRF_WITH_REFERENCE_HEADER", "-I" + REF_INCLUDE])
    cmd = [c for c in cmd if not c.startswith(("-L", "-l", "-Wl"))] + ["-fsyntax-only"]
    cmd.remove("-o")
    cmd.remove("/dev/null")
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


@pytest.mark.gpu
def test_c_caller_runs_exact():
    exe = rb.C_CALLER_BIN
    if not os.path.exists(exe):
        rb.build_c_caller()
    print("case: tests/c/test_dropin seed 2024", flush=True)
    res = subprocess.run([exe, "2024"], capture_output=True, text=True, timeout=600)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "PASSED: 0 failures" in res.stdout
