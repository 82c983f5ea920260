"""Products on the longest transforms, compared in full with GMP.

Plans of 2^23 and 2^24 points normally use 512-bit chunks (entries of 2^31
and 2^32 bits), whose GMP products take minutes on one host core.  A tuning
rule (rf_tune_set) steers 1x1 products of dense random 2^28- and 2^29-bit
entries onto 64-bit-chunk plans of exactly 2^23 and 2^24 points instead, so
that every column, middle and row pass of those lengths, and the twiddle
tables of order 2^24, are exercised on operands GMP multiplies in seconds;
the whole product is compared limb for limb (mpz_cmp).  The reference's GPU
path stops at 2^26 points of 116-bit chunks
(src/include/cufft62_thresholds.h:3); its CPU path has no limit
(src/lib/fft62.c:490-615).
"""
import ctypes

import pytest

import rforest_amd as rf
from rforest_amd import tune, workloads
from rforest_amd.gmp import MPZ, MpzArray, mpz_fn

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield
    tune.reset()


@pytest.mark.parametrize("lgN", [23, 24])
def test_longest_plans_equal_gmp_in_full(lgN):
    bits = 1 << (lgN + 5)  # 2^(lgN-1) chunks of 64 bits per entry
    print(f"case: 1x1 product of {bits}-bit entries on a 2^{lgN}-point plan", flush=True)
    tune.set_rule(1, lgN, 1, 0, 0.0)
    assert rf.plan(bits, bits, 1, 1)[:2] == (1, lgN)
    rng = workloads.GmpRandom(2000 + lgN)
    a, b = rng.urandomb(bits) | (1 << (bits - 1)), -(rng.urandomb(bits) | (1 << (bits - 1)))
    rng.close()
    mA, mB, mC, want = MpzArray([a]), MpzArray([b]), MpzArray.zeros(1), MpzArray.zeros(1)
    del a, b
    P = ctypes.POINTER(MPZ)
    rf.check(rf.lib().rf_mpz_rmatrix_mult(mC.ptr(), mA.ptr(), 1, mB.ptr(), 1))
    mpz_fn("mul", [P, P, P])(want.ref(0), mA.ref(0), mB.ref(0))
    assert mpz_fn("cmp", [P, P], ctypes.c_int)(mC.ref(0), want.ref(0)) == 0
    for m in (mA, mB, mC, want):
        m.close()
