"""The fp64 modular primitives of rforest_amd/csrc/rf_field.h, pinned on the CPU.

Every exactness claim of the GPU arithmetic rests on the bounds stated in
rf_field.h.  Here the same source is compiled for the host (g++,
-ffp-contract=off: every FMA explicit, as in the library's build) and each
function is checked against exact 128-bit integer arithmetic for all 32
primes, at its stated bound (|a| = 2^53 - 1 for rf_mulmod, |a|, |b| = 2^49 for
rf_mulmod_var (whose stated output bound this test corrected: it depends
on |a b|, not |a| alone), |x| = 2^53 - 1 for rf_reduce, |x| < 2^90 for rf_reduce_wide,
|x| = p - 1 for rf_canon1, residues at the lazy bound 12.2 p + 160 of the
transforms), with random inputs and the edges in between.  Each checker is
also run on a planted mistake (the product without its low half, a quotient
rounded down, the canonicalisation offset 2^-45) and must reject it.  The
lazy-range arithmetic of the transforms (rf_plan.h at RF_PMAX) is restated
in exact rationals.  Reference contracts: the lazy ranges of
src/include/mod62.h:23-75 and the Shoup product of src/include/mod62.h:245-262.
"""
import ctypes
import os
import subprocess
from fractions import Fraction

import pytest

import rforest_amd as rf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = 40000  # random inputs per prime and checker


@pytest.fixture(scope="module")
def fh(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("field") / "libfieldhost.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-o", out,
                    os.path.join(ROOT, "tests", "field_host.cpp")], check=True)
    L = ctypes.CDLL(out)
    d, u64, i64, lng, i32 = ctypes.c_double, ctypes.c_uint64, ctypes.c_int64, ctypes.c_long, ctypes.c_int
    for name in ("fh_mulmod", "fh_mulmod_var", "fh_reduce", "fh_reduce_wide", "fh_canon1"):
        getattr(L, name).restype = d
    L.fh_mulmod.argtypes = [d, d, d, d]
    L.fh_mulmod_var.argtypes = [d, d, d, d]
    L.fh_reduce.argtypes = [d, d, d]
    L.fh_reduce_wide.argtypes = [d, d, d]
    L.fh_canon1.argtypes = [d, d, d]
    pd = ctypes.POINTER(d)
    for name in ("fh_check_mulmod", "fh_check_mulmod_var", "fh_check_reduce"):
        getattr(L, name).restype = lng
        getattr(L, name).argtypes = [i32, u64, i64, lng, u64, pd]
    L.fh_check_reduce_wide.restype = lng
    L.fh_check_reduce_wide.argtypes = [u64, i32, lng, u64, pd]
    L.fh_check_canon1.restype = lng
    L.fh_check_canon1.argtypes = [i32, u64, i64, lng, u64]
    return L


PRIMES = rf.primes(32)
LAZY = [int(Fraction(122, 10) * p) + 161 for p in PRIMES]  # just above 12.2 p + 160


def _run(fn, *args):
    worst = ctypes.c_double(0)
    bad = fn(*args, ctypes.byref(worst))
    return bad, worst.value


@pytest.mark.parametrize("q", range(32))
def test_mulmod_at_its_bound(fh, q):
    p = PRIMES[q]
    for amax in ((1 << 53) - 1, LAZY[q], (1 << 51), p):
        bad, worst = _run(fh.fh_check_mulmod, 0, p, amax, N, 1000 + q)
        assert bad == 0, f"rf_mulmod wrong for p={p}, |a|<={amax}"
        assert worst <= 1.0
    # exact edges, one by one
    for a in ((1 << 53) - 1, -((1 << 53) - 1), LAZY[q], -LAZY[q], p - 1, 1 - p, 0, 1, -1):
        for w in (0, 1, p - 1, p - 2, (p - 1) // 2):
            r = fh.fh_mulmod(float(a), float(w), float(w) / float(p), float(p))
            assert r == int(r)
            assert (int(r) - a * w) % p == 0
            assert abs(Fraction(int(r))) <= p * (Fraction(1, 2) + Fraction(abs(a), 1 << 52))


@pytest.mark.parametrize("q", range(32))
def test_mulmod_var_at_its_bound(fh, q):
    p = PRIMES[q]
    bad, worst = _run(fh.fh_check_mulmod_var, 0, p, 1 << 49, N, 2000 + q)
    assert bad == 0 and worst <= 1.0
    # the pointwise products' operands are reduced (|x| <= p/2 + 8): their
    # terms stay below 0.53 p (the kernels budget 0.6 p per term)
    h = p // 2 + 8
    bad, worst = _run(fh.fh_check_mulmod_var, 0, p, h, N, 2100 + q)
    assert bad == 0
    assert Fraction(p, 2) + Fraction(151 * h * h, 100 << 52) < Fraction(53, 100) * p


@pytest.mark.parametrize("q", range(32))
def test_reduce_at_its_bound(fh, q):
    p = PRIMES[q]
    bad, worst = _run(fh.fh_check_reduce, 0, p, (1 << 53) - 1, N, 3000 + q)
    assert bad == 0 and worst <= 1.0
    pinv = 1.0 / float(p)
    for x in ((1 << 53) - 1, -((1 << 53) - 1), LAZY[q], -LAZY[q], p // 2, -(p // 2), p // 2 + 1, 3 * p // 2):
        r = fh.fh_reduce(float(x), float(p), pinv)
        assert r == int(r) and (int(r) - x) % p == 0 and abs(int(r)) <= p // 2 + 8


@pytest.mark.parametrize("q", range(32))
def test_reduce_wide_at_its_bound(fh, q):
    p = PRIMES[q]
    bad, worst = _run(fh.fh_check_reduce_wide, p, 90, N, 4000 + q)
    assert bad == 0 and worst <= 1.0
    pinv = 1.0 / float(p)
    # the pack's three accumulators at their largest: 16 terms of (2^32 - 1)
    # times a 16-bit piece, scaled by 1, 2^16, 2^32
    for x in (16 * ((1 << 32) - 1) * ((1 << 16) - 1) << 16, 16 * ((1 << 32) - 1) * ((1 << 16) - 1) << 32,
              ((1 << 53) - 1) << 37, -(((1 << 53) - 1) << 37)):
        assert abs(x) < 1 << 90 and float(x) == x
        r = fh.fh_reduce_wide(float(x), float(p), pinv)
        assert r == int(r) and (int(r) - x) % p == 0 and abs(int(r)) <= p * (Fraction(1, 2) + Fraction(1, 512))


@pytest.mark.parametrize("q", range(32))
def test_canon1_at_its_bound(fh, q):
    p = PRIMES[q]
    assert fh.fh_check_canon1(0, p, p - 1, N, 5000 + q) == 0
    pinv = 1.0 / float(p)
    for x in (p - 1, -(p - 1), 0, 1, -1, p // 2, -(p // 2)):
        r = fh.fh_canon1(float(x), float(p), pinv)
        assert r == x % p


def test_planted_mistakes_are_rejected(fh):
    """The checkers have teeth: each planted mistake fails on every prime."""
    for q in (0, 11, 31):
        p = PRIMES[q]
        assert _run(fh.fh_check_mulmod, 1, p, (1 << 53) - 1, N, 7)[0] > 0
        assert _run(fh.fh_check_mulmod_var, 1, p, 1 << 49, N, 7)[0] > 0
        assert _run(fh.fh_check_reduce, 1, p, (1 << 53) - 1, N, 7)[0] > 0
        assert fh.fh_check_canon1(1, p, p - 1, N, 7) > 0


def test_bounds_are_tight_enough_to_matter(fh):
    """Outside their domains the functions do fail: rf_reduce_wide on
    integers near 2^110, rf_canon1 on |x| up to 2^53 (it is right up to about
    18 p, its floor's rounding margin; the kernels call it on |x| < p); so the
    domains the tests check are not vacuous."""
    p = PRIMES[0]
    assert _run(fh.fh_check_reduce_wide, p, 110, N, 9)[0] > 0
    pinv = 1.0 / float(p)
    wrong = [m for m in range(1, 64) if fh.fh_canon1(float(m * p - 1), float(p), pinv) != p - 1]
    assert wrong and min(wrong) > 16


def test_lazy_range_of_the_transforms():
    """rf_plan.h (RF_PMAX): primes below 2^51 / 9; from |x| <= p/2 + 8, at
    most four doubling stages (twiddle 1, not multiplied) at the start of a
    window and multiplied stages adding p (1/2 + B 2^-52) keep every value of
    an eight-stage window below 16 (p/2 + 8)(1 + 1/18)^4 + 2.2 p
    < 12.2 p + 160 < 2^51.5 < 2^53 (rf_mulmod's exact range).  Restated in
    exact rationals for every prime and every split of the window; the
    worst order (all doublings first) is the one the kernels can reach."""
    for p in PRIMES:
        assert p < (1 << 51) // 9
        assert Fraction(p, 1 << 52) < Fraction(1, 18)
        start = Fraction(p, 2) + 8
        worst = Fraction(0)
        for dbl in range(0, 5):
            B = start
            for _ in range(dbl):
                B = 2 * B
            for _ in range(8 - dbl):
                B = B + p * (Fraction(1, 2) + B / (1 << 52))
            worst = max(worst, B)
        stated = 16 * start * Fraction(19, 18) ** 4 + Fraction(22, 10) * p
        assert worst <= stated
        assert stated < Fraction(122, 10) * p + 160
        assert (Fraction(122, 10) * p + 160) ** 2 < Fraction(2) ** 103  # 12.2 p + 160 < 2^51.5
        # a product at the lazy bound stays inside rf_mulmod's exact domain
        assert Fraction(122, 10) * p + 160 < (1 << 53)
