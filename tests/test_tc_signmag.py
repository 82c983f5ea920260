"""CPU checks of the host epilogue every host entry point ends in: results of
rn limbs of two's complement to sign and magnitude, converted in pieces of
8192 limbs over the library's copy threads (rf_stage_tc_to_signmag), against
Python integers.  Host code only: no GPU work, no hw_mpz_setup."""
import ctypes
import random

import numpy as np
import pytest

import rforest_amd as rf

_u64p = ctypes.POINTER(ctypes.c_uint64)


def _lib():
    L = rf.lib()
    f = L.rf_stage_tc_to_signmag
    f.restype = ctypes.c_int
    f.argtypes = [_u64p, ctypes.c_uint64, _u64p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_long)]
    return L


def _convert(x, rn, cap=None):
    """(rc, magnitude limbs as written, signed length) for x held in rn limbs."""
    tc = np.frombuffer((x % (1 << (64 * rn))).to_bytes(8 * rn, "little"), dtype=np.uint64).copy()
    cap = rn if cap is None else cap
    dst = np.zeros(max(cap, 1) + 4, dtype=np.uint64)  # 4 guard limbs past cap
    ns = ctypes.c_long(12345)
    rc = _lib().rf_stage_tc_to_signmag(tc.ctypes.data_as(_u64p), rn, dst.ctypes.data_as(_u64p), cap,
                                       ctypes.byref(ns))
    return rc, dst, ns.value


def _check(x, rn):
    rc, dst, ns = _convert(x, rn)
    assert rc == rf.RF_OK
    n = abs(ns)
    assert n == (abs(x).bit_length() + 63) // 64  # no high zero limbs
    assert (ns < 0) == (x < 0)
    mag = int.from_bytes(dst[:n].tobytes(), "little")
    assert (-mag if ns < 0 else mag) == x
    assert not dst[n:].any()  # nothing written past the magnitude


@pytest.mark.parametrize("k", [0, 1, 8191, 8192, 8193, 16384])
def test_powers_of_two_both_signs_around_piece_edges(k):
    # -2^(64 k): limbs below k zero, all ones from k up; its magnitude is one
    # limb longer than the zero run and shorter than the two's complement
    rn = k + 3
    _check(-(1 << (64 * k)), rn)
    _check(1 << (64 * k), rn)
    _check(-(1 << (64 * k + 63)), rn)


def test_lowest_nonzero_limb_in_a_later_piece():
    rng = random.Random(3)
    odd = rng.getrandbits(3000) | 1
    for shift in (64 * 9000 + 5, 64 * 8192, 64 * 8191 + 63, 64 * 20000):
        x = odd << shift
        rn = (x.bit_length() + 64) // 64 + 2
        _check(-x, rn)
        _check(x, rn)


def test_extremes_zero_and_small():
    for rn in (1, 2, 3, 8192, 8193):
        top = 1 << (64 * rn - 1)
        _check(-top, rn)          # the most negative value: magnitude 2^(64 rn - 1)
        _check(top - 1, rn)       # the most positive
        _check(-1, rn)
        _check(1, rn)
        _check(0, rn)
        _check(-(1 << 64) + 1 if rn > 1 else -5, rn)


def test_random_values_of_many_pieces():
    rng = random.Random(7)
    for _ in range(40):
        bits = rng.choice([1, 63, 64, 65, 5000, 8192 * 64 - 1, 8192 * 64 + 1, 300000, 1500000])
        x = rng.getrandbits(bits) * rng.choice([-1, 1])
        if rng.random() < 0.3:
            x <<= 64 * rng.randrange(0, 10000)  # long runs of zero low limbs
        rn = (abs(x).bit_length() + 64) // 64 + rng.randrange(0, 3)
        _check(x, rn)


def test_capacity_is_checked_before_any_write():
    x = -((1 << (64 * 9000)) * 12345)
    rn = 9010
    rc, dst, ns = _convert(x, rn, cap=9000)  # the magnitude needs 9001 limbs
    assert rc == rf.RF_ERR_CAPACITY and not dst.any() and ns == 12345
    rc, dst, ns = _convert(x, rn, cap=9001)
    assert rc == rf.RF_OK and ns == -9001
