"""The drop-in on products beyond one device plan: entries cut into pieces
whose products run on the device and are recombined on the host.

The reference's CPU mpz_rmatrix_mult_fft (src/lib/hwmpz.c:65-130) takes
entries of any size, so the drop-in must too.  The recombination is checked
at small sizes against GMP schoolbook by forcing small pieces
(rf_stage_dropin_piece_bits), then once above the device plans' 2^32-bit
limit on structured entries whose exact product GMP forms cheaply.
"""
import ctypes
import random

import pytest

import rforest_amd as rf
from helpers import naive
from rforest_amd.gmp import MPZ, MpzArray, mpz_fn

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield
    rf.check(rf.lib().rf_stage_dropin_piece_bits(0))


def _dropin(A, r, B, d, entry="rf"):
    mA, mB, mC = MpzArray(A), MpzArray(B), MpzArray.zeros(r * d)
    try:
        if entry == "rf":
            rf.check(rf.lib().rf_mpz_rmatrix_mult(mC.ptr(), mA.ptr(), r, mB.ptr(), d))
        else:
            w = MpzArray.zeros(1)
            rf.lib().mpz_rmatrix_mult_fft(mC.ptr(), mA.ptr(), r, mB.ptr(), d, w.ptr())
            w.close()
        return mC.values()
    finally:
        mA.close()
        mB.close()
        mC.close()


def _entry(rng, bits):
    if rng.random() < 0.08:
        return 0
    b = rng.randint(1, bits)
    x = rng.getrandbits(b) | (1 << (b - 1))
    return -x if rng.random() < 0.5 else x


@pytest.mark.parametrize("r,d,bits,piece", [(2, 2, 20000, 4096), (3, 3, 9000, 1024), (1, 4, 30000, 4096),
                                            (4, 1, 5000, 640), (2, 2, 300000, 65536), (2, 3, 12345, 64)])
def test_pieces_match_schoolbook(r, d, bits, piece):
    print(f"case: pieces r={r} d={d} bits={bits} piece={piece}", flush=True)
    rng = random.Random(r * 1000 + d * 100 + bits)
    rf.check(rf.lib().rf_stage_dropin_piece_bits(piece))
    for entry in ("rf", "dropin"):
        for _ in range(3):
            A = [_entry(rng, bits) for _ in range(r * d)]
            B = [_entry(rng, bits) for _ in range(d * d)]
            A[0] = (1 << (bits - 1)) | 1  # at least one entry of full length
            assert _dropin(A, r, B, d, entry) == naive(A, r, B, d)


def test_pieces_edges():
    """Entries at exact piece multiples, all-ones pieces (carries between
    diagonals), a negative product, zeros."""
    print("case: pieces edges", flush=True)
    P = 4096
    rf.check(rf.lib().rf_stage_dropin_piece_bits(P))
    for A, B in [([(1 << (3 * P)) - 1, 0, -((1 << (2 * P)) - 1), 1 << P], [(1 << (3 * P)) - 1, -1, 0, (1 << (4 * P)) - 1]),
                 ([-(1 << (5 * P)), 1, 1, 1], [1 << (2 * P), 1 << P, -1, (1 << (5 * P)) + 1]),
                 ([0, 0, 0, 0], [(1 << (3 * P)) + 5, 1, 2, 3])]:
        assert _dropin(A, 2, B, 2) == naive(A, 2, B, 2)


def _z():
    z = MPZ()
    mpz_fn("init", [ctypes.POINTER(MPZ)])(ctypes.byref(z))
    return z


def test_above_the_device_limit():
    """1x1 product of entries of 2^32 + 101 and 2^32 + 37 bits (beyond the
    device plans: 2^24 points of 512-bit chunks), with dense low and middle
    parts straddling the 2^31-bit piece boundary:
    a = 2^La + m1 2^K1 + x, b = 2^Lb - m2 2^K2 + y.  The exact product is a
    sum of nine shifted small products, formed by GMP and compared in full."""
    print("case: 1x1 above the 2^32-bit device limit", flush=True)
    P = ctypes.POINTER(MPZ)
    mul = mpz_fn("mul", [P, P, P])
    add = mpz_fn("add", [P, P, P])
    mul_2exp = mpz_fn("mul_2exp", [P, P, ctypes.c_ulong])
    setbit = mpz_fn("setbit", [P, ctypes.c_ulong])
    cmp = mpz_fn("cmp", [P, P], ctypes.c_int)
    sizeinbase = mpz_fn("sizeinbase", [P, ctypes.c_int], ctypes.c_size_t)
    rng = random.Random(51)
    La, Lb = (1 << 32) + 100, (1 << 32) + 36
    K1, K2 = (1 << 31) - 3000, (1 << 31) - 700
    small = MpzArray([rng.getrandbits(1 << 20), rng.getrandbits(4096), rng.getrandbits(1 << 20),
                      -rng.getrandbits(4096)])  # x, m1, y, -m2
    x, m1, y, nm2 = (small.ref(i) for i in range(4))
    # the operands: terms (value, shift) of a and b
    ta = [(None, La), (m1, K1), (x, 0)]
    tb = [(None, Lb), (nm2, K2), (y, 0)]
    a, b, t, want = _z(), _z(), _z(), _z()

    def build(dst, terms):
        for v, sh in terms:
            if v is None:
                setbit(ctypes.byref(dst), sh)
            else:
                mul_2exp(ctypes.byref(t), v, sh)
                add(ctypes.byref(dst), ctypes.byref(dst), ctypes.byref(t))

    build(a, ta)
    build(b, tb)
    assert sizeinbase(ctypes.byref(a), 2) == La + 1 and sizeinbase(ctypes.byref(b), 2) == Lb
    one = MpzArray([1])
    for va, sa in ta:
        for vb, sb in tb:
            pa = one.ref(0) if va is None else va
            pb = one.ref(0) if vb is None else vb
            mul(ctypes.byref(t), pa, pb)
            mul_2exp(ctypes.byref(t), ctypes.byref(t), sa + sb)
            add(ctypes.byref(want), ctypes.byref(want), ctypes.byref(t))
    mA, mB, mC = MpzArray.zeros(1), MpzArray.zeros(1), MpzArray.zeros(1)
    mpz_fn("swap", [P, P])(mA.ref(0), ctypes.byref(a))
    mpz_fn("swap", [P, P])(mB.ref(0), ctypes.byref(b))
    assert rf.lib().mpz_rmatrix_gpu_fft_suitable(mA.ptr(), 1, mB.ptr(), 1) == 0  # no single device plan
    rf.check(rf.lib().rf_mpz_rmatrix_mult(mC.ptr(), mA.ptr(), 1, mB.ptr(), 1))
    assert cmp(mC.ref(0), ctypes.byref(want)) == 0
    clear = mpz_fn("clear", [P])
    for z in (a, b, t, want):
        clear(ctypes.byref(z))
    for m in (mA, mB, mC, small, one):
        m.close()
