"""Remainders X mod D (rf_mod_*, the reference's mpzfft_mod interface,
src/lib/mpzfft.c:59-100) against Python's x % d, which is in [0, d) for
d > 0 like the reference's output."""
import ctypes
import random

import pytest

import rforest_amd as rf
from rforest_amd.mod import Modulus, mod_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield


@pytest.mark.parametrize("dbits,xbits", [(1, 64), (64, 64), (64, 200), (65, 130), (1000, 2000), (4096, 4096),
                                         (50000, 100001), (1 << 20, 1 << 21), (777, 1 << 18), (3 << 20, 5 << 20)])
def test_remainders(dbits, xbits):
    rng = random.Random(dbits * 3 + xbits)
    d = rng.getrandbits(dbits) | (1 << (dbits - 1))
    M = Modulus(d, xbits)
    xs = [0, 1, -1, d, -d, d - 1, d + 1, 2 * d, 3 * d - 1, -(3 * d) + 1,
          (1 << xbits) - 1, -((1 << xbits) - 1)] + [rng.getrandbits(xbits) * rng.choice([-1, 1]) for _ in range(4)]
    for x in xs:
        if abs(x).bit_length() <= max(xbits, dbits):
            assert M.mod(x) == x % d, (dbits, xbits, x.bit_length())


def test_quotient_estimate_edges():
    """Dividends right at the quotient estimate's error bound: D a power of
    two, D = 2^k - 1 (all-ones limbs), and X = q D + r with r near 0 and D."""
    rng = random.Random(3)
    for d in ((1 << 640), (1 << 640) - 1, (1 << 64) + 1, 3):
        M = Modulus(d, 2048)
        for _ in range(20):
            q = rng.getrandbits(2048 - d.bit_length())
            for r in (0, 1, d - 1, d // 2):
                x = q * d + r
                if x.bit_length() <= 2048:
                    assert M.mod(x) == r
                    assert M.mod(-x) == (-x) % d


def test_batch_matches_one_by_one():
    """A remainder-tree level: many moduli of uneven sizes in one call."""
    rng = random.Random(5)
    pairs = []
    for i in range(40):
        db = rng.choice([64, 300, 5000, 40000, 200000])
        d = rng.getrandbits(db) | (1 << (db - 1))
        xb = rng.choice([db, 2 * db, 2 * db + 5])
        M = Modulus(d, xb)
        x = rng.getrandbits(xb) * rng.choice([-1, 1])
        pairs.append((M, x))
    got = mod_batch(pairs)
    assert got == [x % M.d for (M, x) in pairs]
    assert [M.mod(x) for (M, x) in pairs[:6]] == got[:6]


def test_errors():
    with pytest.raises(rf.RfError):
        Modulus(0, 100)
    with pytest.raises(rf.RfError):
        Modulus(-5, 100)
    M = Modulus(12345, 128)
    with pytest.raises(rf.RfError):
        M.mod(1 << 300)


def test_large_batch_of_small_remainders():
    """Thousands of small remainders in one call: the pipelines' task images
    outgrow the staging buffer mid-call (a regression test for reading the
    buffer through a pointer taken before it was replaced)."""
    rng = random.Random(8)
    pairs = []
    mods = {}
    for i in range(3000):
        db = rng.choice([64, 128, 500])
        if db not in mods:
            mods[db] = [Modulus(rng.getrandbits(db) | (1 << (db - 1)), 2 * db) for _ in range(4)]
        M = rng.choice(mods[db])
        pairs.append((M, rng.getrandbits(2 * db) * rng.choice([-1, 1])))
    assert mod_batch(pairs) == [x % M.d for (M, x) in pairs]


def test_matrix_entries_modulo_one_modulus():
    """A 3x3 matrix of 2-Mbit signed entries reduced modulo one 1-Mbit m in
    one call: every entry of the batch names the same prepared modulus."""
    rng = random.Random(13)
    m = rng.getrandbits(1 << 20) | (1 << ((1 << 20) - 1))
    M = Modulus(m, 1 << 21)
    X = [rng.getrandbits(1 << 21) * rng.choice([-1, 1]) for _ in range(9)]
    assert mod_batch([(M, x) for x in X]) == [x % m for x in X]


def test_remainder_with_long_transforms():
    """A 2^26-bit modulus and 2^27-bit dividends: both products of the
    remainder need more than 2^17 points."""
    from rforest_amd.gmp import mod_values
    rng = random.Random(17)
    d = rng.getrandbits(1 << 26) | (1 << ((1 << 26) - 1))
    M = Modulus(d, 1 << 27)
    xs = [rng.getrandbits(1 << 27) * s for s in (1, -1)]
    assert [M.mod(x) for x in xs] == mod_values(xs, [d, d])


def test_cyclic_q_d_product():
    """For a 1-Mbit modulus the Q D product is a cyclic convolution modulo
    2^(64 W) - 1 (W >= limbs of D + 1), folded with X in the fix-up; random
    and boundary dividends, one by one and as one batch, against x % d.  A
    64-bit modulus is listed too, whichever form the cost model picks."""
    rng = random.Random(77)
    for dbits, xbits, cyclic in [(1 << 20, 1 << 21, True), (64, 200, None)]:
        d = rng.getrandbits(dbits) | (1 << (dbits - 1))
        M = Modulus(d, xbits)
        W = M.wrap_limbs()
        if cyclic:
            assert W >= M.limbs() + 1
        xs = [rng.getrandbits(xbits) * rng.choice([-1, 1]) for _ in range(3)] + \
             [(1 << xbits) - 1, -((1 << xbits) - 1), 3 * d - 1, 2 * d, d - 1, 0, -1]
        want = [x % d for x in xs]
        assert [M.mod(x) for x in xs] == want
        assert mod_batch([(M, x) for x in xs]) == want
        M.close()


def test_batch_creation():
    """rf_mod_create_batch: the same remainders as moduli created one by one
    (mixed sizes and dividend bounds), and a bad modulus in the batch is an
    error that leaves nothing created."""
    rng = random.Random(91)
    specs = [(64, 200), (1000, 2000), (1 << 20, 1 << 21), (777, 1 << 18), (4096, 4096)]
    ds = [rng.getrandbits(b) | (1 << (b - 1)) for b, _ in specs]
    mods = Modulus.batch(ds, [x for _, x in specs])
    for M, d, (_, xb) in zip(mods, ds, specs):
        xs = [rng.getrandbits(xb) * rng.choice([-1, 1]) for _ in range(3)] + [d - 1, 2 * d, -3 * d + 1]
        xs = [x for x in xs if abs(x).bit_length() <= xb]
        want = [x % d for x in xs]
        assert [M.mod(x) for x in xs] == want
        one = Modulus(d, xb)
        assert [one.mod(x) for x in xs] == want and one.limbs() == M.limbs() and one.wrap_limbs() == M.wrap_limbs()
        one.close()
        M.close()
    with pytest.raises(rf.RfError):
        Modulus.batch([5, 0, 7], 64)


def _recip(d, max_bits):
    import numpy as np
    L = rf.lib()
    L.rf_stage_mod_reciprocal.restype = ctypes.c_int
    L.rf_stage_mod_reciprocal.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p,
                                          ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    dl = np.frombuffer(d.to_bytes(8 * ((d.bit_length() + 63) // 64), "little"), dtype=np.uint64).copy()
    m = max((max_bits + 63) // 64, len(dl))
    out = np.zeros(m - len(dl) + 2, dtype=np.uint64)
    n = ctypes.c_uint64()
    rf.check(L.rf_stage_mod_reciprocal(dl.ctypes.data, len(dl), max_bits, out.ctypes.data, len(out), ctypes.byref(n)))
    assert n.value == len(out)
    return int.from_bytes(out.tobytes(), "little"), m


@pytest.mark.parametrize("dbits,xbits", [(64, 64), (100, 300), (4096, 8192), (4096 + 1, 8192), (64 * 64, 64 * 64 * 2),
                                         (64 * 65 - 3, 64 * 130), (200000, 400000), (200000, 1 << 20),
                                         (1 << 20, 1 << 21), ((1 << 20) + 7, 1 << 21)])
def test_device_reciprocal_within_two_of_the_floor(dbits, xbits):
    """The reciprocal a modulus keeps, computed by the device's Newton
    recursion, lies in [floor(B^(m+1)/D) - 2, floor(B^(m+1)/D)] (the
    remainder step's quotient estimate needs exactly that), for moduli with
    the top bit set or clear, powers of two, all-ones, and sizes below, at
    and above the recursion's 64-limb host base."""
    rng = random.Random(dbits * 7 + xbits)
    cases = [rng.getrandbits(dbits) | (1 << (dbits - 1)), 1 << (dbits - 1), (1 << dbits) - 1,
             (1 << (dbits - 1)) + 1]
    if dbits > 64:
        cases.append(rng.getrandbits(dbits - 5) | (1 << (dbits - 6)))  # top limb small
    for d in cases:
        I, m = _recip(d, xbits)
        exact = (1 << (64 * (m + 1))) // d
        assert exact - 2 <= I <= exact, (dbits, xbits, exact - I)
