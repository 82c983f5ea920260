"""Product trees kept in device memory (rf_mpz_tree_product,
rf_tree_product_device): the root equals the left-to-right product of the
leaves, whatever the count, sizes and signs."""
import random
from functools import reduce

import pytest

import rforest_amd as rf
from helpers import naive

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield


def _matmul(X, Y, d):
    return naive(X, d, Y, d)


@pytest.mark.parametrize("n,d,bits", [(1, 2, 100), (2, 2, 3000), (3, 2, 500), (5, 3, 2000), (8, 2, 1024),
                                      (13, 2, 700), (7, 1, 9000), (6, 4, 300), (33, 2, 256)])
def test_tree_matches_leftfold(n, d, bits):
    rng = random.Random(n * 1000 + d * 10 + bits)
    mats = []
    for m in range(n):
        b = rng.randrange(1, 2 * bits)
        M = [rng.getrandbits(b) * rng.choice([-1, 1]) for _ in range(d * d)]
        if m % 4 == 1:
            M[0] = 0
        mats.append(M)
    want = reduce(lambda X, Y: _matmul(X, Y, d), mats)
    assert rf.tree_product(mats, d) == want


def test_tree_with_zero_matrix_and_cancellation():
    d = 2
    I = [1, 0, 0, 1]
    Z = [0, 0, 0, 0]
    m = (1 << 5000) - 1
    assert rf.tree_product([I, [m, m, -m, m], Z, I], d) == Z
    # A A^-1-like cancellation: [[1,x],[0,1]] [[1,-x],[0,1]] = I
    x = (1 << 9000) + 12345
    assert rf.tree_product([[1, x, 0, 1], [1, -x, 0, 1]] * 3, d) == I


def test_tree_level_of_headline_sized_leaves():
    from rforest_amd import workloads
    prods = workloads.headline(4, seed=42)
    mats = [p[2] for p in prods] + [p[3] for p in prods[:2]]
    want = reduce(lambda X, Y: _matmul(X, Y, 2), mats)
    assert rf.tree_product(mats, 2) == want


@pytest.mark.parametrize("case", ["low_zero_run", "low_zero_run_later_tile", "power_of_two_negated",
                                  "small_negative_in_long_buffer", "many_tiles_both_signs"])
def test_tree_level_conversion_edges(case):
    # level outputs that stress the in-place conversion to sign and
    # magnitude: the lowest nonzero limb several sweeps of 64 limbs up and in
    # a later tile of 2048, a negative magnitude one limb long, a
    # small negative value under hundreds of all-ones limbs, and entries of
    # several tiles of either sign
    rng = random.Random(sum(map(ord, case)))
    if case == "low_zero_run":
        d, mats = 1, [[-(1 << (64 * 300))], [1 << (64 * 100)], [3]]
    elif case == "low_zero_run_later_tile":
        d, mats = 1, [[-(1 << (64 * 3000 + 17))], [rng.getrandbits(700) | 1], [-(rng.getrandbits(64) | 1)]]
    elif case == "power_of_two_negated":
        d, mats = 1, [[-(1 << (64 * 2500))], [1], [1 << 64], [-1]]
    elif case == "small_negative_in_long_buffer":
        d = 2
        x = (1 << 40000) + 7
        mats = [[1, x, 0, 1], [1, -x - 5, 0, 1], [2, 3, 5, 7], [1, -x, 0, 1], [1, x - 1, 0, 1]]
    else:
        d = 2
        mats = [[rng.getrandbits(rng.randrange(100000, 200000)) * rng.choice([-1, 1]) for _ in range(4)]
                for _ in range(3)]
    want = reduce(lambda X, Y: _matmul(X, Y, d), mats)
    assert rf.tree_product(mats, d) == want


def test_tree_product_device_entry_point():
    # rf_tree_product_device: leaves already in device memory as sign and
    # magnitude descriptors, root magnitudes written to a device array
    import ctypes
    import numpy as np
    from rforest_amd.device import DeviceBuffer

    class DevEntry(ctypes.Structure):
        _fields_ = [("limbs", ctypes.c_void_p), ("nlimbs", ctypes.c_uint64), ("sign", ctypes.c_int32),
                    ("reserved", ctypes.c_int32)]

    rng = random.Random(77)
    n, d = 6, 2
    mats = [[rng.getrandbits(rng.randrange(1, 3000)) * rng.choice([-1, 1]) for _ in range(d * d)] for _ in range(n)]
    mats[2][1] = 0
    want = reduce(lambda X, Y: _matmul(X, Y, d), mats)
    flat = [x for M in mats for x in M]
    nl = [(abs(x).bit_length() + 63) // 64 for x in flat]
    off = np.cumsum([0] + nl)
    limbs = np.zeros(max(int(off[-1]), 1), dtype=np.uint64)
    for i, x in enumerate(flat):
        if nl[i]:
            limbs[off[i]:off[i + 1]] = np.frombuffer(abs(x).to_bytes(8 * nl[i], "little"), dtype=np.uint64)
    src = DeviceBuffer(8 * limbs.size)
    src.upload(limbs)
    ents = (DevEntry * len(flat))()
    for i, x in enumerate(flat):
        ents[i] = DevEntry(src.ptr + 8 * int(off[i]), nl[i], (x > 0) - (x < 0), 0)
    cap = max((abs(w).bit_length() + 63) // 64 for w in want) + 1
    out = DeviceBuffer(8 * cap * d * d)
    sign = (ctypes.c_int * (d * d))()
    nlimbs = (ctypes.c_uint64 * (d * d))()
    L = rf.lib()
    L.rf_tree_product_device.restype = ctypes.c_int
    L.rf_tree_product_device.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                                         ctypes.c_void_p, ctypes.c_void_p]
    rf.check(L.rf_tree_product_device(ents, n, d, ctypes.c_void_p(out.ptr), cap, sign, nlimbs))
    res = out.download(np.uint64, cap * d * d)
    got = []
    for e in range(d * d):
        mag = int.from_bytes(res[e * cap:e * cap + nlimbs[e]].tobytes(), "little")
        assert nlimbs[e] == (mag.bit_length() + 63) // 64  # no high zero limbs
        got.append(sign[e] * mag)
    assert got == want
    short = max(nlimbs) - 1
    assert L.rf_tree_product_device(ents, n, d, ctypes.c_void_p(out.ptr), short, sign, nlimbs) == rf.RF_ERR_CAPACITY
    src.free()
    out.free()


def test_trees_above_and_below_the_kept_blocks():
    # the library keeps a tree's device blocks for the next tree up to 512 MiB
    # in all; a tree of two leaves of 1.25 * 2^30 bits (320 MiB of leaves, 320
    # MiB for the product) frees them at its end.  Large, small, large again,
    # each exact; the values are of the form 2^a - c so that the expected
    # products take linear time to form
    a, b = (1 << 30) + (1 << 28), (1 << 30) + (1 << 28) - 12345
    x, y = (1 << a) - 1, -((1 << b) - 3)
    big_want = [-((1 << (a + b)) - 3 * (1 << a) - (1 << b) + 3)]  # (2^a - 1)(-(2^b - 3)) expanded
    small = [[3, -5, 7, 11], [2, 0, -1, 4], [-(1 << 700), 1, 1, 1 << 900]]
    small_want = reduce(lambda X, Y: _matmul(X, Y, 2), small)
    assert rf.tree_product([[x], [y]], 1) == big_want
    assert rf.tree_product(small, 2) == small_want
    assert rf.tree_product([[x], [y]], 1) == big_want
