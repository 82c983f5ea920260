"""End-to-end parity of the GPU product with schoolbook GMP-equivalent
arithmetic (Python integers) and with the reference CPU path, on the
benchmark configurations and the edge cases the reference's path admits
(signs, zeros, ragged sizes, r != d, all-ones entries)."""
import random

import pytest

import oracle
import rforest_amd as rf
from helpers import naive
from rforest_amd import workloads
from rforest_amd.gmp import MpzArray

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield


def _rand_product(rng, bits, r, d, signed=True):
    s = (lambda: rng.choice([-1, 1])) if signed else (lambda: 1)
    return ([rng.getrandbits(bits) * s() for _ in range(r * d)],
            [rng.getrandbits(bits) * s() for _ in range(d * d)])


@pytest.mark.parametrize("bits,r,d", [(1, 1, 1), (64, 1, 1), (65, 2, 2), (4096, 2, 2), (50000, 2, 2),
                                      (3000, 3, 3), (100, 1, 4), (200000, 2, 2), (12345, 2, 5),
                                      (777, 4, 4), (30000, 5, 1), (100000, 1, 2)])
def test_product_vs_schoolbook(bits, r, d):
    rng = random.Random(bits + 100 * r + d)
    A, B = _rand_product(rng, bits, r, d)
    assert rf.mpz_rmatrix_mult_fft(A, r, B, d) == naive(A, r, B, d)


def test_headline_product_vs_reference_cpu_path():
    (r, d, A, B), = workloads.headline(1, seed=42)
    C = rf.mpz_rmatrix_mult_fft(A, r, B, d)
    assert C == naive(A, r, B, d)
    if oracle.ref_available():
        assert C == oracle.ref_rmatrix_mult(A, r, B, d)


def test_cpu_reference_case_seed_1001():
    r, d, A, B = workloads.cpu_reference_case(1001)
    C = rf.mpz_rmatrix_mult_fft(A, r, B, d)
    assert C == naive(A, r, B, d)
    if oracle.ref_available():
        assert C == oracle.ref_rmatrix_mult(A, r, B, d)


def test_plain_gpu_case_4x4_seed_1002():
    r, d, A, B = workloads.plain_gpu_case(1002)
    assert rf.mpz_rmatrix_mult_fft(A, r, B, d) == naive(A, r, B, d)


@pytest.mark.parametrize("bits,r,d", [(5 << 20, 2, 2), (1 << 24, 2, 2), (3 << 21, 3, 3), (1 << 26, 1, 1),
                                      (3 << 25, 2, 1), (1 << 27, 1, 1), (1 << 29, 1, 1)])
def test_long_transform_products(bits, r, d):
    """Entries whose plans need more than 2^14 points (column + row passes),
    against GMP schoolbook."""
    rng = random.Random(bits + r)
    t, lgN, k = rf.plan(bits, bits, r, d)
    assert lgN > 14
    A, B = _rand_product(rng, bits, r, d)
    assert rf.mpz_rmatrix_mult_fft(A, r, B, d) == naive(A, r, B, d)


def test_product_at_2_22_points_vs_gmp():
    """1x1 product of 2^30-bit entries: a plan of at least 2^22 points (three
    column passes each way, the twiddle tables of order 2^24)."""
    rng = random.Random(22)
    bits = 1 << 30
    assert rf.plan(bits, bits, 1, 1)[1] >= 22
    A, B = _rand_product(rng, bits, 1, 1)
    assert rf.mpz_rmatrix_mult_fft(A, 1, B, 1) == naive(A, 1, B, 1)


def _limb_mods(v):
    """x mod 2^64 - 1 and x mod 2^64 + 1 for x given by its uint64 limbs
    (digit sums in base 2^64; 32-bit halves keep numpy's sums exact)."""
    import numpy as np
    lo, hi = v & np.uint64(0xffffffff), v >> np.uint64(32)
    s = lambda a: int(a.sum(dtype=np.uint64))
    tot = s(lo) + (s(hi) << 32)
    alt = s(lo[0::2]) + (s(hi[0::2]) << 32) - s(lo[1::2]) - (s(hi[1::2]) << 32)
    return tot % ((1 << 64) - 1), alt % ((1 << 64) + 1)


def test_product_at_2_24_points_digit_sums():
    """2x1 product of 2^32-bit entries (2^24 points, the longest plan), on
    limb arrays: each result modulo 2^64 - 1 and 2^64 + 1 (so modulo
    2^128 - 1) against the inputs' residues, its low 4096 limbs exactly, and
    its sign and length."""
    import numpy as np
    bits = 1 << 32
    r, d = 2, 1
    assert rf.plan(bits, bits, r, d)[1] == 24
    gen = np.random.default_rng(24)
    n = bits // 64
    a0 = gen.integers(0, 1 << 64, n, dtype=np.uint64, endpoint=False)
    a1 = gen.integers(0, 1 << 64, n, dtype=np.uint64, endpoint=False)
    a1[-1] |= np.uint64(1 << 63)  # top bit set
    b0 = gen.integers(0, 1 << 64, n, dtype=np.uint64, endpoint=False)
    C = rf.rmatrix_mult_limb_arrays([(1, a0), (-1, a1)], r, [(1, b0)], d)
    mods = [(1 << 64) - 1, (1 << 64) + 1]
    lowb = 64 * 4096
    for (sa, av), (sc, cv) in zip([(1, a0), (-1, a1)], C):
        assert sc == sa and len(cv) in (2 * n - 1, 2 * n) and int(cv[-1]) != 0
        for ma, mb, mc, m in zip(_limb_mods(av), _limb_mods(b0), _limb_mods(cv), mods):
            assert mc == ma * mb % m
        low = lambda v: int.from_bytes(v[:4096].tobytes(), "little")
        assert low(cv) == low(av) * low(b0) % (1 << lowb)


def test_zero_and_ragged_entries():
    rng = random.Random(11)
    A = [0, rng.getrandbits(9000), -rng.getrandbits(10), 0]
    B = [rng.getrandbits(100), 0, 0, -rng.getrandbits(8000)]
    assert rf.mpz_rmatrix_mult_fft(A, 2, B, 2) == naive(A, 2, B, 2)
    assert rf.mpz_rmatrix_mult_fft([0] * 4, 2, B, 2) == [0] * 4
    assert rf.mpz_rmatrix_mult_fft(A, 2, [0] * 4, 2) == [0] * 4


def test_all_ones_entries_and_cancellation():
    L = 70000
    m = (1 << L) - 1
    A = [m, m, -m, m]
    B = [m, -m, m, m]
    assert rf.mpz_rmatrix_mult_fft(A, 2, B, 2) == naive(A, 2, B, 2)  # includes exact zeros
    A = [m, m, m, m]
    assert rf.mpz_rmatrix_mult_fft(A, 2, A, 2) == naive(A, 2, A, 2)


def test_powers_of_two():
    A = [1 << 50000, 1, -(1 << 63), 1 << 64]
    B = [1 << 1, -(1 << 70000), 1, 3]
    assert rf.mpz_rmatrix_mult_fft(A, 2, B, 2) == naive(A, 2, B, 2)


def test_uneven_operand_sizes():
    rng = random.Random(3)
    A = [rng.getrandbits(300000) for _ in range(4)]
    B = [rng.getrandbits(5000) * rng.choice([-1, 1]) for _ in range(4)]
    assert rf.mpz_rmatrix_mult_fft(A, 2, B, 2) == naive(A, 2, B, 2)


def test_dropin_mpz_abi():
    """mpz_rmatrix_mult_fft through its C prototype on GMP mpz_t arrays."""
    rng = random.Random(5)
    for (bits, r, d) in [(5000, 2, 2), (40000, 3, 2), (1, 1, 1)]:
        A, B = _rand_product(rng, bits, r, d)
        mA, mB, mC, w = MpzArray(A), MpzArray(B), MpzArray.zeros(r * d), MpzArray.zeros(1)
        # C entries pre-set to garbage of another size must be overwritten
        mC.set([12345678901234567890 ** 3] * (r * d))
        rf.lib().mpz_rmatrix_mult_fft(mC.ptr(), mA.ptr(), r, mB.ptr(), d, w.ptr())
        assert mC.values() == naive(A, r, B, d)
        assert rf.lib().mpz_rmatrix_gpu_fft_suitable(mA.ptr(), r, mB.ptr(), d) == 1
        mC2 = MpzArray.zeros(r * d)
        rf.lib().cu_mpz_rmatrix_mult_fft(mC2.ptr(), mA.ptr(), r, mB.ptr(), d, w.ptr())
        assert mC2.values() == naive(A, r, B, d)
        assert rf.lib().rf_mpz_rmatrix_mult(mC2.ptr(), mA.ptr(), r, mB.ptr(), d) == rf.RF_OK


def test_dropin_mpz_results_converted_in_pieces():
    """The mpz_t drop-in converts two's complement results to sign and
    magnitude in pieces of 8192 limbs: negative results whose lowest nonzero
    limb lies in a later piece (or starts one), a power of two whose
    magnitude is shorter than its two's complement, and mixed signs over
    results of several pieces, against exact products."""
    rng = random.Random(11)
    odd = rng.getrandbits(3000) | 1
    cases = [
        ([-odd << (64 * 9000 + 5)], 1, [3], 1),                  # zero limbs 0..8999, negative
        ([-(1 << (64 * 8192))], 1, [1], 1),                      # lowest nonzero limb = 8192, a piece start
        ([1 << (64 * 8192)], 1, [-(1 << 64)], 1),                # the same, one limb higher
        (_rand_product(rng, 600000, 2, 2)[0], 2, _rand_product(rng, 600000, 2, 2)[1], 2),  # ~18,750 limbs
        ([-(rng.getrandbits(700000) << 640), rng.getrandbits(5000)], 1,
         [rng.getrandbits(640000), -(1 << 100), 1, -rng.getrandbits(90000)], 2),
    ]
    w = MpzArray.zeros(1)
    for (A, r, B, d) in cases:
        mA, mB, mC = MpzArray(A), MpzArray(B), MpzArray.zeros(r * d)
        mC.set([-(12345678901234567890 ** 5)] * (r * d))
        rf.lib().mpz_rmatrix_mult_fft(mC.ptr(), mA.ptr(), r, mB.ptr(), d, w.ptr())
        assert mC.values() == naive(A, r, B, d)


def test_plain_array_path_refuses_sizes_without_a_plan():
    """rf_rmatrix_mult_limbs (plain arrays, one device plan per call) returns
    RF_ERR_UNSUPPORTED for a product of two 2^33-bit entries, which needs 2^25
    points at the widest chunks (the longest plan has 2^24); the mpz_t drop-in computes
    such products from pieces instead (test_gpu_pieces.py)."""
    import numpy as np
    big = np.zeros(1 << 27, dtype=np.uint64)  # 2^33 bits, sparse
    big[-1] = 1 << 63
    big[0] = 12345
    with pytest.raises(rf.RfError) as ei:
        rf.rmatrix_mult_limb_arrays([(1, big)], 1, [(-1, big)], 1)
    assert ei.value.code == rf.RF_ERR_UNSUPPORTED
    mA = MpzArray([1 << 40])
    assert rf.lib().mpz_rmatrix_gpu_fft_suitable(mA.ptr(), 1, mA.ptr(), 1) == 1
    mA.close()


def test_device_batch_mixed_level_matches_alone():
    """A product-tree level of uneven products in one call; each result equals
    the product computed alone and the schoolbook value."""
    prods = workloads.tree_level(48, seed=1003)
    prods += [workloads.plain_gpu_case(1002)]
    batch = rf.DeviceBatch(prods)
    batch.run()
    res = batch.results()
    for (r, d, A, B), C in zip(prods, res):
        assert C == naive(A, r, B, d)
    for (r, d, A, B), C in list(zip(prods, res))[:5]:
        assert rf.mpz_rmatrix_mult_fft(A, r, B, d) == C


def test_tree_level_all_1024_products_through_the_prepared_lanes():
    """The whole product-tree level workload (1024 products 2x2 . 2x2, entries
    log-uniform in [2^10, 2^17] bits, seed 1003) as the benchmark runs it:
    a prepared batch of three lanes replayed from its captured graph, twice;
    every product equals GMP schoolbook (mpz_mul / mpz_addmul)."""
    from rforest_amd.gmp import rmatrix_mult_schoolbook
    prods = workloads.tree_level(1024, seed=1003)
    want = [rmatrix_mult_schoolbook(A, r, B, d) for (r, d, A, B) in prods]
    batch = rf.DeviceBatch(prods)
    batch.prepare(lanes=3)
    for _ in range(2):
        batch.clear_outputs()
        batch.launch(graph=True)
        rf.device_sync()
        got = batch.results()
        bad = [i for i, (g, w) in enumerate(zip(got, want)) if g != w]
        assert not bad, f"products {bad[:8]} differ from GMP"
    batch.release()


def test_device_batch_repeat_is_deterministic():
    prods = workloads.headline(3, seed=42)
    batch = rf.DeviceBatch(prods)
    batch.run()
    first = batch.results()
    for _ in range(3):
        batch.run(sync=False)
    assert batch.results() == first
    assert first[0] == naive(prods[0][2], 2, prods[0][3], 2)


def test_prepared_batch_graph_replays_match_eager():
    """The prepared pipeline, replayed from a captured hipGraph several times
    with eager launches and synchronisations in between, rewrites every output
    with the eager result each time (outputs are clobbered between runs)."""
    import numpy as np
    prods = workloads.tree_level(16, seed=7) + workloads.headline(2, seed=42)
    batch = rf.DeviceBatch(prods)
    batch.run()
    eager = batch.results()
    for (r, d, A, B), C in zip(prods, eager):
        assert C == naive(A, r, B, d)
    batch.prepare()
    junk = np.full(max(batch.total_out, 1), 0x5A5A5A5A5A5A5A5A, dtype=np.uint64)
    for _ in range(4):
        batch.out.upload(junk)
        batch.launch(graph=True)
        rf.device_sync()
        assert batch.results() == eager
        batch.out.upload(junk)
        batch.launch(graph=False)
        rf.device_sync()
        assert batch.results() == eager
    batch.release()


@pytest.mark.parametrize("lanes", [2, 3])
def test_prepared_batch_lanes_match_one_lane(lanes):
    """Products split over several overlapping lanes (own streams, forked and
    joined inside the captured graph) give the one-lane results."""
    import numpy as np
    prods = workloads.tree_level(20, seed=8) + workloads.headline(2, seed=42)
    batch = rf.DeviceBatch(prods)
    batch.run()
    want = batch.results()
    batch.prepare(lanes=lanes)
    junk = np.full(max(batch.total_out, 1), 0x77, dtype=np.uint64)
    for graph in (True, False, True):
        batch.out.upload(junk)
        batch.launch(graph=graph)
        rf.device_sync()
        assert batch.results() == want
    batch.release()


@pytest.mark.parametrize("t,split,bits", [(1, 0, 9000), (8, 0, 9000), (3, 0, 70000), (1, 1, 400000), (5, 1, 2500000), (2, 0, 1 << 20)])
def test_tuned_plans_give_the_same_products(t, split, bits):
    """Planner rules force other chunk widths (and the two-row split of
    2^14-point transforms); the products stay exact."""
    from rforest_amd import tune
    rng = random.Random(t * 100 + bits)
    A, B = _rand_product(rng, bits, 2, 2)
    tune.clear()
    tune.force(2, t, split)
    try:
        assert rf.plan(max(abs(x).bit_length() for x in A), max(abs(x).bit_length() for x in B), 2, 2)[0] == t
        assert rf.mpz_rmatrix_mult_fft(A, 2, B, 2) == naive(A, 2, B, 2)
        batch = rf.DeviceBatch([(2, 2, A, B)])
        batch.prepare()
        batch.launch(graph=True)
        rf.device_sync()
        assert batch.results()[0] == naive(A, 2, B, 2)
        batch.release()
    finally:
        tune.reset()


@pytest.mark.parametrize("bits,r", [(5000, 2), (300000, 3), (3000000, 2), (6 << 20, 2)])
def test_every_chunk_width_gives_the_same_products(bits, r):
    """Each valid chunk width (forced through a planner rule) packs, transforms
    and recomposes to the schoolbook product."""
    from rforest_amd import tune
    rng = random.Random(bits)
    A, B = _rand_product(rng, bits, r, r)
    want = naive(A, r, B, r)
    seen = 0
    try:
        for t in range(1, 9):
            tune.clear()
            tune.force(r, t)
            try:
                pt = rf.plan(max(abs(x).bit_length() for x in A), max(abs(x).bit_length() for x in B), r, r)[0]
            except rf.RfError:
                continue
            if pt != t:
                continue
            seen += 1
            batch = rf.DeviceBatch([(r, r, A, B)])
            batch.run()
            assert batch.results()[0] == want, t
    finally:
        tune.reset()
    assert seen >= 3
