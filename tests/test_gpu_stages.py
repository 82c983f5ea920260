"""Stage-level parity of the GPU kernels with the oracle, bit-exact.

Each stage runs through librforest_hip.so's stage entry points (the same
kernels as the full product) and is compared with the oracle's restatement
of the corresponding reference stage on the same seeded inputs."""
import random

import pytest

import oracle
import rforest_amd as rf
from helpers import from_tc
from rforest_amd import stages

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield


def _expect_fwd(limbs, sign, t, nchunk, lgN, p, scale):
    if sign == 0:  # a zero entry (mpz_sgn == 0, src/lib/mpnfft.c:122-126): zero polynomial
        return [0] * (1 << lgN)
    w = oracle.root(p, 24, lgN)
    ch = oracle.split(limbs, nchunk, 64 * t)
    red = oracle.reduce(ch, nchunk, t, sign, p)
    if scale:
        s = pow(pow(2, scale, p), p - 2, p)
        red = [x * s % p for x in red]
    return oracle.ntt_fwd(red, lgN, p, w)


@pytest.mark.parametrize("lgN", list(range(1, 15)))
def test_fwd_stage(lgN):
    rng = random.Random(lgN)
    P = rf.primes(32)
    N = 1 << lgN
    for t in ((1, 3, 4, 8) if lgN < 11 else (1, 2, 3, 4, 5, 6, 7, 8)):
        nchunk = max(1, N // 2)
        ents = []
        for sign in (1, -1, 0):
            nl = rng.randrange(max(1, nchunk * t - 3), nchunk * t + 1)  # ragged last chunk
            ents.append(([rng.getrandbits(64) for _ in range(nl)], sign))
        ents.append(([(1 << 64) - 1] * (nchunk * t), 1))  # all-ones magnitude
        for q in (0, 17, 23):
            for scale in (0, lgN):
                got = stages.fwd(ents, t, nchunk, lgN, q, scale)
                for e, (lv, s) in enumerate(ents):
                    assert got[e] == _expect_fwd(lv, s, t, nchunk, lgN, P[q], scale), (t, q, scale, e)


def test_fwd_stage_full_length_input():
    # nchunk = N (no zero padding): every input position populated
    rng = random.Random(9)
    P = rf.primes(1)
    for lgN in (1, 4, 10, 13):
        N = 1 << lgN
        ents = [([rng.getrandbits(64) for _ in range(2 * N)], -1)]
        got = stages.fwd(ents, 2, N, lgN, 0, 0)
        assert got[0] == _expect_fwd(ents[0][0], -1, 2, N, lgN, P[0], 0)


@pytest.mark.parametrize("lgN", [18, 19, 20, 21])
def test_fwd_inv_stage_longest(lgN):
    """Transforms above 2^17 points (column, middle and row passes; row and two
    column passes back), checked without an O(N log N) Python transform:
    known answers (delta_0 -> all ones; the chunk-1 delta -> omega^rev(m),
    which pins every twiddle), and the round trip inv(fwd(x) . 1) = N x for a
    random x filling half the points and for one filling all of them."""
    rng = random.Random(lgN)
    P = rf.primes(24)
    N = 1 << lgN
    for q in (0, 23):
        p = P[q]
        w = oracle.root(p, 24, lgN)
        ones = stages.fwd([([1], 1)], 1, 1, lgN, q, 0)[0]
        assert ones == [1] * N
        e1 = stages.fwd([([0, 1], 1)], 1, 2, lgN, q, 0)[0]
        rev = [0]
        for _ in range(lgN):  # bit-reversal permutation of lgN bits
            rev = [2 * r for r in rev] + [2 * r + 1 for r in rev]
        pw = [1] * N
        for i in range(1, N):
            pw[i] = pw[i - 1] * w % p
        assert e1 == [pw[r] for r in rev]
        for nchunk, t in ((N // 2, 1), (N, 2)):
            limbs = [rng.getrandbits(64) for _ in range(nchunk * t)]
            spec = stages.fwd([(limbs, -1)], t, nchunk, lgN, q, 0)
            back = stages.inv(spec, [ones], 1, 1, lgN, q)[0]
            red = oracle.reduce(oracle.split(limbs, nchunk, 64 * t), nchunk, t, -1, p)
            assert back == [x * N % p for x in red + [0] * (N - nchunk)]


@pytest.mark.parametrize("lgN,q", [(22, 23), (23, 5), (24, 0)])
def test_fwd_inv_stage_beyond_2_21(lgN, q):
    """Transforms of 2^22 .. 2^24 points (three column passes each way and the
    twiddle tables of order 2^24), on numpy arrays: delta_0 -> all ones; the
    chunk-1 delta -> omega^rev(m) at 4096 sampled positions and the pairs
    y[2m + 1] = -y[2m] everywhere; the round trip inv(fwd(x) . 1) = N x for a
    random negative x filling every point."""
    import numpy as np
    p = rf.primes(24)[q]
    N = 1 << lgN
    ones = stages.fwd_array(np.array([1], np.uint64), 1, 1, 1, lgN, q)
    assert bool(np.all(ones == 1))
    e1 = stages.fwd_array(np.array([0, 1], np.uint64), 1, 1, 2, lgN, q)
    w = oracle.root(p, 24, lgN)
    rng = random.Random(lgN)
    for m in [0, 1, N // 2, N - 1] + [rng.randrange(N) for _ in range(4096)]:
        rev = int(format(m, "0%db" % lgN)[::-1], 2)
        assert int(e1[m]) == pow(w, rev, p), m
    assert bool(np.all((e1[0::2] + e1[1::2]) % np.uint64(p) == 0))
    gen = np.random.default_rng(lgN)
    x = gen.integers(0, 1 << 63, size=N, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    spec = stages.fwd_array(x, -1, 1, N, lgN, q)
    back = stages.inv_array(spec, ones, 1, 1, lgN, q)
    pp = np.uint64(p)
    want = x % pp
    for _ in range(lgN):  # N x mod p by doubling (values stay below 2^52)
        want = (want * np.uint64(2)) % pp
    want = (pp - want) % pp  # the entry is negative
    assert bool(np.array_equal(back, want))


def test_long_tables_made_again_after_clear():
    """The order-2^24 twiddle tables are freed by hw_mpz_clear and made again
    by the next plan that needs them after hw_mpz_setup."""
    import numpy as np
    lgN, q = 22, 7
    p = rf.primes(24)[q]
    w = oracle.root(p, 24, lgN)
    for _ in range(2):
        e1 = stages.fwd_array(np.array([0, 1], np.uint64), 1, 1, 2, lgN, q)
        for m in (1, 3, 1 << 20, (1 << 22) - 1):
            rev = int(format(m, "0%db" % lgN)[::-1], 2)
            assert int(e1[m]) == pow(w, rev, p)
        rf.hw_mpz_clear()
        rf.hw_mpz_setup()


@pytest.mark.parametrize("lgN", [15, 16, 17])
def test_fwd_stage_long(lgN):
    """Column pass + row pass transforms (beyond one workgroup's LDS): half
    filled (the column pass's copy stage), full length, and a short entry
    (whole rows of zeros)."""
    rng = random.Random(lgN)
    P = rf.primes(32)
    N = 1 << lgN
    for t, nchunk, scale in ((1, N // 2, 0), (8, N // 2, lgN), (2, N, 0), (3, 1000, lgN), (5, N // 2, 0),
                             (7, N // 4, lgN), (6, N // 2 - 3, 0)):
        ents = [([rng.getrandbits(64) for _ in range(nchunk * t - 1)], 1),
                ([rng.getrandbits(64) for _ in range(nchunk * t)], -1)]
        for q in (0, 23):
            got = stages.fwd(ents, t, nchunk, lgN, q, scale)
            for e, (lv, s) in enumerate(ents):
                assert got[e] == _expect_fwd(lv, s, t, nchunk, lgN, P[q], scale), (t, nchunk, q, e)


@pytest.mark.parametrize("lgN", [15, 16, 17])
@pytest.mark.parametrize("r,d", [(1, 1), (2, 2), (1, 3)])
def test_inv_stage_long(lgN, r, d):
    rng = random.Random(lgN * 100 + r * 10 + d)
    P = rf.primes(32)
    N = 1 << lgN
    q = rng.randrange(24)
    p = P[q]
    w = oracle.root(p, 24, lgN)
    A = [[rng.randrange(p) for _ in range(N)] for _ in range(r * d)]
    B = [[rng.randrange(p) for _ in range(N)] for _ in range(d * d)]
    A[0][: N // 2] = [p - 1] * (N // 2)
    got = stages.inv(A, B, r, d, lgN, q)
    C = oracle.pointwise(sum(A, []), sum(B, []), r, d, N, p)
    for e in range(r * d):
        assert got[e] == oracle.ntt_inv(C[e * N:(e + 1) * N], lgN, p, w, False)


@pytest.mark.parametrize("lgN", [1, 2, 3, 6, 10, 13, 14])
@pytest.mark.parametrize("r,d", [(1, 1), (2, 2), (3, 2), (2, 3), (4, 4), (1, 9)])
def test_inv_stage(lgN, r, d):
    rng = random.Random(lgN * 100 + r * 10 + d)
    P = rf.primes(32)
    N = 1 << lgN
    q = rng.randrange(24)
    p = P[q]
    w = oracle.root(p, 24, lgN)
    A = [[rng.randrange(p) for _ in range(N)] for _ in range(r * d)]
    B = [[rng.randrange(p) for _ in range(N)] for _ in range(d * d)]
    A[0] = [p - 1] * N  # extreme residues
    got = stages.inv(A, B, r, d, lgN, q)
    C = oracle.pointwise(sum(A, []), sum(B, []), r, d, N, p)
    for e in range(r * d):
        assert got[e] == oracle.ntt_inv(C[e * N:(e + 1) * N], lgN, p, w, False)


@pytest.mark.parametrize("k", [1, 2, 3, 6, 11, 12, 17, 22, 24])
def test_crt_stage(k):
    rng = random.Random(k)
    P = rf.primes(32)[:k]
    Pk = 1
    for p in P:
        Pk *= p
    for t in (1, 2, 4, 8):
        for nout in (1, 2, 257, 700):
            cs = []
            for _ in range(nout):
                u = rng.random()
                if u < 0.05:
                    # within one mixed-radix digit of the lift threshold
                    # (P - 1) / 2: the digits above position j tie with it
                    j = rng.randrange(k)
                    Mj = 1
                    for p in P[:j]:
                        Mj *= p
                    c = rng.choice([1, -1]) * ((Pk - 1) // 2 - rng.choice([0, 1]) * Mj)
                elif u < 0.1:
                    c = (Pk - 1) // 2
                elif u < 0.2:
                    c = -((Pk - 1) // 2)
                elif u < 0.3:
                    c = rng.choice([-1, 0, 1])
                else:
                    c = rng.randrange(-(Pk // 2) + 1, Pk // 2)
                cs.append(c)
            res = [c % p for p in P for c in cs]
            total = sum(c << (64 * t * i) for i, c in enumerate(cs))
            rn = (abs(total).bit_length() + 1 + 63) // 64 + rng.randrange(0, 3)
            got = stages.crt(res, nout, k, t, rn)
            assert from_tc(got) == total, (t, nout)


def test_crt_stage_long_carry_chains():
    # coefficients that recompose to 2^(64 rn - 1) - 1 and to -2^j: carries
    # and borrows ripple across every digit block
    P = rf.primes(32)
    k, t = 3, 1
    Pk = P[0] * P[1] * P[2]
    top = (1 << 64) - 1
    nout = 2000
    cs = [top] * nout                      # sum = 2^(64 nout) - 1 with t = 1
    total = sum(c << (64 * i) for i, c in enumerate(cs))
    assert total == (1 << (64 * nout)) - 1
    res = [c % p for p in P[:k] for c in cs]
    rn = nout + 1
    assert from_tc(stages.crt(res, nout, k, t, rn)) == total
    cs = [0] * (nout - 1) + [-1]            # one negative top coefficient
    cs[0] = 1
    total = sum(c << (64 * i) for i, c in enumerate(cs))
    res = [c % p for p in P[:k] for c in cs]
    assert from_tc(stages.crt(res, nout, k, t, rn)) == total
    assert Pk > 2 * top
