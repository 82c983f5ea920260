"""rf_batch_mult_device_noalloc: the same products as rf_batch_mult_device
with the pipeline's scratch in the library's fixed device array, checked
bit-exact against that entry point on two calls in a row (the second with
smaller inputs, so stale contents of the shared array would show), and
against GMP schoolbook; products too large for the array are refused."""
import random

import pytest

import rforest_amd as rf
from helpers import naive
from rforest_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield


def _signed(rng, bits):
    x = rng.getrandbits(bits) | (1 << (bits - 1))
    return -x if rng.random() < 0.5 else x


def _mixed(seed, n, lo, hi):
    rng = random.Random(seed)
    out = []
    for i in range(n):
        r = rng.choice([1, 2, 3])
        d = rng.choice([1, 2, 3])
        bits = int(2 ** rng.uniform(lo, hi))
        A = [_signed(rng, bits) if rng.random() > 0.05 else 0 for _ in range(r * d)]
        B = [_signed(rng, bits) if rng.random() > 0.05 else 0 for _ in range(d * d)]
        out.append((r, d, A, B))
    return out


def _both(prods):
    batch = rf.DeviceBatch(prods)
    batch.run()
    want = batch.results()
    batch.clear_outputs()
    batch.run_noalloc()
    got = batch.results()
    return want, got


def test_noalloc_matches_the_allocating_entry_point_twice():
    print("case: noalloc, headline product + mixed level, then a smaller set", flush=True)
    first = workloads.headline(1, seed=42) + _mixed(61, 40, 10, 17)
    want, got = _both(first)
    assert got == want
    r, d, A, B = first[0]
    assert got[0] == naive(A, r, B, d)
    second = _mixed(62, 12, 8, 13)  # smaller: stale scratch from the first call would show
    want2, got2 = _both(second)
    assert got2 == want2
    for (r, d, A, B), C in zip(second, got2):
        assert C == naive(A, r, B, d)


def test_noalloc_refuses_what_does_not_fit():
    print("case: noalloc capacity refusal", flush=True)
    batch = rf.DeviceBatch(workloads.headline(8, seed=42))  # 8 x 8.4 MiB of scratch > 64 MiB
    with pytest.raises(rf.RfError) as ei:
        batch.run_noalloc()
    assert ei.value.code == rf.RF_ERR_CAPACITY
    # and the fixed scratch still serves a product that fits
    small = _mixed(63, 4, 10, 12)
    want, got = _both(small)
    assert got == want
