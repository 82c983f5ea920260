"""Seeded random sweep of the product over shapes and sizes (r, d in 1..5,
entries log-uniform up to 2^22 bits, uneven entries, random signs and
zeros), one product at a time and as one device batch, against GMP
schoolbook."""
import random

import pytest

import rforest_amd as rf
from helpers import naive

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield


def _case(rng):
    r, d = rng.randint(1, 5), rng.randint(1, 5)
    top = int(2 ** rng.uniform(0, 22))
    def ent():
        if rng.random() < 0.05:
            return 0
        b = max(1, int(top * rng.uniform(0.25, 1.0)))
        return rng.getrandbits(b) * rng.choice([-1, 1])
    return r, d, [ent() for _ in range(r * d)], [ent() for _ in range(d * d)]


def test_random_products_one_by_one():
    rng = random.Random(2024)
    for _ in range(40):
        r, d, A, B = _case(rng)
        assert rf.mpz_rmatrix_mult_fft(A, r, B, d) == naive(A, r, B, d), (r, d)


def test_random_products_as_one_batch():
    rng = random.Random(2025)
    prods = [_case(rng) for _ in range(60)]
    batch = rf.DeviceBatch(prods)
    batch.run()
    got = batch.results()
    batch.release()
    assert got == [naive(A, r, B, d) for (r, d, A, B) in prods]
