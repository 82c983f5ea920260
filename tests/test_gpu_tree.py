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
