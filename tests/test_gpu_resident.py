"""Resident left operands (rf_operand_*): a left matrix transformed once and
reused across products must give exactly the products computed from
scratch, through every entry point that accepts it."""
import random

import numpy as np
import pytest

import rforest_amd as rf
from helpers import naive
from rforest_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield


def _signed(rng, bits, n):
    return [rng.getrandbits(bits) * rng.choice([-1, 1]) for _ in range(n)]


@pytest.mark.parametrize("la,lb,r", [(50000, 20000, 3), (3000, 3000, 2), (200000, 1000, 4), (64, 64, 1)])
def test_resident_batch_matches_schoolbook_and_plain(la, lb, r):
    rng = random.Random(la + lb + r)
    A = _signed(rng, la, r * r)
    Bs = [_signed(rng, rng.randrange(1, lb + 1), r * r) for _ in range(5)]
    op = rf.ResidentOperand(A, r, r, lb)
    batch = rf.DeviceBatch([(r, r, op, B) for B in Bs])
    batch.run()
    got = batch.results()
    for B, C in zip(Bs, got):
        assert C == naive(A, r, B, r)
    plain = rf.DeviceBatch([(r, r, A, B) for B in Bs])
    plain.run()
    assert plain.results() == got
    op.release()


def test_resident_host_mult_and_plan():
    rng = random.Random(2)
    A = _signed(rng, 30000, 4)
    op = rf.ResidentOperand(A, 2, 2, 40000)
    t, lgN, k = op.plan()
    assert (t, lgN, k) == rf.plan(64 * ((30000 + 63) // 64), 40000, 2, 2)
    for bits in (40000, 1, 777):
        B = _signed(rng, bits, 4)
        assert op.mult(B) == naive(A, 2, B, 2)
    op.release()


def test_resident_right_operand_too_large_is_refused():
    rng = random.Random(3)
    op = rf.ResidentOperand(_signed(rng, 5000, 4), 2, 2, 1000)
    with pytest.raises(rf.RfError) as e:
        op.mult(_signed(rng, 200000, 4))
    assert e.value.code == rf.RF_ERR_ARG
    op.release()


def test_resident_zero_operands():
    rng = random.Random(4)
    op = rf.ResidentOperand([0, 0, 0, 0], 2, 2, 5000)
    assert op.mult(_signed(rng, 5000, 4)) == [0] * 4
    op.release()
    A = _signed(rng, 5000, 4)
    A[1] = 0
    op = rf.ResidentOperand(A, 2, 2, 5000)
    assert op.mult([0] * 4) == [0] * 4
    B = _signed(rng, 4000, 4)
    assert op.mult(B) == naive(A, 2, B, 2)
    op.release()


def test_resident_mixed_with_plain_products_and_graph_replay():
    """Resident and ordinary products in one prepared batch, replayed from a
    captured hipGraph with clobbered outputs in between."""
    rng = random.Random(5)
    A = _signed(rng, 80000, 9)
    op = rf.ResidentOperand(A, 3, 3, 30000)
    prods = [(3, 3, op, _signed(rng, 30000, 9)) for _ in range(3)]
    prods += [(2, 2, _signed(rng, 10000, 4), _signed(rng, 9000, 4)) for _ in range(3)]
    prods.insert(2, (3, 3, A, prods[0][3]))
    batch = rf.DeviceBatch(prods)
    batch.run()
    eager = batch.results()
    for (r, d, X, B), C in zip(prods, eager):
        assert C == naive(A if X is op else X, r, B, d)
    assert eager[2] == eager[0]
    batch.prepare()
    junk = np.full(max(batch.total_out, 1), 0x3C3C3C3C3C3C3C3C, dtype=np.uint64)
    for _ in range(3):
        batch.out.upload(junk)
        batch.launch(graph=True)
        rf.device_sync()
        assert batch.results() == eager
    batch.release()
    op.release()


def test_resident_long_transform():
    """A left operand whose plan needs more than 2^14 points."""
    rng = random.Random(6)
    A = _signed(rng, 5 << 20, 4)
    op = rf.ResidentOperand(A, 2, 2, 5 << 20)
    assert op.plan()[1] > 14
    Bs = [_signed(rng, 5 << 20, 4), _signed(rng, 100000, 4)]
    batch = rf.DeviceBatch([(2, 2, op, B) for B in Bs])
    batch.run()
    for B, C in zip(Bs, batch.results()):
        assert C == naive(A, 2, B, 2)
    op.release()


def test_resident_case_all_sixteen_products():
    """The resident-operand workload in full: one 3x3 left operand of
    4,194,304-bit entries, transformed once, times all 16 right 3x3 operands
    of 1,048,576-bit entries (seed 1004) in one batch; every product equals
    GMP schoolbook."""
    from rforest_amd.gmp import rmatrix_mult_schoolbook
    A, Bs = workloads.resident_case(1004)
    assert len(Bs) == 16
    op = rf.ResidentOperand(A, 3, 3, 1 << 20)
    batch = rf.DeviceBatch([(3, 3, op, B) for B in Bs])
    batch.run()
    got = batch.results()
    for i, B in enumerate(Bs):
        assert got[i] == rmatrix_mult_schoolbook(A, 3, B, 3), f"right operand {i}"
    batch.release()
    op.release()


def test_resident_case_two_products():
    """The resident-operand workload's shapes (3x3, 4 Mbit left, 1 Mbit
    right) on two of its right operands."""
    A, Bs = workloads.resident_case(1004, n_right=2)
    op = rf.ResidentOperand(A, 3, 3, 1 << 20)
    batch = rf.DeviceBatch([(3, 3, op, B) for B in Bs])
    batch.run()
    for B, C in zip(Bs, batch.results()):
        assert C == naive(A, 3, B, 3)
    op.release()
