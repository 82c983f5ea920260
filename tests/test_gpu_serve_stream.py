"""The request server on the served mixed stream at its real sizes, and with
resident left operands in the mix.

* A slice of the 20,000-request stream (r in {2,3,4}, entries log-uniform in
  [2^12, 2^24] bits, random signs, seed 1005) fed from host memory
  (RF_IN_HOST) with results copied back (RF_OUT_HOST), under a budget small
  enough that admission waits for memory: every result equals GMP
  schoolbook and equals the same request run alone, compared in full.
* Requests naming resident left operands (rf_request.left) mixed with plain
  ones: every result equals GMP schoolbook and the request alone.

The reference's end-to-end semantics are those of src/lib/hwmpz.c:65-130;
its transform-domain reuse is mpzfft_fft (src/lib/mpzfft.c:10-23).
"""
import random

import numpy as np
import pytest

import rforest_amd as rf
from helpers import naive
from rforest_amd import serve, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield


def _slice():
    """Requests of the stream's first 400: the first three with r >= 3 and
    entries of at least 2^22 bits, the largest with r = 2, and 30 of the
    others spread over the size range."""
    stream = workloads.served_stream(400, seed=1005)
    big = [q for q in stream if q[0] >= 3 and q[1][0] >= (1 << 22)][:3]
    big2 = max((q for q in stream if q[0] == 2), key=lambda q: q[1][0])
    rest = sorted((q for q in stream if q[1][0] < (1 << 22)), key=lambda q: q[1][0])
    picked = big + [big2] + rest[::max(1, len(rest) // 30)][:30]
    assert len(big) == 3 and min(q[1][0] for q in picked) < (1 << 14) and big2[1][0] >= (1 << 22)
    return picked


def _limbs(q):
    r, bits, signs, seed, _ = q
    return [(sg, serve.gen_limbs(b, seed, e), b) for e, (b, sg) in enumerate(zip(bits, signs))]


def _values(q):
    r, bits, signs, seed, _ = q
    vals = [serve.gen_value(b, sg, seed, e) for e, (b, sg) in enumerate(zip(bits, signs))]
    return vals[:r * r], vals[r * r:]


def _run(reqs, budget, streams, max_batch):
    srv = serve.Server(budget, streams=streams, max_batch=max_batch)
    ids = [srv.submit_host_limbs(q[0], q[0], _limbs(q), tag=i) for i, q in enumerate(reqs)]
    comps = {c["tag"]: c for c in srv.wait_all(len(ids), timeout_s=900)}
    out = []
    for i, rid in enumerate(ids):
        assert comps[i]["status"] == rf.RF_OK, i
        out.append(srv.host_result(rid, comps[i]["out_limbs"]))
    st = srv.stats()
    srv.close()
    return out, st


def test_real_size_stream_host_fed_equals_gmp_and_alone():
    reqs = _slice()
    print("case: served stream slice,", len(reqs), "requests, max entry bits", max(q[1][0] for q in reqs), flush=True)
    budget = 3 << 29  # 1.5 GiB: the largest request alone needs most of it
    mix, st = _run(reqs, budget, streams=3, max_batch=16)
    assert st["peak_bytes"] <= budget
    assert st["batches"] < len(reqs)
    alone, _ = _run(reqs, budget, streams=1, max_batch=1)
    assert mix == alone
    for q, C in zip(reqs, mix):
        A, B = _values(q)
        assert C == naive(A, q[0], B, q[0]), (q[0], q[1][0])


def _signed(rng, bits):
    x = rng.getrandbits(bits)
    return -x if rng.random() < 0.5 else x


def test_resident_operands_in_the_mix():
    print("case: resident operands in a served mix", flush=True)
    rng = random.Random(91)
    A1 = [_signed(rng, 1 << 20) for _ in range(9)]
    A2 = [_signed(rng, 1 << 16) for _ in range(4)]
    op1 = rf.ResidentOperand(A1, 3, 3, 1 << 20)
    op2 = rf.ResidentOperand(A2, 2, 2, 1 << 17)
    reqs = []  # (r, A, B, left)
    for i in range(26):
        kind = i % 3
        if kind == 0:
            reqs.append((3, A1, [_signed(rng, rng.randint(1000, 1 << 20)) for _ in range(9)], op1))
        elif kind == 1:
            reqs.append((2, A2, [_signed(rng, rng.randint(64, 1 << 17)) for _ in range(4)], op2))
        else:
            r = rng.choice([2, 3])
            b = rng.randint(100, 1 << 18)
            reqs.append((r, [_signed(rng, b) for _ in range(r * r)], [_signed(rng, b) for _ in range(r * r)], None))

    def run(streams, max_batch):
        srv = serve.Server(2 << 30, streams=streams, max_batch=max_batch)
        ids = [srv.submit_host(r, r, A, B, tag=i, left=left) for i, (r, A, B, left) in enumerate(reqs)]
        comps = {c["tag"]: c for c in srv.wait_all(len(ids))}
        out = []
        for i, rid in enumerate(ids):
            assert comps[i]["status"] == rf.RF_OK
            out.append(srv.host_result(rid, comps[i]["out_limbs"]))
        srv.close()
        return out

    mix = run(3, 8)
    alone = run(1, 1)
    assert mix == alone
    for (r, A, B, _), C in zip(reqs, mix):
        assert C == naive(A, r, B, r)
    # a right operand outside the operand's plan is refused at submission
    srv = serve.Server(1 << 30)
    with pytest.raises(rf.RfError) as ei:
        srv.submit_host(2, 2, A2, [1 << (1 << 19)] * 4, left=op2)
    assert ei.value.code == rf.RF_ERR_ARG
    srv.close()
    op1.release()
    op2.release()
