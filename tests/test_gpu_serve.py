"""The request server: exact results for mixed requests, each identical
alone and in the mix, the HBM budget respected, kept results readable."""
import math
import random

import numpy as np
import pytest

import rforest_amd as rf
from helpers import naive
from rforest_amd import serve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield


def _mixed_stream(n, seed, lo=12, hi=20):
    """Requests of the served mixed stream's shape: r in {2, 3, 4} with
    probabilities {0.6, 0.3, 0.1}, bits log-uniform in [2^lo, 2^hi], random
    signs (a few zero entries)."""
    rng = random.Random(seed)
    out = []
    for i in range(n):
        u = rng.random()
        r = 2 if u < 0.6 else (3 if u < 0.9 else 4)
        L = int(2 ** (lo + rng.random() * (hi - lo)))
        ents = 2 * r * r
        bits = [L] * ents
        signs = [rng.choice([-1, 1]) if rng.random() > 0.03 else 0 for _ in range(ents)]
        out.append((r, bits, signs, (seed << 20) + i))
    return out


def _values(r, bits, signs, seed):
    vals = [serve.gen_value(b, s, seed, e) for e, (b, s) in enumerate(zip(bits, signs))]
    return vals[:r * r], vals[r * r:]


def test_synthetic_stream_on_device_matches_host():
    buf = rf.DeviceBuffer(8 * 4096)
    for bits, seed, e in [(1, 5, 0), (64, 6, 3), (65, 7, 1), (100000, 8, 31), (64 * 4096, 9, 2)]:
        serve.gen_entry_device(buf.ptr, bits, seed, e)
        n = (bits + 63) // 64
        got = buf.download(np.uint64, n)
        assert np.array_equal(got, serve.gen_limbs(bits, seed, e))
    buf.free()


def test_checksum_on_device_matches_host():
    rng = np.random.default_rng(3)
    buf = rf.DeviceBuffer(8 * 10000)
    for n, e in [(1, 0), (2047, 1), (2048, 2), (2049, 3), (10000, 15)]:
        a = rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64, endpoint=True)
        buf.upload(a)
        assert serve.checksum_device(buf.ptr, n, e) == serve.checksum_limbs(a, e)
    buf.free()


def test_server_host_inputs_exact():
    rng = random.Random(11)
    srv = serve.Server(1 << 30)
    reqs = []
    for i in range(40):
        r = rng.choice([1, 2, 3, 4])
        bits = rng.choice([1, 64, 700, 5000, 40000, 150000])
        A = [rng.getrandbits(bits) * rng.choice([-1, 1]) for _ in range(r * r)]
        B = [rng.getrandbits(max(1, bits // rng.choice([1, 3]))) * rng.choice([-1, 1]) for _ in range(r * r)]
        if i % 7 == 0:
            A[0] = 0
        if i == 5:
            B = [0] * (r * r)
        rid = srv.submit_host(r, r, A, B, tag=i)
        reqs.append((rid, r, A, B))
    comps = {c["id"]: c for c in srv.wait_all(len(reqs))}
    for rid, r, A, B in reqs:
        c = comps[rid]
        assert c["status"] == rf.RF_OK
        C = srv.host_result(rid, c["out_limbs"])
        assert C == naive(A, r, B, r)
        assert c["checksum"] == serve.checksum_values(C, c["out_limbs"])
        assert c["t_submit"] <= c["t_admit"] <= c["t_done"]
    srv.close()


def test_zero_result_alone_in_its_batch_has_checksum_zero():
    # A batch whose every result has no limbs sums nothing on the device; its
    # checksums must be 0, not what the slot's previous batch left behind.
    srv = serve.Server(1 << 30, streams=1, max_batch=1)
    for i, (A, B) in enumerate([([3 << 200], [5 << 100]), ([0], [7]), ([1 << 64], [1]), ([5], [0])]):
        rid = srv.submit_host(1, 1, A, B, tag=i)
        (c,) = srv.wait_all(1)
        C = srv.host_result(rid, c["out_limbs"])
        assert C == [A[0] * B[0]]
        assert c["checksum"] == serve.checksum_values(C, c["out_limbs"]), i
    srv.close()


def test_server_mix_equals_alone_and_gmp():
    stream = _mixed_stream(150, 77)
    srv = serve.Server(8 << 30, streams=3, max_batch=64)
    ids = [srv.submit_synth(r, r, bits, signs, seed, tag=i, output=serve.RF_OUT_CHECKSUM | serve.RF_OUT_KEEP)
           for i, (r, bits, signs, seed) in enumerate(stream)]
    mix = {c["tag"]: c for c in srv.wait_all(len(ids))}
    st = srv.stats()
    assert st["batches"] < len(ids)  # requests shared batches
    full = [srv.read(ids[i], r, r, mix[i]["out_limbs"]) for i, (r, _, _, _) in enumerate(stream)]
    # a sample against GMP schoolbook on the host-generated inputs
    for i in range(0, len(stream), 15):
        r, bits, signs, seed = stream[i]
        A, B = _values(r, bits, signs, seed)
        assert full[i] == naive(A, r, B, r)
    for rid in ids:
        srv.release(rid)
    srv.close()
    # each request alone in a fresh server: the same full value (and checksum)
    srv = serve.Server(8 << 30, streams=1, max_batch=1)
    for i, (r, bits, signs, seed) in enumerate(stream):
        rid = srv.submit_synth(r, r, bits, signs, seed, tag=i, output=serve.RF_OUT_CHECKSUM | serve.RF_OUT_KEEP)
        (c,) = srv.wait_all(1)
        assert c["status"] == rf.RF_OK
        assert c["checksum"] == mix[i]["checksum"] and c["out_limbs"] == mix[i]["out_limbs"], i
        assert srv.read(rid, r, r, c["out_limbs"]) == full[i], i
        srv.release(rid)
    srv.close()


def test_server_budget_is_respected():
    budget = 96 << 20
    srv = serve.Server(budget, streams=3, max_batch=32)
    stream = _mixed_stream(60, 5, lo=14, hi=19)
    for i, (r, bits, signs, seed) in enumerate(stream):
        srv.submit_synth(r, r, bits, signs, seed, tag=i)
    comps = srv.wait_all(len(stream))
    assert all(c["status"] == rf.RF_OK for c in comps)
    st = srv.stats()
    assert 0 < st["peak_bytes"] <= budget
    with pytest.raises(rf.RfError) as e:
        srv.submit_synth(4, 4, [1 << 23] * 32, [1] * 32, 1)
    assert e.value.code == rf.RF_ERR_CAPACITY
    srv.close()


def test_server_unsupported_request_is_refused():
    srv = serve.Server(1 << 30)
    with pytest.raises(rf.RfError) as e:
        srv.submit_synth(2, 2, [1 << 27] * 8, [1] * 8, 1)
    assert e.value.code in (rf.RF_ERR_UNSUPPORTED, rf.RF_ERR_CAPACITY)
    srv.close()


def test_synthetic_device_batch_checksums_match_host():
    """Device batches of synthetic entries (generated in place) and their
    device checksums agree with GMP products of the host-generated entries."""
    rng = random.Random(21)
    specs = []
    for i in range(12):
        r = rng.choice([1, 2, 3])
        bits = [rng.randrange(1, 30000) for _ in range(2 * r * r)]
        signs = [rng.choice([-1, 0, 1]) for _ in range(2 * r * r)]
        specs.append((r, r, bits, signs, rng.getrandbits(64)))
    batch = rf.DeviceBatch.synthetic(specs)
    batch.run()
    sums = batch.checksums()
    res = batch.results()
    for (r, _, bits, signs, seed), h, C in zip(specs, sums, res):
        vals = [serve.gen_value(b, s, seed, e) for e, (b, s) in enumerate(zip(bits, signs))]
        want = naive(vals[:r * r], r, vals[r * r:], r)
        assert C == want
        la = max([(b + 63) // 64 for b, s in zip(bits[:r * r], signs[:r * r]) if s] or [0])
        lb = max([(b + 63) // 64 for b, s in zip(bits[r * r:], signs[r * r:]) if s] or [0])
        rn = int(rf.lib().rf_out_limbs(la, lb, r))
        assert h == serve.checksum_values(want, rn)
