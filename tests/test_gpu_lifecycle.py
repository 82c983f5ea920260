"""Library lifecycle on the GPU: asynchronous batches on a caller's stream
followed at once by other calls, prepared batches and servers across
hw_mpz_clear / hw_mpz_setup, every result checked against GMP schoolbook.

The library's one device arena and pinned staging buffer serve every call;
an asynchronous batch on a caller's stream still reads them after
returning, so later calls must be ordered after it.  Prepared batches
capture table addresses into their graphs, which hw_mpz_clear frees; they
are captured again under the new set-up.  The reference's interface has the
same lifecycle calls (hw_mpz_setup / hw_mpz_clear, src/lib/hwmpz.c:24-52).
"""
import random

import pytest

import rforest_amd as rf
from helpers import naive
from rforest_amd import serve
from rforest_amd.gmp import MpzArray

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield
    rf.hw_mpz_setup()


def _signed(rng, bits):
    x = rng.getrandbits(bits) | (1 << (bits - 1))
    return -x if rng.random() < 0.5 else x


def _products(seed, n, bits, r=2):
    rng = random.Random(seed)
    return [(r, r, [_signed(rng, bits) for _ in range(r * r)], [_signed(rng, bits) for _ in range(r * r)])
            for _ in range(n)]


def _dropin(A, r, B, d):
    """C = A B through the mpz_t drop-in (rf_mpz_rmatrix_mult)."""
    mA, mB, mC = MpzArray(A), MpzArray(B), MpzArray.zeros(r * d)
    try:
        rf.check(rf.lib().rf_mpz_rmatrix_mult(mC.ptr(), mA.ptr(), r, mB.ptr(), d))
        return mC.values()
    finally:
        mA.close()
        mB.close()
        mC.close()


@pytest.mark.parametrize("grow", [False, True])
def test_async_batch_on_user_stream_then_dropin(grow):
    """An asynchronous one-shot batch on a caller's stream, then at once a
    drop-in call (which reuses the arena and staging buffer, and with
    grow=True replaces the arena by a larger one): both exact."""
    print("case: async user-stream batch, then drop-in, grow =", grow, flush=True)
    prods = _products(11, 48, 200000)
    batch = rf.DeviceBatch(prods)
    st = rf.HipStream()
    try:
        batch.run(sync=False, stream=st)
        rng = random.Random(12)
        bits = 4_000_000 if grow else 150000
        A = [_signed(rng, bits) for _ in range(4)]
        B = [_signed(rng, bits) for _ in range(4)]
        C = _dropin(A, 2, B, 2)
        st.synchronize()
        assert C == naive(A, 2, B, 2)
        got = batch.results()
        for (r, d, a, b), c in zip(prods, got):
            assert c == naive(a, r, b, d)
    finally:
        st.destroy()


def test_two_async_batches_on_different_streams():
    """Asynchronous batches on the library's stream and on a caller's stream
    back to back (both use the arena): both exact."""
    print("case: async batches on two streams", flush=True)
    p1, p2 = _products(21, 32, 120000), _products(22, 32, 90000, r=3)
    b1, b2 = rf.DeviceBatch(p1), rf.DeviceBatch(p2)
    st = rf.HipStream()
    try:
        b1.run(sync=False)
        b2.run(sync=False, stream=st)
        st.synchronize()
        for prods, b in ((p1, b1), (p2, b2)):
            for (r, d, a, bb), c in zip(prods, b.results()):
                assert c == naive(a, r, bb, d)
    finally:
        st.destroy()


def test_graph_replay_after_clear_and_setup():
    """A prepared batch replayed from its graph on a caller's stream, then
    hw_mpz_clear and hw_mpz_setup (new table addresses), then replayed
    again: the second replay captures a new graph and is exact."""
    print("case: graph replay across clear/setup", flush=True)
    prods = _products(31, 16, 300000)
    batch = rf.DeviceBatch(prods)
    batch.prepare(lanes=2)
    st = rf.HipStream()
    try:
        batch.launch(graph=True, stream=st)
        st.synchronize()
        want = [naive(a, r, b, d) for (r, d, a, b) in prods]
        assert batch.results() == want
        rf.hw_mpz_clear()
        rf.hw_mpz_setup()
        batch.clear_outputs()
        batch.launch(graph=True, stream=st)
        st.synchronize()
        assert batch.results() == want
        # and once more on the library's stream (another capture)
        batch.clear_outputs()
        batch.launch(graph=True)
        rf.device_sync()
        assert batch.results() == want
    finally:
        st.destroy()
        batch.release()


def test_server_across_clear_and_setup():
    """A server created before hw_mpz_clear: requests submitted while the
    engine is cleared are refused (RF_ERR_NOT_INIT); after hw_mpz_setup the
    same server serves exact results again."""
    print("case: server across clear/setup", flush=True)
    rng = random.Random(41)
    srv = serve.Server(hbm_budget=2 << 30, streams=2, max_batch=8)
    try:
        A = [_signed(rng, 50000) for _ in range(4)]
        B = [_signed(rng, 50000) for _ in range(4)]
        rid = srv.submit_host(2, 2, A, B)
        (c,) = srv.wait_all(1)
        assert c["status"] == rf.RF_OK and srv.host_result(rid, c["out_limbs"]) == naive(A, 2, B, 2)
        rf.hw_mpz_clear()
        with pytest.raises(rf.RfError) as ei:
            srv.submit_host(2, 2, A, B)
        assert ei.value.code == rf.RF_ERR_NOT_INIT
        rf.hw_mpz_setup()
        A2 = [_signed(rng, 80000) for _ in range(9)]
        B2 = [_signed(rng, 80000) for _ in range(9)]
        rid = srv.submit_host(3, 3, A2, B2)
        (c,) = srv.wait_all(1)
        assert c["status"] == rf.RF_OK and srv.host_result(rid, c["out_limbs"]) == naive(A2, 3, B2, 3)
    finally:
        srv.close()
