"""The general transform API (rf_fft_*, the reference's mpzfft interface,
src/include/mpzfft.h:13-120): integers moved into the transform domain,
combined there and brought back, against Python integer arithmetic."""
import random

import pytest

import rforest_amd as rf
from rforest_amd.device import DeviceBuffer
from rforest_amd.fft import FftParams, matrix_mul

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def setup():
    rf.hw_mpz_setup()
    yield


def _r(rng, bits):
    return rng.getrandbits(bits) * rng.choice([-1, 1])


@pytest.mark.parametrize("bits", [1, 64, 65, 1000, 1 << 16, 1 << 20, 5 << 20, 1 << 23])
def test_forward_inverse_round_trip(bits):
    rng = random.Random(bits)
    P = FftParams(bits, 1)
    for x in (0, 1, -1, (1 << bits) - 1, -((1 << bits) - 1), _r(rng, bits)):
        assert P.value(x).inverse() == x, (P.plan(), x.bit_length())


@pytest.mark.parametrize("bx,by", [(1, 1), (64, 64), (3000, 17), (1 << 19, 1 << 19), (1 << 21, 3 << 20), (12345, 1 << 22)])
def test_product(bx, by):
    rng = random.Random(bx * 7 + by)
    P = FftParams(bx + by, 1)
    x, y = _r(rng, bx), _r(rng, by)
    Z = P.value().mul(P.value(x), P.value(y))
    assert Z.inverse() == x * y, P.plan()


def test_linear_combinations_and_aliasing():
    rng = random.Random(5)
    n = 200000
    P = FftParams(2 * n + 3, 4)
    a, b, c, d, e = (_r(rng, n) for _ in range(5))
    A, B, C, D, E = (P.value(v) for v in (a, b, c, d, e))
    Z = P.value().mul(A, B)
    W = P.value().mul(C, D)
    Z.add(Z, W)                      # a b + c d (aliased output)
    Z.sub(Z, P.value().mul(E, E))    # - e^2
    assert Z.inverse() == a * b + c * d - e * e
    Z.neg(Z)
    assert Z.inverse() == -(a * b + c * d - e * e)
    Y = P.value().set(Z)
    Y.mul(Y, P.value(1))             # times one, in place
    assert Y.inverse() == -(a * b + c * d - e * e)
    A.mul(A, A)                      # a^2 in place
    assert A.inverse() == a * a


@pytest.mark.parametrize("dims,bits", [((2, 2, 2), 30000), ((2, 3, 4), 5000), ((3, 3, 3), 1 << 20), ((1, 9, 1), 777)])
def test_matrix_mul(dims, bits):
    d1, d2, d3 = dims
    rng = random.Random(d1 * 100 + d2 * 10 + d3)
    P = FftParams(2 * bits + d2.bit_length() + 1, d2)
    X = [_r(rng, bits) for _ in range(d1 * d2)]
    Y = [_r(rng, bits) for _ in range(d2 * d3)]
    FX = [P.value(v) for v in X]
    FY = [P.value(v) for v in Y]
    FZ = [P.value() for _ in range(d1 * d3)]
    matrix_mul(FZ, FX, FY, d1, d2, d3)
    want = [sum(X[i * d2 + l] * Y[l * d3 + j] for l in range(d2)) for i in range(d1) for j in range(d3)]
    assert [z.inverse() for z in FZ] == want


def test_matrix_mul_in_place():
    """rop aliasing op1 (the reference allows parameter aliasing)."""
    rng = random.Random(9)
    bits = 4000
    P = FftParams(2 * bits + 3, 2)
    X = [_r(rng, bits) for _ in range(4)]
    Y = [_r(rng, bits) for _ in range(4)]
    FX = [P.value(v) for v in X]
    FY = [P.value(v) for v in Y]
    matrix_mul(FX, FX, FY, 2, 2, 2)
    want = [sum(X[i * 2 + l] * Y[l * 2 + j] for l in range(2)) for i in range(2) for j in range(2)]
    assert [z.inverse() for z in FX] == want


def test_device_forward_and_inverse():
    import numpy as np
    rng = random.Random(11)
    bits = 300000
    x, y = _r(rng, bits), _r(rng, bits)
    P = FftParams(2 * bits, 1)
    X, Y = P.value(), P.value()
    for F, v in ((X, x), (Y, y)):
        m = abs(v)
        n = (m.bit_length() + 63) // 64
        buf = DeviceBuffer(8 * n)
        buf.upload(np.frombuffer(m.to_bytes(8 * n, "little"), dtype=np.uint64))
        e = rf.DevEntry()
        e.limbs, e.nlimbs, e.sign = buf.ptr, n, (v > 0) - (v < 0)
        F.forward_device(e)
        buf.free()
    Z = P.value().mul(X, Y)
    cap = Z.out_limbs()
    out = DeviceBuffer(8 * cap)
    sign, n = Z.inverse_device(out.ptr, cap)
    mag = int.from_bytes(out.download(np.uint64, cap)[:n].tobytes(), "little")
    assert sign * mag == x * y
    with pytest.raises(rf.RfError):
        Z.inverse_device(out.ptr, cap - 1)


def test_errors():
    P1, P2 = FftParams(1000, 1), FftParams(1 << 20, 1)
    with pytest.raises(rf.RfError):
        P1.value(1 << 1000)          # more bits than the bound
    with pytest.raises(rf.RfError):
        P1.value().add(P1.value(1), P2.value(1))   # different plans


def test_longest_transforms_round_trip_and_product():
    """Values whose plans need more than 2^17 points (three-pass forward,
    row plus two column passes back) through the transform domain."""
    rng = random.Random(21)
    bits = 1 << 27
    P = FftParams(2 * bits + 1, 1)
    t, lg, k = P.plan()
    assert lg > 17
    x, y = _r(rng, bits), _r(rng, bits)
    X, Y = P.value(x), P.value(y)
    assert X.inverse() == x
    Z = P.value().mul(X, Y)
    from rforest_amd.gmp import rmatrix_mult_schoolbook
    assert Z.inverse() == rmatrix_mult_schoolbook([x], 1, [y], 1)[0]


def test_transforms_beyond_2_21_points():
    """Values whose plans need 2^22 points or more (twiddle tables of order
    2^24): round trip and a product against GMP."""
    rng = random.Random(22)
    bits = 1 << 29
    P = FftParams(2 * bits + 1, 1)
    t, lg, k = P.plan()
    assert lg >= 22
    x, y = _r(rng, bits), _r(rng, bits)
    X, Y = P.value(x), P.value(y)
    assert X.inverse() == x
    Z = P.value().mul(X, Y)
    from rforest_amd.gmp import rmatrix_mult_schoolbook
    assert Z.inverse() == rmatrix_mult_schoolbook([x], 1, [y], 1)[0]
