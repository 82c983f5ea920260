"""CPU-side checks of librforest_hip.so: it loads, exports every function the
public headers declare, and its host-only parts (prime set, roots, planner,
error returns before setup) behave as documented.  No GPU work here."""
import ctypes
import math
import os
import re

import pytest

import rforest_amd as rf
from helpers import is_prime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("rforest_hip.h", "rforest_hip_stages.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", src, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof", "defined"):
                names.add(name)
    return sorted(names)


def test_library_loads_and_exports_every_declared_symbol():
    L = rf.lib()
    decl = declared_functions()
    assert "mpz_rmatrix_mult_fft" in decl and "rf_batch_mult_device" in decl and "rf_stage_fwd" in decl
    for name in decl:
        assert hasattr(L, name), f"{name} declared in include/ but not exported"
    for var in ("hw_disable_fft", "mpzfft_threads"):
        ctypes.c_int.in_dll(L, var)


def test_prime_set():
    ps = rf.primes(32)
    assert len(set(ps)) == 32
    assert ps == sorted(ps, reverse=True)
    for p in ps:
        assert (1 << 47) < p < (1 << 51) // 9
        assert p % (1 << 24) == 1
        assert is_prime(p)


def test_roots_match_oracle_rule():
    import oracle
    ps = rf.primes(32)
    for q in (0, 1, 7, 31):
        for lgN in (1, 2, 9, 13, 14, 24):
            w = rf.root(q, lgN)
            assert w == oracle.root(ps[q], 24, lgN)
            assert pow(w, 1 << (lgN - 1), ps[q]) == ps[q] - 1


def _bound_ok(t, lgN, k, la, lb, d):
    b = 64 * t
    na, nb = -(-la // b), -(-lb // b)
    P = 1
    for p in rf.primes(k):
        P *= p
    return na + nb - 1 <= (1 << lgN) and 2 * d * min(na, nb) * ((1 << b) - 1) ** 2 < P


@pytest.mark.parametrize("la,lb,r,d", [(1 << 20, 1 << 20, 2, 2), (1 << 20, 1 << 20, 4, 4), (4096, 4096, 2, 2),
                                       (1 << 10, 1 << 10, 2, 2), (1 << 18, 1 << 17, 2, 2), (64, 64, 1, 1),
                                       (4 << 20, 1 << 20, 3, 3), (1, 1, 1, 1), (3000, 50000, 2, 5),
                                       (1 << 24, 1 << 24, 4, 4), (5 << 20, 5 << 20, 2, 2), (1 << 25, 1 << 24, 2, 2)])
def test_planner_bound_is_proven_and_minimal(la, lb, r, d):
    t, lgN, k = rf.plan(la, lb, r, d)
    assert 1 <= t <= 8 and 1 <= lgN <= 17
    assert _bound_ok(t, lgN, k, la, lb, d)
    if k > 1:  # one prime fewer must not satisfy the exact bound
        b = 64 * t
        na, nb = -(-la // b), -(-lb // b)
        P = 1
        for p in rf.primes(k - 1):
            P *= p
        assert not 2 * d * min(na, nb) * ((1 << b) - 1) ** 2 < P


def test_planner_headline_plan():
    # 2x2 of 1,048,576-bit entries: 256-bit chunks, 4096 chunks, 2^13 points
    from rforest_amd import tune
    tune.clear()
    assert rf.plan(1 << 20, 1 << 20, 2, 2) == (4, 13, 11)


def test_tuning_rules_steer_the_planner(tmp_path):
    from rforest_amd import tune
    tune.clear()
    base = rf.plan(3000, 3000, 2, 2)
    tune.force(2, 8)
    t, lgN, k = rf.plan(3000, 3000, 2, 2)
    assert t == 8 and _bound_ok(t, lgN, k, 3000, 3000, 2)
    assert rf.plan(3000, 3000, 3, 3) == rf.plan(3000, 3000, 3, 3)
    tune.clear()
    assert rf.plan(3000, 3000, 2, 2) == base
    # the cheapest measured valid plan wins; unmeasured plans do not compete
    tune.set_rule(2, 13, 4, 0, 5.0)   # 2^20-bit entries at 256-bit chunks
    tune.set_rule(2, 14, 3, 0, 4.0)   # ... at 192-bit chunks
    assert rf.plan(1 << 20, 1 << 20, 2, 2)[:2] == (3, 14)
    tune.set_rule(2, 14, 3, 0, 6.0)
    assert rf.plan(1 << 20, 1 << 20, 2, 2)[:2] == (4, 13)
    tune.clear()
    f = tmp_path / "rules.txt"
    f.write_text("# comment\n2 12 2 0 1.5\n2 12 3 0 9.0\nbad line\n")
    assert tune.load(str(f)) == 2
    assert rf.plan(3000 * 64, 3000 * 64, 2, 2)[0] in (2, 3)
    tune.clear()
    with pytest.raises(rf.RfError):
        tune.set_rule(2, 11, 9, 0, 1.0)


def test_long_transform_plans():
    # entries beyond 4 Mbit need more than 2^14 points: column + row passes
    from rforest_amd import tune
    tune.clear()
    for la, r in [(1 << 24, 2), (1 << 24, 4), (5 << 20, 2)]:
        t, lgN, k = rf.plan(la, la, r, r)
        assert 14 < lgN <= 17 and _bound_ok(t, lgN, k, la, la, r)


def test_unsupported_sizes_are_reported():
    # 2^33-bit entries would need 2^25 points at the widest (512-bit) chunks
    with pytest.raises(rf.RfError) as e:
        rf.plan(1 << 33, 1 << 33, 2, 2)
    assert e.value.code == rf.RF_ERR_UNSUPPORTED


def test_longest_transform_plans():
    """Entries up to 2^32 bits plan within 2^24 points, k <= 24 primes; the
    spectrum strides d k 2^lgN stay 32-bit."""
    for bits in (1 << 26, 3 << 26, 1 << 28, 1 << 29, 1 << 30, 1 << 31, 1 << 32):
        t, lgN, k = rf.plan(bits, bits, 2, 2)
        assert 17 < lgN <= 24 and k <= 24 and _bound_ok(t, lgN, k, bits, bits, 2)
        assert (2 * k) << lgN < 1 << 32
    assert rf.plan(1 << 30, 1 << 30, 1, 1)[1] >= 22
    assert rf.plan(1 << 32, 1 << 32, 1, 1)[1] == 24


def test_gmp_schoolbook_matches_python_integers():
    import random
    from helpers import naive_py
    from rforest_amd.gmp import rmatrix_mult_schoolbook
    rng = random.Random(9)
    for bits, r, d in [(1, 1, 1), (64, 2, 2), (5000, 3, 2), (100000, 2, 4)]:
        A = [rng.getrandbits(bits) * rng.choice([-1, 1]) for _ in range(r * d)]
        B = [rng.getrandbits(bits) * rng.choice([-1, 1]) for _ in range(d * d)]
        A[0] = 0
        assert rmatrix_mult_schoolbook(A, r, B, d) == naive_py(A, r, B, d)


def test_out_limbs_formula():
    assert rf.out_limbs(16384, 16384, 2) == (64 * 32768 + 1 + 1 + 63) // 64
    assert rf.out_limbs(0, 5, 2) == 0
    assert rf.out_limbs(1, 1, 1) == 3


def test_calls_before_setup_return_not_init():
    import numpy as np
    L = rf.lib()
    z = np.zeros(4, dtype=np.uint64)
    zs = np.zeros(4, dtype=np.uintp)
    zi = np.zeros(4, dtype=np.intc)
    zl = np.zeros(4, dtype=np.int_)
    u64, sz, ip, lp = rf._u64p, rf._szp, rf._intp, rf._longp
    rc = L.rf_rmatrix_mult_limbs(z.ctypes.data_as(u64), zs.ctypes.data_as(sz), zs.ctypes.data_as(sz),
                                 zl.ctypes.data_as(lp), z.ctypes.data_as(u64), zs.ctypes.data_as(sz),
                                 zs.ctypes.data_as(sz), zi.ctypes.data_as(ip), z.ctypes.data_as(u64),
                                 zs.ctypes.data_as(sz), zs.ctypes.data_as(sz), zi.ctypes.data_as(ip), 1, 1)
    if rc == rf.RF_OK:
        pytest.skip("library already initialised in this process")
    assert rc == rf.RF_ERR_NOT_INIT
    assert b"hw_mpz_setup" in L.rf_last_error()


@pytest.mark.parametrize("bits,terms", [(1, 1), (64, 1), (4096, 3), (1 << 20, 1), (1 << 21, 8), (3 << 22, 2)])
def test_general_transform_plan_bound_is_proven(bits, terms):
    """rf_fft_params_create: the plan's points cover any x*y with
    |x y| < 2^bits, and terms * ceil(points/2) * (2^b - 1)^2 * 2 < P_k with
    k minimal (exact integers; src/lib/mpnfft.c:18-78 states the same
    conditions with a floating-point estimate)."""
    from rforest_amd.fft import FftParams
    P = FftParams(bits, terms)
    t, lg, k = P.plan()
    P.close()
    b = 64 * t
    points = -(-(bits + 1) // b) + 1
    assert points <= 1 << lg
    ps = rf.primes(32)
    f = 2 * terms * ((points + 1) // 2) * (2 ** b - 1) ** 2
    assert f < math.prod(ps[:k])
    assert k == 1 or f >= math.prod(ps[:k - 1])
    # the extreme operands: bits(x) + bits(y) = bits + 1 chunks still fit
    for bx in (1, bits // 2 + 1, bits):
        by = bits + 1 - bx
        assert -(-bx // b) + -(-max(by, 1) // b) - 1 <= points


def test_general_transform_plan_rejects_what_no_plan_holds():
    from rforest_amd.fft import FftParams
    with pytest.raises(rf.RfError):
        FftParams(1 << 34, 1)
    assert FftParams(1 << 31, 1).plan()[1] >= 22
    with pytest.raises(rf.RfError):
        FftParams(0, 1)
