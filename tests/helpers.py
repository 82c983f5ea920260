"""Shared test helpers (pure Python arithmetic used as independent checks)."""
import hashlib

REF_PRIMES = [0x3fffc00000000001, 0x3ffdf00000000001, 0x3ffd900000000001, 0x3ffd500000000001]


def bitrev(x, bits):
    r = 0
    for i in range(bits):
        r |= ((x >> i) & 1) << (bits - 1 - i)
    return r


def naive_dft_bitrev(a, lgN, p, w):
    """y_m = sum_i a_i w^(i rev(m)) mod p, by definition (O(N^2))."""
    N = 1 << lgN
    a = list(a) + [0] * (N - len(a))
    return [sum(a[i] * pow(w, i * bitrev(m, lgN), p) for i in range(N)) % p for m in range(N)]


def cyclic_conv(a, b, N, p):
    c = [0] * N
    for i, x in enumerate(a):
        if x:
            for j, y in enumerate(b):
                c[(i + j) % N] = (c[(i + j) % N] + x * y) % p
    return c


def crt_signed(res, primes):
    """Symmetric CRT lift into (-P/2, P/2)."""
    P = 1
    for p in primes:
        P *= p
    x = 0
    for r, p in zip(res, primes):
        m = P // p
        x = (x + r * m * pow(m, -1, p)) % P
    return x - P if x > P // 2 else x


def to_tc(x, limbs):
    return [(x % (1 << (64 * limbs)) >> (64 * i)) & ((1 << 64) - 1) for i in range(limbs)]


def from_tc(limbs):
    v = 0
    for i, l in enumerate(limbs):
        v |= int(l) << (64 * i)
    if limbs and (int(limbs[-1]) >> 63):
        v -= 1 << (64 * len(limbs))
    return v


def sha(x):
    """SHA-256 of a sign byte followed by the little-endian magnitude."""
    x = int(x)
    m = abs(x)
    body = m.to_bytes((m.bit_length() + 7) // 8, "little")
    return hashlib.sha256((b"-" if x < 0 else b"+") + body).hexdigest()


def naive(A, r, B, d):
    """Schoolbook product: Python integers for small entries, GMP mpz_mul /
    mpz_addmul for large ones (the two agree: test_abi checks it)."""
    if max((abs(x).bit_length() for x in list(A) + list(B)), default=0) > 200000:
        from rforest_amd.gmp import rmatrix_mult_schoolbook
        return rmatrix_mult_schoolbook(A, r, B, d)
    return naive_py(A, r, B, d)


def naive_py(A, r, B, d):
    return [sum(A[i * d + k] * B[k * d + j] for k in range(d)) for i in range(r) for j in range(d)]


def is_prime(n):
    if n < 2:
        return False
    for q in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True
