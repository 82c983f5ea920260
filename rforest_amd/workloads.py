"""Seeded synthetic inputs in GMP's own generator (gmp_randinit_default +
gmp_randseed_ui, as src/benchmark/bench_matmul.c:58-60 seeds it).

Conventions: A row-major, then B; "L-bit uniform" = mpz_urandomb(x, L);
"top bit set" = mpz_setbit(x, L - 1); "random sign" = one further
mpz_urandomb(t, 1), negate if set; "log-uniform in [2^a, 2^b]" =
floor(2^(a + u (b - a))) with u = mpz_urandomb(53) / 2^53.
"""
import ctypes
import math

from .gmp import MPZ, libgmp


class GmpRandom:
    def __init__(self, seed):
        import ctypes.util  # noqa: F401
        self._g = ctypes.CDLL(__import__("rforest_amd.gmp", fromlist=["GMP_LIB"]).GMP_LIB)
        self.state = ctypes.create_string_buffer(128)  # gmp_randstate_t (32 bytes on x86-64)
        getattr(self._g, "__gmp_randinit_default")(self.state)
        getattr(self._g, "__gmp_randseed_ui")(self.state, ctypes.c_ulong(seed))
        self.z = MPZ()
        libgmp()["init"](ctypes.byref(self.z))
        self._urandomb = getattr(self._g, "__gmpz_urandomb")
        self._urandomb.argtypes = [ctypes.POINTER(MPZ), ctypes.c_void_p, ctypes.c_ulong]
        self._setbit = getattr(self._g, "__gmpz_setbit")
        self._setbit.argtypes = [ctypes.POINTER(MPZ), ctypes.c_ulong]

    def _value(self):
        n = abs(self.z.size)
        if n == 0:
            return 0
        return int.from_bytes(ctypes.string_at(self.z.d, 8 * n), "little")

    def urandomb(self, bits):
        self._urandomb(ctypes.byref(self.z), self.state, bits)
        return self._value()

    def top(self, bits):
        """Uniform bits-bit integer with its top bit set."""
        self._urandomb(ctypes.byref(self.z), self.state, bits)
        self._setbit(ctypes.byref(self.z), bits - 1)
        return self._value()

    def signed(self, bits):
        x = self.urandomb(bits)
        return -x if self.urandomb(1) else x

    def uniform(self):
        return self.urandomb(53) / float(1 << 53)

    def log_uniform(self, a, b):
        return int(math.floor(2.0 ** (a + self.uniform() * (b - a))))

    def close(self):
        if self.state is not None:
            getattr(self._g, "__gmp_randclear")(self.state)
            libgmp()["clear"](ctypes.byref(self.z))
            self.state = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def topbit_product(rng, r, d, bits):
    """r x d times d x d, unsigned entries of exactly `bits` bits."""
    A = [rng.top(bits) for _ in range(r * d)]
    B = [rng.top(bits) for _ in range(d * d)]
    return (r, d, A, B)


def signed_product(rng, r, d, bits):
    A = [rng.signed(bits) for _ in range(r * d)]
    B = [rng.signed(bits) for _ in range(d * d)]
    return (r, d, A, B)


def headline(n, seed=42, bits=1 << 20):
    """n products of the headline workload: 2x2 . 2x2, 1,048,576-bit entries
    with the top bit set, one generator seeded `seed`."""
    rng = GmpRandom(seed)
    out = [topbit_product(rng, 2, 2, bits) for _ in range(n)]
    rng.close()
    return out


def cpu_reference_case(seed=1001):
    """2x2 . 2x2, 4096-bit entries with random signs."""
    rng = GmpRandom(seed)
    out = signed_product(rng, 2, 2, 4096)
    rng.close()
    return out


def plain_gpu_case(seed=1002, bits=1 << 20):
    """4x4 . 4x4, 1,048,576-bit entries with the top bit set."""
    rng = GmpRandom(seed)
    out = topbit_product(rng, 4, 4, bits)
    rng.close()
    return out


def tree_level(n=1024, seed=1003, lo=10, hi=17):
    """n products 2x2 . 2x2: per product L0 log-uniform in [2^lo, 2^hi], each
    entry floor(L0 2^u) bits (all within 2x), random signs."""
    rng = GmpRandom(seed)
    out = []
    for _ in range(n):
        L0 = rng.log_uniform(lo, hi)
        ents = []
        for _ in range(8):
            L = max(1, int(math.floor(L0 * 2.0 ** rng.uniform())))
            ents.append(rng.signed(L))
        out.append((2, 2, ents[:4], ents[4:]))
    rng.close()
    return out


def resident_case(seed=1004, left_bits=1 << 22, right_bits=1 << 20, n_right=16, r=3):
    """One r x r left matrix of uniform left_bits-bit entries and n_right
    right r x r matrices of uniform right_bits-bit entries, one generator."""
    rng = GmpRandom(seed)
    A = [rng.urandomb(left_bits) for _ in range(r * r)]
    Bs = [[rng.urandomb(right_bits) for _ in range(r * r)] for _ in range(n_right)]
    rng.close()
    return A, Bs


def served_stream(n=20000, seed=1005, lo=12, hi=24):
    """The mixed request stream: per request r in {2, 3, 4} with
    probabilities {0.6, 0.3, 0.1}, entry bit length L log-uniform in
    [2^lo, 2^hi] (all 2 r^2 entries of L bits), random signs, an
    exponential inter-arrival gap of mean 1 (scale by 1/rate), and its own
    seed seed * 2^20 + id for the synthetic entry stream.  Returns a list of
    (r, bits, signs, request_seed, gap)."""
    rng = GmpRandom(seed)
    out = []
    for i in range(n):
        u = rng.uniform()
        r = 2 if u < 0.6 else (3 if u < 0.9 else 4)
        L = rng.log_uniform(lo, hi)
        ents = 2 * r * r
        signs = [(-1 if rng.urandomb(1) else 1) for _ in range(ents)]
        gap = -math.log(1.0 - rng.uniform())
        out.append((r, [L] * ents, signs, seed * (1 << 20) + i, gap))
    rng.close()
    return out


def remainder_level(n=256, seed=1007, dbits=1 << 20, xbits=1 << 21):
    """n remainders of a remainder-tree level: D with dbits bits (top bit
    set), X uniform of xbits bits with a random sign, one generator seeded
    `seed` (D then X per pair)."""
    rng = GmpRandom(seed)
    out = []
    for _ in range(n):
        d = rng.top(dbits)
        x = rng.urandomb(xbits) * (-1 if rng.urandomb(1) else 1)
        out.append((d, x))
    rng.close()
    return out
