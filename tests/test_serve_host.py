"""CPU checks of the request server's host-side definitions: the synthetic
entry stream and the result checksum (csrc/rf_util.h) restated in numpy,
pinned against plain Python integer arithmetic and SplitMix64's published
first output."""
import random

from rforest_amd import serve

M = (1 << 64) - 1


def mix64_int(z):
    z &= M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_mix64_known_answer():
    # SplitMix64 seeded 0: the first output is mix64(0x9E3779B97F4A7C15)
    assert mix64_int(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF
    assert int(serve._mix64([0x9E3779B97F4A7C15])[0]) == 0xE220A8397B1DCDAF


def test_gen_stream_matches_integer_definition():
    rng = random.Random(1)
    for bits in (1, 2, 63, 64, 65, 127, 128, 129, 4096, 100003):
        seed, e = rng.getrandbits(64), rng.randrange(32)
        key = mix64_int(seed ^ ((e + 1) * 0xD1B54A32D192ED03))
        n = (bits + 63) // 64
        limbs = [mix64_int(key + (i + 1) * 0x9E3779B97F4A7C15) for i in range(n)]
        tb = bits - 64 * (n - 1)
        limbs[-1] = (limbs[-1] & ((1 << tb) - 1)) | (1 << (tb - 1))
        want = sum(l << (64 * i) for i, l in enumerate(limbs))
        got = serve.gen_value(bits, 1, seed, e)
        assert got == want and got.bit_length() == bits
        assert serve.gen_value(bits, -1, seed, e) == -want


def test_checksum_matches_integer_definition():
    rng = random.Random(2)
    C = [rng.getrandbits(1000) * rng.choice([-1, 1]) for _ in range(4)] + [0]
    rn = 17
    h = 0
    for e, c in enumerate(C):
        x = c % (1 << (64 * rn))
        for i in range(rn):
            limb = (x >> (64 * i)) & M
            h = (h + mix64_int(limb + ((e << 40) + i + 1) * 0x9E3779B97F4A7C15)) & M
    assert serve.checksum_values(C, rn) == h
    C2 = list(C)
    C2[1] += 1
    assert serve.checksum_values(C2, rn) != h
