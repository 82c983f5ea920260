"""Writes the golden fixtures of this directory.

products.json: seeded product cases (GMP generator, conventions of
rforest_amd/workloads.py) with the SHA-256 of every input and output entry;
outputs are computed with Python integers (schoolbook), independently of
every implementation under test, and cross-checked against the reference
CPU path when oracle/_ref is built.

ntt_vectors.json: tiny transforms by the definition y_m = sum a_i w^(i rev m)
for one reference prime (src/include/fft62.h:24) and one library prime.

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from helpers import naive, naive_dft_bitrev, sha  # noqa: E402
from rforest_amd import workloads  # noqa: E402

CASES = [
    # name, seed, r, d, bits, kind
    ("cpu_reference_2x2_4096_signed", 1001, 2, 2, 4096, "signed"),
    ("headline_2x2_1mbit_topbit", 42, 2, 2, 1 << 20, "top"),
    ("plain_gpu_4x4_1mbit_topbit", 1002, 4, 4, 1 << 20, "top"),
    ("rect_3x2_50000_signed", 7, 3, 2, 50000, "signed"),
    ("tiny_1x1_64_signed", 8, 1, 1, 64, "signed"),
]


def gen(seed, r, d, bits, kind):
    rng = workloads.GmpRandom(seed)
    f = rng.signed if kind == "signed" else rng.top
    A = [f(bits) for _ in range(r * d)]
    B = [f(bits) for _ in range(d * d)]
    rng.close()
    return A, B


def main():
    out = []
    for (name, seed, r, d, bits, kind) in CASES:
        A, B = gen(seed, r, d, bits, kind)
        C = naive(A, r, B, d)
        rec = {"name": name, "seed": seed, "r": r, "d": d, "bits": bits, "kind": kind,
               "inputs_sha256": [sha(x) for x in A + B], "outputs_sha256": [sha(x) for x in C],
               "output_bits": [abs(x).bit_length() for x in C], "output_signs": [(x > 0) - (x < 0) for x in C]}
        if bits <= 4096:
            rec["inputs"] = [hex(x) for x in A + B]
            rec["outputs"] = [hex(x) for x in C]
        try:
            import oracle
            if oracle.ref_available():
                assert oracle.ref_rmatrix_mult(A, r, B, d) == C, name
                rec["reference_cpu_path_agrees"] = True
        except ImportError:
            pass
        out.append(rec)
    with open(os.path.join(HERE, "products.json"), "w") as f:
        json.dump(out, f, indent=1)
    vec = []
    for (p, s) in [(0x3fffc00000000001, 44), (281474305622017, 24)]:
        import oracle
        for lgN in (3, 4):
            w = oracle.root(p, s, lgN)
            a = [(i * 0x9E3779B97F4A7C15 + 12345) % p for i in range(1 << lgN)]
            vec.append({"p": p, "s": s, "lgN": lgN, "w": w, "a": a, "y": naive_dft_bitrev(a, lgN, p, w)})
    with open(os.path.join(HERE, "ntt_vectors.json"), "w") as f:
        json.dump(vec, f, indent=1)


if __name__ == "__main__":
    main()
