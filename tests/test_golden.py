"""Golden fixtures (tests/golden/, written by make_golden.py): seeded inputs
must regenerate bit for bit, and every implementation must reproduce the
stored products."""
import json
import os

import pytest

import oracle
from helpers import sha

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "products.json")))
VECS = json.load(open(os.path.join(HERE, "golden", "ntt_vectors.json")))


def _gen(c):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.gen(c["seed"], c["r"], c["d"], c["bits"], c["kind"])


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_inputs_regenerate(c):
    A, B = _gen(c)
    assert [sha(x) for x in A + B] == c["inputs_sha256"]
    if "inputs" in c:
        assert [hex(x) for x in A + B] == c["inputs"]


@pytest.mark.parametrize("c", [c for c in CASES if c["bits"] <= 50000], ids=lambda c: c["name"])
def test_oracle_reproduces_golden(c):
    A, B = _gen(c)
    pl, _ = oracle.ref_plan(max((abs(x).bit_length() + 63) // 64 for x in A + B), c["d"])
    C = oracle.rmatrix_mult(A, c["r"], B, c["d"], pl)
    assert [sha(x) for x in C] == c["outputs_sha256"]
    if oracle.ref_available():
        assert [sha(x) for x in oracle.ref_rmatrix_mult(A, c["r"], B, c["d"])] == c["outputs_sha256"]


@pytest.mark.parametrize("v", VECS, ids=lambda v: f"p{v['p'] % 1000}-lg{v['lgN']}")
def test_ntt_vectors(v):
    assert oracle.root(v["p"], v["s"], v["lgN"]) == v["w"]
    assert oracle.ntt_fwd(v["a"], v["lgN"], v["p"], v["w"]) == v["y"]


@pytest.mark.gpu
@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_gpu_reproduces_golden(c):
    import rforest_amd as rf
    rf.hw_mpz_setup()
    A, B = _gen(c)
    C = rf.mpz_rmatrix_mult_fft(A, c["r"], B, c["d"])
    assert [sha(x) for x in C] == c["outputs_sha256"]
    assert [abs(x).bit_length() for x in C] == c["output_bits"]
