"""CPU checks of the library's segmented host copy (rf_stage_copy_segments):
the segments are laid end to end and split into 64 KiB pieces over the copy
threads, so a piece may hold many small segments, part of a large one, or
the end of one and the start of the next.  Every byte lands where a plain
per-segment copy puts it and nothing around a destination is touched.  Host
code only: no GPU work, no hw_mpz_setup."""
import ctypes
import random

import numpy as np
import pytest

import rforest_amd as rf

PIECE = 64 << 10


def _copy(sizes, seed):
    rng = np.random.default_rng(seed)
    srcs = [rng.integers(0, 256, size=n, dtype=np.uint8) for n in sizes]
    # each destination with 64 guard bytes on both sides
    dsts = [np.full(n + 128, 0xA5, dtype=np.uint8) for n in sizes]
    k = len(sizes)
    dst_p = (ctypes.c_void_p * k)(*[d.ctypes.data + 64 for d in dsts])
    src_p = (ctypes.c_void_p * k)(*[s.ctypes.data for s in srcs])
    n_p = (ctypes.c_uint64 * k)(*sizes)
    L = rf.lib()
    L.rf_stage_copy_segments.restype = ctypes.c_int
    L.rf_stage_copy_segments.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    assert L.rf_stage_copy_segments(dst_p, src_p, n_p, k) == rf.RF_OK
    for s, d, n in zip(srcs, dsts, sizes):
        assert np.array_equal(d[64:64 + n], s)
        assert (d[:64] == 0xA5).all() and (d[64 + n:] == 0xA5).all()


@pytest.mark.parametrize("sizes", [
    [8] * 5000,                                   # thousands of segments per piece
    [PIECE - 8, 16, PIECE, PIECE + 8, 0, 3],      # segments across piece edges
    [0, 0, 5 * PIECE + 123, 0, 0],                # empty segments around a long one
    [1, 3 * PIECE, 1, 0, 2 * PIECE - 1, 7],
])
def test_segments_against_plain_copies(sizes):
    _copy(sizes, len(sizes))


def test_random_mixtures():
    r = random.Random(11)
    for t in range(20):
        sizes = [r.choice([0, 1, 8, 512, 4096, PIECE - 1, PIECE, PIECE + 1, 3 * PIECE + 5])
                 for _ in range(r.randrange(1, 200))]
        _copy(sizes, t)


def test_no_segments_and_bad_arguments():
    L = rf.lib()
    L.rf_stage_copy_segments.restype = ctypes.c_int
    L.rf_stage_copy_segments.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    assert L.rf_stage_copy_segments(None, None, None, 0) == rf.RF_OK
    assert L.rf_stage_copy_segments(None, None, None, 2) == rf.RF_ERR_ARG
    assert L.rf_stage_copy_segments(None, None, None, -1) == rf.RF_ERR_ARG
