"""Single-stage calls into librforest_hip.so (rforest_hip_stages.h): each runs
one pipeline stage through the product kernels on host data, so that each
stage can be checked on its own."""
import ctypes

from . import _intp, _szp, _u64p, check, lib


def _a(vals, dtype=None):
    import numpy as np
    return np.ascontiguousarray(np.array(vals, dtype=dtype or np.uint64))


def fwd(entries, t, nchunk, lgN, q, scale=0):
    """entries: list of (limbs, sign).  Returns one list of 2^lgN canonical
    spectrum values per entry."""
    import numpy as np
    limbs, offs, ns, signs = [], [], [], []
    for (lv, s) in entries:
        offs.append(len(limbs))
        ns.append(len(lv))
        signs.append(s)
        limbs += list(lv)
    la = _a(limbs + [0])
    oa, na, sa = _a(offs, np.uintp), _a(ns, np.uintp), _a(signs, np.intc)
    N = 1 << lgN
    out = _a([0] * (len(entries) * N))
    check(lib().rf_stage_fwd(la.ctypes.data_as(_u64p), oa.ctypes.data_as(_szp), na.ctypes.data_as(_szp),
                             sa.ctypes.data_as(_intp), len(entries), t, nchunk, lgN, q, scale,
                             out.ctypes.data_as(_u64p)))
    return [[int(v) for v in out[e * N:(e + 1) * N]] for e in range(len(entries))]


def inv(specA, specB, r, d, lgN, q):
    """specA: r*d lists, specB: d*d lists of 2^lgN values.  Returns r*d
    lists of N * c_i (canonical)."""
    N = 1 << lgN
    a = _a([v for s in specA for v in s])
    b = _a([v for s in specB for v in s])
    out = _a([0] * (r * d * N))
    check(lib().rf_stage_inv(a.ctypes.data_as(_u64p), b.ctypes.data_as(_u64p), r, d, lgN, q,
                             out.ctypes.data_as(_u64p)))
    return [[int(v) for v in out[e * N:(e + 1) * N]] for e in range(r * d)]


def fwd_array(limbs, sign, t, nchunk, lgN, q, scale=0):
    """One entry given as a numpy uint64 limb array; returns its 2^lgN
    canonical spectrum values as a numpy uint64 array (for transforms too
    long for Python lists)."""
    import numpy as np
    la = np.ascontiguousarray(np.concatenate([np.asarray(limbs, dtype=np.uint64), np.zeros(1, np.uint64)]))
    oa, na, sa = _a([0], np.uintp), _a([len(limbs)], np.uintp), _a([sign], np.intc)
    out = np.zeros(1 << lgN, dtype=np.uint64)
    check(lib().rf_stage_fwd(la.ctypes.data_as(_u64p), oa.ctypes.data_as(_szp), na.ctypes.data_as(_szp),
                             sa.ctypes.data_as(_intp), 1, t, nchunk, lgN, q, scale, out.ctypes.data_as(_u64p)))
    return out


def inv_array(specA, specB, r, d, lgN, q):
    """inv() on numpy uint64 arrays: specA holds r*d spectra, specB d*d, each
    of 2^lgN values, back to back; returns the r*d outputs the same way."""
    import numpy as np
    a = np.ascontiguousarray(specA, dtype=np.uint64)
    b = np.ascontiguousarray(specB, dtype=np.uint64)
    out = np.zeros(r * d << lgN, dtype=np.uint64)
    check(lib().rf_stage_inv(a.ctypes.data_as(_u64p), b.ctypes.data_as(_u64p), r, d, lgN, q,
                             out.ctypes.data_as(_u64p)))
    return out


def crt(res, nout, k, t, rn):
    """res: k*nout residues (prime-major).  Returns rn limbs (two's complement)."""
    a = _a(res)
    out = _a([0] * rn)
    check(lib().rf_stage_crt(a.ctypes.data_as(_u64p), nout, k, t, rn, out.ctypes.data_as(_u64p)))
    return [int(v) for v in out]
