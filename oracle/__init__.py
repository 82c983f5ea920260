"""Python access to the oracle: TEST INFRASTRUCTURE ONLY.

Wraps liboracle.so (this repository's C restatement of the reference
algorithm, rf_oracle.c) and _ref/libzzref.so (the reference's own CPU path
compiled from its sources).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this package, and only to check or to
time the CPU baseline; the product library never uses it.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libzzref.so")
REF_SRC = os.environ.get("RF_REFERENCE", "/root/reference")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_szp = ctypes.POINTER(ctypes.c_size_t)
_intp = ctypes.POINTER(ctypes.c_int)
_longp = ctypes.POINTER(ctypes.c_long)

ORC_MAXP = 64


class Plan(ctypes.Structure):
    _fields_ = [("b", ctypes.c_uint), ("lgN", ctypes.c_uint), ("k", ctypes.c_uint),
                ("p", ctypes.c_uint64 * ORC_MAXP), ("w", ctypes.c_uint64 * ORC_MAXP)]


def build():
    """Builds liboracle.so, and _ref/libzzref.so where the reference sources exist."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if os.path.isdir(os.path.join(REF_SRC, "src", "lib")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", "REF=" + REF_SRC], check=True)


_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_root.restype = ctypes.c_uint64
        L.orc_root.argtypes = [ctypes.c_uint64, ctypes.c_uint, ctypes.c_uint]
        L.orc_powmod.restype = ctypes.c_uint64
        L.orc_powmod.argtypes = [ctypes.c_uint64] * 3
        L.orc_ntt_fwd.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_uint,
                                  ctypes.c_uint64, ctypes.c_uint64]
        L.orc_ntt_inv.argtypes = [_u64p, _u64p, ctypes.c_uint, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
        L.orc_split.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_uint]
        L.orc_reduce.argtypes = [_u64p, _u64p, ctypes.c_size_t, ctypes.c_uint, ctypes.c_int, ctypes.c_uint64]
        L.orc_pointwise.argtypes = [_u64p, _u64p, _u64p, ctypes.c_uint, ctypes.c_uint, ctypes.c_size_t,
                                    ctypes.c_uint64]
        L.orc_crt.argtypes = [_u64p, ctypes.c_uint, _u64p, ctypes.c_size_t, ctypes.c_uint, _u64p]
        L.orc_recompose.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint]
        L.orc_ref_params.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint, _u64p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint), _szp]
        L.orc_ref_plan.argtypes = [ctypes.POINTER(Plan), ctypes.c_size_t, ctypes.c_uint, _szp]
        L.orc_next_size.restype = ctypes.c_size_t
        L.orc_next_size.argtypes = [ctypes.c_size_t, ctypes.c_uint]
        L.orc_log2.restype = ctypes.c_uint
        L.orc_log2.argtypes = [ctypes.c_size_t]
        L.orc_rmatrix_mult.restype = ctypes.c_int
        L.orc_rmatrix_mult.argtypes = [_u64p, _szp, _szp, _longp, _u64p, _szp, _szp, _intp,
                                       _u64p, _szp, _szp, _intp, ctypes.c_uint, ctypes.c_uint,
                                       ctypes.POINTER(Plan)]
        _orc = L
    return _orc


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            raise FileNotFoundError(REF_SO + " is not built (needs the reference sources)")
        L = ctypes.CDLL(REF_SO)
        L.zzref_setup.restype = ctypes.c_int
        L.zzref_prime.restype = ctypes.c_uint64
        L.zzref_prime.argtypes = [ctypes.c_uint]
        L.zzref_root.restype = ctypes.c_uint64
        L.zzref_root.argtypes = [ctypes.c_uint, ctypes.c_uint]
        L.zzref_params.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint), _szp]
        L.zzref_next_size.restype = ctypes.c_size_t
        L.zzref_next_size.argtypes = [ctypes.c_size_t, ctypes.c_uint]
        L.zzref_split_reduce.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                         _u64p, ctypes.c_size_t, ctypes.c_uint]
        L.zzref_fft.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint]
        L.zzref_ifft.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_uint, ctypes.c_uint]
        L.zzref_crt.argtypes = [_u64p, _u64p, ctypes.c_size_t, ctypes.c_uint]
        L.zzref_recompose.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_uint,
                                      ctypes.c_uint, ctypes.c_uint]
        L.zzref_crt_recompose.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_uint, _u64p, ctypes.c_size_t,
                                          ctypes.c_uint]
        L.zzref_rmatrix_mult_limbs.restype = ctypes.c_int
        L.zzref_rmatrix_mult_limbs.argtypes = [_u64p, _szp, _szp, _longp, _u64p, _szp, _szp, _intp,
                                               _u64p, _szp, _szp, _intp, ctypes.c_int, ctypes.c_int]
        L.zzref_setup()
        _ref = L
    return _ref


# ---- helpers ----

def _arr(vals, dtype=None):
    import numpy as np
    return np.ascontiguousarray(np.array(vals, dtype=dtype or np.uint64))


def _p(a, t=_u64p):
    return a.ctypes.data_as(t)


def _pack(entries):
    import numpy as np
    offs, ns, signs, chunks = [], [], [], []
    pos = 0
    for x in entries:
        x = int(x)
        n = (abs(x).bit_length() + 63) // 64
        offs.append(pos)
        ns.append(n)
        signs.append((x > 0) - (x < 0))
        if n:
            chunks.append(np.frombuffer(abs(x).to_bytes(8 * n, "little"), dtype=np.uint64))
        pos += n
    limbs = np.concatenate(chunks + [np.zeros(1, dtype=np.uint64)])
    return (np.ascontiguousarray(limbs), _arr(offs, np.uintp), _arr(ns, np.uintp), _arr(signs, np.intc))


def _mult(fn, A, B, r, d, extra=()):
    import numpy as np
    a, ao, an, asg = _pack(A)
    b, bo, bn, bsg = _pack(B)
    cap = int(an.max()) + int(bn.max()) + 2
    coff = np.arange(r * d, dtype=np.uintp) * cap
    ccap = np.full(r * d, cap, dtype=np.uintp)
    c = np.zeros(r * d * cap + 1, dtype=np.uint64)
    cs = np.zeros(r * d, dtype=np.int_)
    rc = fn(_p(c), _p(coff, _szp), _p(ccap, _szp), _p(cs, _longp), _p(a), _p(ao, _szp), _p(an, _szp),
            _p(asg, _intp), _p(b), _p(bo, _szp), _p(bn, _szp), _p(bsg, _intp), r, d, *extra)
    if rc != 0:
        raise RuntimeError(f"oracle product failed rc={rc}")
    out = []
    for e in range(r * d):
        n = abs(int(cs[e]))
        v = int.from_bytes(c[int(coff[e]):int(coff[e]) + n].tobytes(), "little")
        out.append(-v if cs[e] < 0 else v)
    return out


def ref_rmatrix_mult(A, r, B, d):
    """The reference CPU path (src/lib/hwmpz.c:65-130) on Python ints."""
    return _mult(ref().zzref_rmatrix_mult_limbs, A, B, r, d)


def ref_plan(max_limbs, d):
    pl = Plan()
    pts = ctypes.c_size_t()
    orc().orc_ref_plan(ctypes.byref(pl), max_limbs, d, ctypes.byref(pts))
    return pl, pts.value


def make_plan(b, lgN, primes, s=24):
    """An oracle plan for chunk width b, length 2^lgN, the given primes with
    roots derived by the rule of orc_root (2-adic order s)."""
    pl = Plan()
    pl.b, pl.lgN, pl.k = b, lgN, len(primes)
    for j, p in enumerate(primes):
        pl.p[j] = p
        pl.w[j] = orc().orc_root(p, s, lgN)
    return pl


def rmatrix_mult(A, r, B, d, plan):
    """The oracle's own restatement of the six stages with the given plan."""
    return _mult(orc().orc_rmatrix_mult, A, B, r, d, (ctypes.byref(plan),))


def root(p, s, lgN):
    return int(orc().orc_root(p, s, lgN))


def ntt_fwd(x, lgN, p, w, yn=None):
    N = 1 << lgN
    xa = _arr(list(x) + [0])
    y = _arr([0] * N)
    orc().orc_ntt_fwd(_p(y), yn or N, _p(xa), len(x), lgN, p, w)
    return [int(v) for v in y[: (yn or N)]]


def ntt_inv(y, lgN, p, w, scale):
    N = 1 << lgN
    ya = _arr(y)
    a = _arr([0] * N)
    orc().orc_ntt_inv(_p(a), _p(ya), lgN, p, w, 1 if scale else 0)
    return [int(v) for v in a]


def pointwise(A, B, r, d, n, p):
    Aa, Ba = _arr(A), _arr(B)
    C = _arr([0] * (r * d * n))
    orc().orc_pointwise(_p(C), _p(Aa), _p(Ba), r, d, n, p)
    return [int(v) for v in C]


def split(limbs, rn, b):
    t = (b + 63) // 64
    up = _arr(list(limbs) + [0])
    out = _arr([0] * (rn * t))
    orc().orc_split(_p(out), rn, _p(up), len(limbs), b)
    return [int(v) for v in out]


def reduce(chunks, rn, t, sign, p):
    ca = _arr(list(chunks) + [0])
    out = _arr([0] * rn)
    orc().orc_reduce(_p(out), _p(ca), rn, t, sign, p)
    return [int(v) for v in out]


def crt(res, un, primes, s):
    ra = _arr(res)
    pa = _arr(primes)
    out = _arr([0] * (un * s))
    orc().orc_crt(_p(out), s, _p(ra), un, len(primes), _p(pa))
    return [int(v) for v in out]


def recompose(coeffs, un, s, b, rn):
    ca = _arr(list(coeffs) + [0])
    out = _arr([0] * rn)
    orc().orc_recompose(_p(out), rn, _p(ca), un, s, b)
    return [int(v) for v in out]
