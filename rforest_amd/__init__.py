"""rforest_amd -- MI355X-native exact integer matrix products.

Python mirror of the reference's matrix-product interface
(src/include/hwmpz.h:74-139): ``hw_mpz_setup``, ``mpz_rmatrix_mult_fft``
(C = A B with A r x d and B d x d, row-major lists of Python ints) and
``mpz_rmatrix_gpu_fft_suitable``, all calling the C ABI of
librforest_hip.so through ctypes.  The library is required: nothing here
computes a product on the CPU (``mpz_rmatrix_mult_naive`` is the
reference's schoolbook helper, src/include/hwmpz.h:117-124, kept for
checking).

Beside it: ``device`` (device-resident batches, resident operands),
``tree_product`` (product trees in device memory), ``fft`` (the mpzfft
transform-domain interface), ``mod`` (remainders), ``serve`` (the request
server), ``tune`` (the plan autotuner), ``workloads`` (the seeded inputs of
the benchmarks) and ``stages`` (per-stage entry points for parity tests).
"""
import ctypes
import os

from . import gmp  # noqa: F401  (mpz_t helpers for the drop-in ABI)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librforest_hip.so")

RF_OK = 0
RF_ERR_NOT_INIT = -1
RF_ERR_UNSUPPORTED = -2
RF_ERR_HIP = -3
RF_ERR_ARG = -4
RF_ERR_NOMEM = -5
RF_ERR_CAPACITY = -6

STAGE_NAMES = ("fwd", "inv", "crt", "carry", "apply")

_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_szp = ctypes.POINTER(ctypes.c_size_t)
_intp = ctypes.POINTER(ctypes.c_int)
_longp = ctypes.POINTER(ctypes.c_long)


class RfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rforest_hip error {code}: {msg}")
        self.code = code


class DevEntry(ctypes.Structure):
    _fields_ = [("limbs", ctypes.c_void_p), ("nlimbs", ctypes.c_uint64),
                ("sign", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class DevOut(ctypes.Structure):
    _fields_ = [("limbs", ctypes.c_void_p), ("cap", ctypes.c_uint64)]


class DevProduct(ctypes.Structure):
    _fields_ = [("r", ctypes.c_int32), ("d", ctypes.c_int32),
                ("A", ctypes.POINTER(DevEntry)), ("B", ctypes.POINTER(DevEntry)),
                ("C", ctypes.POINTER(DevOut)), ("left", ctypes.c_void_p)]


def lib():
    """Loads librforest_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run rforest_amd/build.py (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    L.hw_mpz_setup.restype = None
    L.hw_mpz_clear.restype = None
    L.rf_last_error.restype = ctypes.c_char_p
    L.rf_rmatrix_mult_limbs.restype = ctypes.c_int
    L.rf_rmatrix_mult_limbs.argtypes = [_u64p, _szp, _szp, _longp, _u64p, _szp, _szp, _intp,
                                        _u64p, _szp, _szp, _intp, ctypes.c_int, ctypes.c_int]
    L.rf_mpz_rmatrix_mult.restype = ctypes.c_int
    L.rf_out_limbs.restype = ctypes.c_uint64
    L.rf_out_limbs.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
    L.rf_batch_mult_device.restype = ctypes.c_int
    L.rf_batch_mult_device.argtypes = [ctypes.POINTER(DevProduct), ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    L.rf_plan_query.restype = ctypes.c_int
    L.rf_plan_query.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                _intp, _intp, _intp]
    L.rf_primes.restype = ctypes.c_int
    L.rf_primes.argtypes = [_u64p, ctypes.c_int]
    L.rf_root.restype = ctypes.c_uint64
    L.rf_root.argtypes = [ctypes.c_int, ctypes.c_int]
    L.rf_stage_fwd.restype = ctypes.c_int
    L.rf_stage_fwd.argtypes = [_u64p, _szp, _szp, _intp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int, _u64p]
    L.rf_stage_inv.restype = ctypes.c_int
    L.rf_stage_inv.argtypes = [_u64p, _u64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u64p]
    L.rf_stage_crt.restype = ctypes.c_int
    L.rf_stage_crt.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _u64p]
    L.rf_profile_enable.argtypes = [ctypes.c_int]
    L.rf_profile_read.argtypes = [ctypes.POINTER(ctypes.c_double), _u64p, ctypes.c_int]
    L.mpz_rmatrix_gpu_fft_suitable.restype = ctypes.c_int
    L.cu_device_malloc.restype = ctypes.c_void_p
    L.cu_device_malloc.argtypes = [ctypes.c_size_t]
    L.cu_device_free.restype = None
    L.cu_device_free.argtypes = [ctypes.c_void_p]
    L.cu_malloc_pinned.restype = ctypes.c_void_p
    L.cu_malloc_pinned.argtypes = [ctypes.c_size_t]
    L.cu_free_pinned.restype = None
    L.cu_free_pinned.argtypes = [ctypes.c_void_p]
    L.cu_memcpy_d2h.restype = None
    L.cu_memcpy_d2h.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    L.rf_memcpy_h2d.restype = ctypes.c_int
    L.rf_memcpy_h2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    L.rf_device_sync.restype = ctypes.c_int
    L.rf_memcpy_d2d.restype = ctypes.c_int
    L.rf_memcpy_d2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    L.rf_batch_create_lanes.restype = ctypes.c_int
    L.rf_batch_create_lanes.argtypes = [ctypes.POINTER(DevProduct), ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]
    L.rf_batch_create.restype = ctypes.c_int
    L.rf_batch_create.argtypes = [ctypes.POINTER(DevProduct), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.rf_batch_launch.restype = ctypes.c_int
    L.rf_batch_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    L.rf_batch_destroy.restype = None
    L.rf_batch_destroy.argtypes = [ctypes.c_void_p]
    L.rf_operand_create_device.restype = ctypes.c_int
    L.rf_operand_create_device.argtypes = [ctypes.POINTER(DevEntry), ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                           ctypes.POINTER(ctypes.c_void_p)]
    L.rf_operand_create.restype = ctypes.c_int
    L.rf_operand_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                    ctypes.POINTER(ctypes.c_void_p)]
    L.rf_operand_mult.restype = ctypes.c_int
    L.rf_operand_mult.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.rf_operand_plan.restype = ctypes.c_int
    L.rf_operand_plan.argtypes = [ctypes.c_void_p, _intp, _intp, _intp]
    L.rf_operand_destroy.restype = None
    L.rf_operand_destroy.argtypes = [ctypes.c_void_p]
    L.rf_batch_mult_device_noalloc.restype = ctypes.c_int
    L.rf_batch_mult_device_noalloc.argtypes = [ctypes.POINTER(DevProduct), ctypes.c_int, ctypes.c_void_p]
    L.rf_stage_dropin_piece_bits.restype = ctypes.c_int
    L.rf_stage_dropin_piece_bits.argtypes = [ctypes.c_uint64]
    L.mpz_rmatrix_mult_fft.restype = None
    L.mpz_rmatrix_mult_fft.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.c_void_p]
    L.rf_mpz_rmatrix_mult.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_int]
    _lib = L
    return L


def hip_runtime():
    """The HIP runtime librforest_hip.so runs on (libamdhip64.so.7, loaded by
    the library itself), for callers that pass streams of their own."""
    lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipStreamCreateWithFlags.restype = ctypes.c_int
    h.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
    h.hipStreamSynchronize.restype = ctypes.c_int
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamDestroy.restype = ctypes.c_int
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return h


class HipStream:
    """A caller-owned HIP stream (non-blocking, default priority) of the
    library's runtime; .ptr is the hipStream_t handed to the C ABI."""

    def __init__(self):
        self._h = hip_runtime()
        s = ctypes.c_void_p()
        if self._h.hipStreamCreateWithFlags(ctypes.byref(s), 1) != 0:  # hipStreamNonBlocking
            raise RfError(RF_ERR_HIP, "hipStreamCreateWithFlags failed")
        self.ptr = s.value

    def synchronize(self):
        if self._h.hipStreamSynchronize(self.ptr) != 0:
            raise RfError(RF_ERR_HIP, "hipStreamSynchronize failed")

    def destroy(self):
        if self.ptr:
            self._h.hipStreamDestroy(self.ptr)
            self.ptr = None


def check(rc):
    if rc != RF_OK:
        raise RfError(rc, lib().rf_last_error().decode())
    return rc


def hw_mpz_setup():
    """src/include/hwmpz.h:74 -- prime tables and GPU state, idempotent."""
    lib().hw_mpz_setup()


def hw_mpz_clear():
    """src/include/hwmpz.h:75."""
    lib().hw_mpz_clear()


# ---- limb conversion ----

def to_limbs(x):
    """Magnitude of a Python int as little-endian 64-bit limbs."""
    x = abs(int(x))
    n = (x.bit_length() + 63) // 64
    b = x.to_bytes(n * 8, "little")
    return [int.from_bytes(b[8 * i:8 * i + 8], "little") for i in range(n)]


def from_limbs(limbs):
    """Little-endian 64-bit limbs to a non-negative Python int."""
    b = b"".join(int(v).to_bytes(8, "little") for v in limbs)
    return int.from_bytes(b, "little")


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
    limbs = np.concatenate(chunks) if chunks else np.zeros(0, dtype=np.uint64)
    limbs = np.ascontiguousarray(np.concatenate([limbs, np.zeros(1, dtype=np.uint64)]))
    return (limbs, np.array(offs, dtype=np.uintp), np.array(ns, dtype=np.uintp),
            np.array(signs, dtype=np.intc))


def _ptr(a, t):
    return a.ctypes.data_as(t)


def mpz_rmatrix_mult_fft(A, r, B, d):
    """C = A (r x d) * B (d x d), exact; lists of Python ints, row-major.

    Mirrors mpz_rmatrix_mult_fft (src/include/hwmpz.h:129) through the
    plain-array ABI rf_rmatrix_mult_limbs."""
    import numpy as np
    if len(A) != r * d or len(B) != d * d:
        raise ValueError("A must have r*d entries and B d*d entries")
    L = lib()
    a, ao, an, asg = _pack(A)
    b, bo, bn, bsg = _pack(B)
    la = int(an.max()) if len(an) else 0
    lb = int(bn.max()) if len(bn) else 0
    cap = la + lb + 2
    coff = np.arange(r * d, dtype=np.uintp) * cap
    ccap = np.full(r * d, cap, dtype=np.uintp)
    c = np.zeros(r * d * cap + 1, dtype=np.uint64)
    cs = np.zeros(r * d, dtype=np.int_)
    check(L.rf_rmatrix_mult_limbs(_ptr(c, _u64p), _ptr(coff, _szp), _ptr(ccap, _szp), _ptr(cs, _longp),
                                  _ptr(a, _u64p), _ptr(ao, _szp), _ptr(an, _szp), _ptr(asg, _intp),
                                  _ptr(b, _u64p), _ptr(bo, _szp), _ptr(bn, _szp), _ptr(bsg, _intp),
                                  r, d))
    out = []
    for e in range(r * d):
        n = abs(int(cs[e]))
        v = int.from_bytes(c[int(coff[e]):int(coff[e]) + n].tobytes(), "little")
        out.append(-v if cs[e] < 0 else v)
    return out


def rmatrix_mult_limb_arrays(A, r, B, d):
    """mpz_rmatrix_mult_fft on entries given as (sign, uint64 limb array)
    pairs (little-endian magnitudes), without Python integers: for entries
    of billions of bits.  Returns r*d (sign, limb array) pairs, the arrays
    views into one output buffer."""
    import numpy as np
    if len(A) != r * d or len(B) != d * d:
        raise ValueError("A must have r*d entries and B d*d entries")

    def pack(ents):
        arrs = [np.asarray(v, dtype=np.uint64) if s else np.zeros(0, np.uint64) for s, v in ents]
        ns = np.array([len(a) for a in arrs], dtype=np.uintp)
        offs = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.uintp)
        limbs = np.ascontiguousarray(np.concatenate(arrs + [np.zeros(1, np.uint64)]))
        return limbs, offs, ns, np.array([s for s, _ in ents], dtype=np.intc)

    a, ao, an, asg = pack(A)
    b, bo, bn, bsg = pack(B)
    cap = int(an.max()) + int(bn.max()) + 2
    coff = np.arange(r * d, dtype=np.uintp) * cap
    ccap = np.full(r * d, cap, dtype=np.uintp)
    c = np.zeros(r * d * cap + 1, dtype=np.uint64)
    cs = np.zeros(r * d, dtype=np.int_)
    check(lib().rf_rmatrix_mult_limbs(_ptr(c, _u64p), _ptr(coff, _szp), _ptr(ccap, _szp), _ptr(cs, _longp),
                                      _ptr(a, _u64p), _ptr(ao, _szp), _ptr(an, _szp), _ptr(asg, _intp),
                                      _ptr(b, _u64p), _ptr(bo, _szp), _ptr(bn, _szp), _ptr(bsg, _intp),
                                      r, d))
    return [(int(np.sign(cs[e])), c[int(coff[e]):int(coff[e]) + abs(int(cs[e]))]) for e in range(r * d)]


def mpz_matrix_mult_fft(A, B, d):
    """src/include/hwmpz.h:130 -- square case r = d."""
    return mpz_rmatrix_mult_fft(A, d, B, d)


def mpz_rmatrix_mult_naive(A, r, B, d):
    """Schoolbook product with GMP mpz_mul / mpz_addmul
    (src/include/hwmpz.h:117-124)."""
    from .gmp import rmatrix_mult_schoolbook
    return rmatrix_mult_schoolbook(A, r, B, d)


def tree_product(mats, d):
    """M_0 M_1 ... M_{n-1} of d x d matrices (lists of d*d Python ints,
    row-major) as a product tree kept in device memory
    (rf_mpz_tree_product)."""
    from .gmp import MpzArray
    flat = [x for M in mats for x in M]
    mA, mC = MpzArray(flat), MpzArray.zeros(d * d)
    try:
        L = lib()
        L.rf_mpz_tree_product.restype = ctypes.c_int
        L.rf_mpz_tree_product.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        check(L.rf_mpz_tree_product(mC.ptr(), mA.ptr(), len(mats), d))
        return mC.values()
    finally:
        mA.close()
        mC.close()


def plan(la_bits, lb_bits, r, d):
    """(t, lgN, k): chunk width 64 t bits, transform 2^lgN, k primes."""
    t, lg, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(lib().rf_plan_query(la_bits, lb_bits, r, d, ctypes.byref(t), ctypes.byref(lg), ctypes.byref(k)))
    return t.value, lg.value, k.value


def primes(n=32):
    import numpy as np
    out = np.zeros(n, dtype=np.uint64)
    lib().rf_primes(_ptr(out, _u64p), n)
    return [int(x) for x in out]


def root(q, lgN):
    return int(lib().rf_root(q, lgN))


def out_limbs(la_limbs, lb_limbs, d):
    return int(lib().rf_out_limbs(la_limbs, lb_limbs, d))


def profile(enable=None, reset=False):
    """Per-stage GPU time (ms) and launch counts measured with HIP events on
    the stream each kernel runs on."""
    import numpy as np
    L = lib()
    if enable is not None:
        L.rf_profile_enable(1 if enable else 0)
    if reset:
        L.rf_profile_reset()
        return None
    ms = (ctypes.c_double * 5)()
    n = (ctypes.c_uint64 * 5)()
    L.rf_profile_read(ms, n, 5)
    return {STAGE_NAMES[i]: (ms[i], int(n[i])) for i in range(5)}


from .device import DeviceBatch, DeviceBuffer, ResidentOperand, device_sync  # noqa: E402,F401
