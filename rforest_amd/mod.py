"""Remainders X mod D on the device (rf_mod_*), the reference's remainder
step (mpzfft_mod_init / mpzfft_mod_mod, src/lib/mpzfft.c:59-100).

    M = Modulus(d, max_bits)     # D > 0, dividends |X| < 2^max_bits
    M.mod(x)                     # x mod d in [0, d), any sign of x
    mod_batch([(M1, x1), ...])   # many remainders in one call (device)
"""
import ctypes

from . import DevEntry, check, lib
from .device import DeviceBuffer, _upload_entries
from .gmp import MpzArray

_vp = ctypes.c_void_p


def _bind():
    L = lib()
    if getattr(L, "_rf_mod_bound", False):
        return L
    L.rf_mod_create.restype = ctypes.c_int
    L.rf_mod_create.argtypes = [ctypes.c_uint64, _vp, ctypes.POINTER(_vp)]
    L.rf_mod_create_batch.restype = ctypes.c_int
    L.rf_mod_create_batch.argtypes = [ctypes.POINTER(ctypes.c_uint64), _vp, ctypes.c_int, ctypes.POINTER(_vp)]
    L.rf_mod_limbs.restype = ctypes.c_uint64
    L.rf_mod_limbs.argtypes = [_vp]
    L.rf_mod_wrap_limbs.restype = ctypes.c_uint64
    L.rf_mod_wrap_limbs.argtypes = [_vp]
    L.rf_mod_mod.restype = ctypes.c_int
    L.rf_mod_mod.argtypes = [_vp, _vp, _vp]
    L.rf_mod_mod_device.restype = ctypes.c_int
    L.rf_mod_mod_device.argtypes = [ctypes.POINTER(_vp), ctypes.POINTER(DevEntry), ctypes.c_int,
                                    ctypes.POINTER(_vp)]
    L.rf_mod_destroy.restype = None
    L.rf_mod_destroy.argtypes = [_vp]
    L._rf_mod_bound = True
    return L


class Modulus:
    def __init__(self, d, max_bits, handle=None):
        if handle is None:
            L = _bind()
            m = MpzArray([int(d)])
            handle = _vp()
            try:
                check(L.rf_mod_create(int(max_bits), ctypes.cast(m.ptr(), _vp), ctypes.byref(handle)))
            finally:
                m.close()
        self.handle, self.d, self.max_bits = handle, int(d), int(max_bits)

    @classmethod
    def batch(cls, ds, max_bits):
        """rf_mod_create_batch: one Modulus per d in ds (max_bits: one value
        or one per modulus)."""
        L = _bind()
        ds = [int(d) for d in ds]
        mb = [int(max_bits)] * len(ds) if isinstance(max_bits, int) else [int(b) for b in max_bits]
        m = MpzArray(ds)
        hs = (_vp * max(len(ds), 1))()
        try:
            check(L.rf_mod_create_batch((ctypes.c_uint64 * max(len(mb), 1))(*mb), ctypes.cast(m.ptr(), _vp),
                                        len(ds), hs))
        finally:
            m.close()
        return [cls(d, b, handle=_vp(hs[i])) for i, (d, b) in enumerate(zip(ds, mb))]

    def limbs(self):
        return int(_bind().rf_mod_limbs(self.handle))

    def wrap_limbs(self):
        """W if Q D is formed modulo 2^(64 W) - 1 (cyclic product), else 0."""
        return int(_bind().rf_mod_wrap_limbs(self.handle))

    def mod(self, x):
        mx, mr = MpzArray([int(x)]), MpzArray([0])
        try:
            check(_bind().rf_mod_mod(self.handle, ctypes.cast(mr.ptr(), _vp), ctypes.cast(mx.ptr(), _vp)))
            return mr.values()[0]
        finally:
            mx.close()
            mr.close()

    def close(self):
        if self.handle is not None:
            _bind().rf_mod_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ModBatch:
    """Dividends uploaded once; run() computes every x_i mod d_i in one
    call of rf_mod_mod_device; results() reads them back."""

    def __init__(self, pairs):
        import numpy as np
        self.pairs = list(pairs)
        self.inbuf, ents = _upload_entries([x for (_, x) in self.pairs])
        self.xs = (DevEntry * max(len(self.pairs), 1))()
        for i, (off, n, s) in enumerate(ents):
            self.xs[i].limbs = self.inbuf.ptr + 8 * off
            self.xs[i].nlimbs = n
            self.xs[i].sign = s
        self.lens = [M.limbs() for (M, _) in self.pairs]
        self.offs = [int(v) for v in np.cumsum([0] + self.lens)]
        self.out = DeviceBuffer(8 * max(self.offs[-1], 1))
        self.mods = (_vp * max(len(self.pairs), 1))(*[M.handle.value for (M, _) in self.pairs])
        self.outs = (_vp * max(len(self.pairs), 1))(*[self.out.ptr + 8 * o for o in self.offs[:-1]])

    def run(self):
        check(_bind().rf_mod_mod_device(self.mods, self.xs, len(self.pairs), self.outs))

    def results(self):
        import numpy as np
        host = self.out.download(np.uint64, max(self.offs[-1], 1))
        return [int.from_bytes(host[o:o + n].tobytes(), "little") for o, n in zip(self.offs, self.lens)]

    def release(self):
        self.inbuf.free()
        self.out.free()


def mod_batch(pairs):
    """[(Modulus, x), ...] -> [x mod d, ...] through one device call."""
    b = ModBatch(pairs)
    try:
        b.run()
        return b.results()
    finally:
        b.release()
