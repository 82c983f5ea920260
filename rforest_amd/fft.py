"""Integers held in the transform domain on the device: the general transform
API of librforest_hip.so (rf_fft_*), mirroring the reference's mpzfft
interface (src/include/mpzfft.h:13-120, src/lib/mpzfft.c:10-55).

    P = FftParams(bits, terms)          # mpzfft_params_init
    X, Y, Z = P.value(x), P.value(y), P.value()
    Z.mul(X, Y); Z.add(Z, X)            # mpzfft_mul / mpzfft_add
    Z.inverse()                         # mpzfft_ifft -> Python int

Any sum of at most `terms` products x*y with |x1*y1| + ... < 2^bits comes
back exactly; larger results are unspecified, as in the reference.
"""
import ctypes

from . import DevEntry, RfError, check, lib
from .gmp import MpzArray

_vp = ctypes.c_void_p


def _bind():
    L = lib()
    if getattr(L, "_rf_fft_bound", False):
        return L
    L.rf_fft_params_create.restype = ctypes.c_int
    L.rf_fft_params_create.argtypes = [ctypes.c_uint64, ctypes.c_uint, ctypes.POINTER(_vp)]
    L.rf_fft_params_plan.restype = ctypes.c_int
    L.rf_fft_params_plan.argtypes = [_vp] + [ctypes.POINTER(ctypes.c_int)] * 3
    L.rf_fft_params_destroy.restype = None
    L.rf_fft_params_destroy.argtypes = [_vp]
    L.rf_fft_create.restype = ctypes.c_int
    L.rf_fft_create.argtypes = [_vp, ctypes.POINTER(_vp)]
    L.rf_fft_destroy.restype = None
    L.rf_fft_destroy.argtypes = [_vp]
    L.rf_fft_forward.restype = ctypes.c_int
    L.rf_fft_forward.argtypes = [_vp, _vp]
    L.rf_fft_forward_device.restype = ctypes.c_int
    L.rf_fft_forward_device.argtypes = [_vp, ctypes.POINTER(DevEntry)]
    L.rf_fft_inverse.restype = ctypes.c_int
    L.rf_fft_inverse.argtypes = [_vp, _vp]
    L.rf_fft_inverse_device.restype = ctypes.c_int
    L.rf_fft_inverse_device.argtypes = [_vp, _vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_uint64)]
    L.rf_fft_out_limbs.restype = ctypes.c_uint64
    L.rf_fft_out_limbs.argtypes = [_vp]
    for name in ("rf_fft_set", "rf_fft_neg"):
        getattr(L, name).restype = ctypes.c_int
        getattr(L, name).argtypes = [_vp, _vp]
    for name in ("rf_fft_add", "rf_fft_sub", "rf_fft_mul"):
        getattr(L, name).restype = ctypes.c_int
        getattr(L, name).argtypes = [_vp, _vp, _vp]
    L.rf_fft_matrix_mul.restype = ctypes.c_int
    L.rf_fft_matrix_mul.argtypes = [ctypes.POINTER(_vp)] * 3 + [ctypes.c_uint] * 3
    L._rf_fft_bound = True
    return L


class FftParams:
    """A plan for sums of at most `terms` products whose |x1*y1| + ... < 2^bits."""

    def __init__(self, bits, terms=1):
        L = _bind()
        h = _vp()
        check(L.rf_fft_params_create(int(bits), int(terms), ctypes.byref(h)))
        self.handle, self.bits, self.terms = h, int(bits), int(terms)

    def plan(self):
        t, lg, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(_bind().rf_fft_params_plan(self.handle, ctypes.byref(t), ctypes.byref(lg), ctypes.byref(k)))
        return t.value, lg.value, k.value

    def value(self, x=None):
        f = Fft(self)
        if x is not None:
            f.forward(x)
        return f

    def close(self):
        if self.handle is not None:
            _bind().rf_fft_params_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Fft:
    """One integer in the transform domain (an rf_fft value, initially 0)."""

    def __init__(self, params):
        L = _bind()
        h = _vp()
        check(L.rf_fft_create(params.handle, ctypes.byref(h)))
        self.handle = h

    def forward(self, x):
        m = MpzArray([int(x)])
        try:
            check(_bind().rf_fft_forward(self.handle, ctypes.cast(m.ptr(), _vp)))
        finally:
            m.close()
        return self

    def forward_device(self, entry):
        check(_bind().rf_fft_forward_device(self.handle, ctypes.byref(entry)))
        return self

    def inverse(self):
        m = MpzArray([0])
        try:
            check(_bind().rf_fft_inverse(ctypes.cast(m.ptr(), _vp), self.handle))
            return m.values()[0]
        finally:
            m.close()

    def inverse_device(self, out_ptr, cap):
        """Into device limbs at out_ptr (cap limbs): returns (sign, nlimbs)."""
        s, n = ctypes.c_int(), ctypes.c_uint64()
        check(_bind().rf_fft_inverse_device(self.handle, _vp(out_ptr), int(cap), ctypes.byref(s), ctypes.byref(n)))
        return s.value, n.value

    def out_limbs(self):
        return int(_bind().rf_fft_out_limbs(self.handle))

    def set(self, a):
        check(_bind().rf_fft_set(self.handle, a.handle))
        return self

    def neg(self, a):
        check(_bind().rf_fft_neg(self.handle, a.handle))
        return self

    def add(self, a, b):
        check(_bind().rf_fft_add(self.handle, a.handle, b.handle))
        return self

    def sub(self, a, b):
        check(_bind().rf_fft_sub(self.handle, a.handle, b.handle))
        return self

    def mul(self, a, b):
        check(_bind().rf_fft_mul(self.handle, a.handle, b.handle))
        return self

    def close(self):
        if self.handle is not None:
            _bind().rf_fft_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def matrix_mul(rop, op1, op2, dim1, dim2, dim3):
    """rop (dim1 x dim3) := op1 (dim1 x dim2) * op2 (dim2 x dim3), row-major
    lists of Fft values (mpzfft_matrix_mul)."""
    def arr(vals):
        a = (_vp * len(vals))()
        for i, v in enumerate(vals):
            a[i] = v.handle
        return a
    check(_bind().rf_fft_matrix_mul(arr(rop), arr(op1), arr(op2), dim1, dim2, dim3))


__all__ = ["FftParams", "Fft", "matrix_mul", "RfError"]
