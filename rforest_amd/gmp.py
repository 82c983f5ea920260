"""mpz_t arrays for calling the drop-in C ABI (mpz_rmatrix_mult_fft) from
Python through GMP's own library, the one librforest_hip.so links."""
import ctypes
import os

GMP_LIB = os.environ.get("RF_GMP_LIB", "libgmp.so.10")


class MPZ(ctypes.Structure):
    # struct __mpz_struct (gmp.h): _mp_alloc, _mp_size, _mp_d
    _fields_ = [("alloc", ctypes.c_int), ("size", ctypes.c_int), ("d", ctypes.POINTER(ctypes.c_uint64))]


_g = None


def libgmp():
    global _g
    if _g is None:
        g = ctypes.CDLL(GMP_LIB)
        # mpz_init & co. are macros for the __gmpz_* symbols (gmp.h)
        _g = {
            "init": getattr(g, "__gmpz_init"),
            "clear": getattr(g, "__gmpz_clear"),
            "import": getattr(g, "__gmpz_import"),
            "neg": getattr(g, "__gmpz_neg"),
            "mul": getattr(g, "__gmpz_mul"),
            "addmul": getattr(g, "__gmpz_addmul"),
            "mod": getattr(g, "__gmpz_mod"),
        }
        _g["init"].argtypes = [ctypes.POINTER(MPZ)]
        _g["clear"].argtypes = [ctypes.POINTER(MPZ)]
        _g["import"].argtypes = [ctypes.POINTER(MPZ), ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t,
                                 ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p]
        _g["neg"].argtypes = [ctypes.POINTER(MPZ), ctypes.POINTER(MPZ)]
        for f in ("mul", "addmul", "mod"):
            _g[f].argtypes = [ctypes.POINTER(MPZ), ctypes.POINTER(MPZ), ctypes.POINTER(MPZ)]
    return _g


class MpzArray:
    """n initialised mpz_t in one contiguous C array (a mpz_t* for the ABI)."""

    def __init__(self, values):
        g = libgmp()
        self.n = len(values)
        self.arr = (MPZ * max(self.n, 1))()
        for i in range(self.n):
            g["init"](ctypes.byref(self.arr[i]))
        if values is not None:
            self.set(values)

    @classmethod
    def zeros(cls, n):
        return cls([0] * n)

    def set(self, values):
        g = libgmp()
        for i, v in enumerate(values):
            v = int(v)
            m = abs(v)
            nb = (m.bit_length() + 7) // 8
            buf = m.to_bytes(max(nb, 1), "little")
            g["import"](ctypes.byref(self.arr[i]), len(buf), -1, 1, 0, 0, buf)
            if v < 0:
                g["neg"](ctypes.byref(self.arr[i]), ctypes.byref(self.arr[i]))

    def values(self):
        out = []
        for i in range(self.n):
            z = self.arr[i]
            n = abs(z.size)
            if n == 0:
                out.append(0)
                continue
            raw = ctypes.string_at(z.d, 8 * n)
            v = int.from_bytes(raw, "little")
            out.append(-v if z.size < 0 else v)
        return out

    def ptr(self):
        return ctypes.cast(self.arr, ctypes.c_void_p)

    def ref(self, i):
        """Pointer to entry i (an mpz_ptr)."""
        return ctypes.byref(self.arr[i])

    def close(self):
        if self.arr is not None:
            g = libgmp()
            for i in range(self.n):
                g["clear"](ctypes.byref(self.arr[i]))
            self.arr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_fn = {}


def mpz_fn(name, argtypes, restype=None):
    """A GMP mpz function (gmp.h's mpz_<name>, symbol __gmpz_<name>) with
    the given argument types; for arithmetic on entries too large for
    Python integers."""
    f = _fn.get(name)
    if f is None:
        f = getattr(ctypes.CDLL(GMP_LIB), "__gmpz_" + name)
        f.argtypes = argtypes
        f.restype = restype
        _fn[name] = f
    return f


def rmatrix_mult_schoolbook(A, r, B, d):
    """C = A (r x d) * B (d x d) with GMP mpz_mul / mpz_addmul, the schoolbook
    product of the reference's mpz_rmatrix_mult_naive
    (src/include/hwmpz.h:117-124).  Lists of Python integers, row-major."""
    g = libgmp()
    mA, mB, mC = MpzArray(A), MpzArray(B), MpzArray.zeros(r * d)
    try:
        for i in range(r):
            for j in range(d):
                c = ctypes.byref(mC.arr[i * d + j])
                g["mul"](c, ctypes.byref(mA.arr[i * d]), ctypes.byref(mB.arr[j]))
                for l in range(1, d):
                    g["addmul"](c, ctypes.byref(mA.arr[i * d + l]), ctypes.byref(mB.arr[l * d + j]))
        return mC.values()
    finally:
        mA.close()
        mB.close()
        mC.close()


def mod_values(xs, ds):
    """[x mod d] (in [0, d)) through GMP's mpz_mod."""
    g = libgmp()
    mx, md, mr = MpzArray(xs), MpzArray(ds), MpzArray.zeros(len(xs))
    try:
        for i in range(len(xs)):
            g["mod"](ctypes.byref(mr.arr[i]), ctypes.byref(mx.arr[i]), ctypes.byref(md.arr[i]))
        return mr.values()
    finally:
        mx.close()
        md.close()
        mr.close()
