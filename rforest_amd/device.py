"""Device-resident batches: many products of one product-tree level in one
call of rf_batch_mult_device, inputs and outputs in HBM.  Device memory
comes from the library itself (cu_device_malloc), so no second GPU runtime
shares the process."""
import ctypes

from . import DevEntry, DevOut, DevProduct, RfError, RF_ERR_NOMEM, check, lib


class DeviceBuffer:
    """nbytes of device memory owned by librforest_hip.so."""

    def __init__(self, nbytes):
        L = lib()
        self.nbytes = max(int(nbytes), 8)
        self.ptr = L.cu_device_malloc(self.nbytes)
        if not self.ptr:
            raise RfError(RF_ERR_NOMEM, L.rf_last_error().decode())

    def upload(self, host_array):
        import numpy as np
        a = np.ascontiguousarray(host_array)
        assert a.nbytes <= self.nbytes
        check(lib().rf_memcpy_h2d(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes))

    def download(self, dtype, count):
        import numpy as np
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        lib().cu_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr), out.nbytes)
        return out

    def free(self):
        if self.ptr:
            lib().cu_device_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_sync():
    check(lib().rf_device_sync())


def _upload_entries(values):
    """Python ints -> (DeviceBuffer of their magnitudes, [(limb offset,
    limbs, sign)])."""
    import numpy as np
    chunks, ents, pos = [], [], 0
    for x in values:
        x = int(x)
        n = (abs(x).bit_length() + 63) // 64
        if n:
            chunks.append(np.frombuffer(abs(x).to_bytes(8 * n, "little"), dtype=np.uint64))
        ents.append((pos, n, (x > 0) - (x < 0)))
        pos += n
    host = np.concatenate(chunks) if chunks else np.zeros(1, dtype=np.uint64)
    buf = DeviceBuffer(host.nbytes)
    buf.upload(host)
    return buf, ents


class ResidentOperand:
    """A left operand (r x d, Python ints) transformed once into the NTT
    domain in device memory (rf_operand_create_device), planned for right
    operands of up to b_max_bits bits per entry.  Pass it as A in a
    DeviceBatch product, or call mult() for a host product."""

    def __init__(self, A, r, d, b_max_bits):
        self.r, self.d = r, d
        buf, ents = _upload_entries(A)
        E = (DevEntry * (r * d))()
        for e, (off, n, s) in enumerate(ents):
            E[e].limbs = buf.ptr + 8 * off
            E[e].nlimbs = n
            E[e].sign = s
        self.la_limbs = max([n for (_, n, s) in ents if s] or [0])
        h = ctypes.c_void_p()
        check(lib().rf_operand_create_device(E, r, d, int(b_max_bits), ctypes.byref(h)))
        self.handle = h
        buf.free()  # the spectra are all the operand keeps

    def plan(self):
        t, lg, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().rf_operand_plan(self.handle, ctypes.byref(t), ctypes.byref(lg), ctypes.byref(k)))
        return t.value, lg.value, k.value

    def mult(self, B):
        """C = A * B through rf_operand_mult on GMP mpz_t."""
        from .gmp import MpzArray
        mB, mC = MpzArray(B), MpzArray.zeros(self.r * self.d)
        try:
            check(lib().rf_operand_mult(self.handle, mC.ptr(), mB.ptr()))
            return mC.values()
        finally:
            mB.close()
            mC.close()

    def release(self):
        if self.handle is not None:
            lib().rf_operand_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class DeviceBatch:
    """Uploads a list of products [(r, d, A, B), ...] (Python ints; A may be a
    ResidentOperand) once; run() enqueues the whole pipeline over all of
    them; results() reads back."""

    def __init__(self, products, synthetic=None):
        import numpy as np
        self.products = products
        L = lib()
        chunks = []
        pos = 0
        plan_in = []
        if synthetic is not None:
            # entries of the synthetic stream (rf_util.h), generated in place
            gen = []
            for (r, d, bits, signs, seed) in synthetic:
                ents = []
                for e, (b, sg) in enumerate(zip(bits, signs)):
                    n = (int(b) + 63) // 64 if sg else 0
                    ents.append((pos, n, int(sg) if n else 0))
                    if n:
                        gen.append((pos, int(b), seed, e))
                    pos += n + (n % 2)
                la = max([n for (_, n, s) in ents[:r * d] if s] or [0])
                lb = max([n for (_, n, s) in ents[r * d:] if s] or [0])
                plan_in.append((r, d, ents, int(L.rf_out_limbs(la, lb, d)), None))
            self.products = [None] * len(synthetic)
            self.limbs = DeviceBuffer(8 * max(pos, 1))
            from .serve import gen_entries_device
            gen_entries_device([self.limbs.ptr + 8 * off for (off, _, _, _) in gen], [b for (_, b, _, _) in gen],
                               [sd for (_, _, sd, _) in gen], [e for (_, _, _, e) in gen])
        for (r, d, A, B) in ([] if synthetic is not None else products):
            left = A if isinstance(A, ResidentOperand) else None
            ents = []
            for x in ([] if left else list(A)) + list(B):
                x = int(x)
                n = (abs(x).bit_length() + 63) // 64
                if n:
                    chunks.append(np.frombuffer(abs(x).to_bytes(8 * n, "little"), dtype=np.uint64))
                ents.append((pos, n, (x > 0) - (x < 0)))
                pos += n
                if n % 2:  # entries start 16-byte aligned
                    chunks.append(np.zeros(1, dtype=np.uint64))
                    pos += 1
            na = 0 if left else r * d
            la = left.la_limbs if left else max([n for (_, n, s) in ents[:na] if s] or [0])
            lb = max([n for (_, n, s) in ents[na:] if s] or [0])
            plan_in.append((r, d, ents, int(L.rf_out_limbs(la, lb, d)), left))
        if synthetic is None:
            host = np.concatenate(chunks) if chunks else np.zeros(1, dtype=np.uint64)
            self.limbs = DeviceBuffer(host.nbytes)
            self.limbs.upload(host)
        total_out = sum(max(rn, 1) * r * d for (r, d, _, rn, _) in plan_in)
        self.out = DeviceBuffer(8 * max(total_out, 1))
        self.total_out = total_out
        base, obase = self.limbs.ptr, self.out.ptr
        self._keep = []
        prods = (DevProduct * len(plan_in))()
        opos = 0
        self.layout = []
        for pi, (r, d, ents, rn, left) in enumerate(plan_in):
            A = (DevEntry * (r * d))()
            B = (DevEntry * (d * d))()
            C = (DevOut * (r * d))()
            na = 0 if left else r * d
            for e, (off, n, s) in enumerate(ents):
                tgt = A[e] if e < na else B[e - na]
                tgt.limbs = base + 8 * off
                tgt.nlimbs = n
                tgt.sign = s
            cpos = []
            for e in range(r * d):
                C[e].limbs = obase + 8 * opos
                C[e].cap = max(rn, 1)
                cpos.append((opos, rn))
                opos += max(rn, 1)
            self.layout.append(cpos)
            prods[pi].r, prods[pi].d = r, d
            prods[pi].A = ctypes.cast(A, ctypes.POINTER(DevEntry))
            prods[pi].B = ctypes.cast(B, ctypes.POINTER(DevEntry))
            prods[pi].C = ctypes.cast(C, ctypes.POINTER(DevOut))
            prods[pi].left = left.handle.value if left else None
            self._keep += [A, B, C]
        self.prods = prods

        self.handle = None

    @classmethod
    def synthetic(cls, specs):
        """Products whose entries are the synthetic stream of rf_util.h,
        generated in device memory: specs = [(r, d, bits, signs, seed)], bits
        and signs per entry (A's r*d, then B's d*d)."""
        return cls(None, synthetic=specs)

    def checksums(self):
        """Per product, the checksum of its result (rf_util.h), computed on
        the device."""
        from .serve import checksums_device
        device_sync()
        ptrs, ns, ents, slots = [], [], [], []
        for pi, cpos in enumerate(self.layout):
            for e, (o, rn) in enumerate(cpos):
                if rn:
                    ptrs.append(self.out.ptr + 8 * o)
                    ns.append(rn)
                    ents.append(e)
                    slots.append(pi)
        return checksums_device(ptrs, ns, ents, slots, len(self.layout))

    def run(self, sync=True, stream=None):
        """One-shot: plans, builds task tables and runs (rf_batch_mult_device),
        on the library's stream or on `stream` (a HipStream)."""
        check(lib().rf_batch_mult_device(self.prods, len(self.products), stream.ptr if stream else None,
                                         0 if sync else 1))

    def run_noalloc(self, stream=None):
        """The same products through rf_batch_mult_device_noalloc (scratch in
        the library's fixed device array; returns after completion)."""
        check(lib().rf_batch_mult_device_noalloc(self.prods, len(self.products), stream.ptr if stream else None))

    def run_one(self, i, sync=True):
        """Product i alone, one-shot (the plain per-product path)."""
        ptr = ctypes.cast(ctypes.byref(self.prods, i * ctypes.sizeof(DevProduct)), ctypes.POINTER(DevProduct))
        check(lib().rf_batch_mult_device(ptr, 1, None, 0 if sync else 1))

    def prepare(self, lanes=1):
        """Fixes plans, task tables and scratch once (rf_batch_create_lanes):
        the products run as `lanes` overlapping pipelines."""
        if self.handle is None:
            h = ctypes.c_void_p()
            check(lib().rf_batch_create_lanes(self.prods, len(self.products), lanes, ctypes.byref(h)))
            self.handle = h
        return self

    def launch(self, graph=True, stream=None):
        """Enqueues the prepared pipeline (on `stream`, a HipStream, or the
        library's stream); returns without waiting."""
        self.prepare()
        check(lib().rf_batch_launch(self.handle, stream.ptr if stream else None, 1 if graph else 0))

    def clear_outputs(self):
        """Zeroes the output buffer (so a later launch is seen to write it)."""
        import numpy as np
        self.out.upload(np.zeros(max(self.total_out, 1), dtype=np.uint64))

    def release(self):
        if self.handle is not None:
            lib().rf_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def results(self):
        """Per product, the r*d entries of C as Python ints."""
        import numpy as np
        device_sync()
        host = self.out.download(np.uint64, max(self.total_out, 1))
        res = []
        for cpos in self.layout:
            vals = []
            for (o, rn) in cpos:
                if rn == 0:
                    vals.append(0)
                    continue
                v = int.from_bytes(host[o:o + rn].tobytes(), "little")
                if v >> (64 * rn - 1):
                    v -= 1 << (64 * rn)
                vals.append(v)
            res.append(vals)
        return res
