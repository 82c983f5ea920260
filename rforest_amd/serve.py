"""The request server of librforest_hip.so (rf_server_*): continuous batching
of matrix-product requests under an HBM budget.

Host restatements of the synthetic entry stream and of the result checksum
(both defined in csrc/rf_util.h) live here too, so that a result computed
inside the server can be checked against other paths."""
import ctypes

import numpy as np

from . import RfError, check, lib

RF_IN_HOST, RF_IN_SYNTH = 0, 1
RF_OUT_KEEP, RF_OUT_HOST, RF_OUT_CHECKSUM = 1, 2, 4

_K_GEN = 0xD1B54A32D192ED03
_K_LIMB = 0x9E3779B97F4A7C15
_M64 = (1 << 64) - 1


class ServeEntry(ctypes.Structure):
    _fields_ = [("limbs", ctypes.c_void_p), ("bits", ctypes.c_uint64), ("sign", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [("r", ctypes.c_int32), ("d", ctypes.c_int32), ("input", ctypes.c_int32), ("output", ctypes.c_int32),
                ("A", ctypes.POINTER(ServeEntry)), ("B", ctypes.POINTER(ServeEntry)), ("seed", ctypes.c_uint64),
                ("out_host", ctypes.c_void_p), ("out_cap", ctypes.c_uint64), ("tag", ctypes.c_uint64),
                ("left", ctypes.c_void_p)]


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("status", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("out_limbs", ctypes.c_uint64), ("checksum", ctypes.c_uint64),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("streams", ctypes.c_int32), ("max_batch", ctypes.c_int32),
                ("max_batch_bytes", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("products_in_batches", ctypes.c_uint64), ("peak_bytes", ctypes.c_uint64),
                ("budget", ctypes.c_uint64)]


def _bind():
    L = lib()
    if getattr(L, "_serve_bound", False):
        return L
    L.rf_server_create.restype = ctypes.c_int
    L.rf_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(ctypes.c_void_p)]
    L.rf_server_submit.restype = ctypes.c_int
    L.rf_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
    L.rf_server_poll.restype = ctypes.c_int
    L.rf_server_poll.argtypes = [ctypes.c_void_p, ctypes.POINTER(Completion), ctypes.c_int, ctypes.c_int]
    L.rf_server_read.restype = ctypes.c_int
    L.rf_server_read.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64]
    L.rf_server_release.restype = ctypes.c_int
    L.rf_server_release.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    L.rf_server_stats_get.restype = ctypes.c_int
    L.rf_server_stats_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerStats)]
    L.rf_server_clock.restype = ctypes.c_double
    L.rf_server_clock.argtypes = [ctypes.c_void_p]
    L.rf_server_destroy.restype = ctypes.c_int
    L.rf_server_destroy.argtypes = [ctypes.c_void_p]
    L.rf_gen_entry_device.restype = ctypes.c_int
    L.rf_gen_entry_device.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]
    L.rf_checksum_device.restype = ctypes.c_int
    L.rf_checksum_device.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                     ctypes.POINTER(ctypes.c_uint64)]
    L.rf_gen_entries_device.restype = ctypes.c_int
    L.rf_gen_entries_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_int]
    L.rf_checksums_device.restype = ctypes.c_int
    L.rf_checksums_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    L._serve_bound = True
    return L


# ---- host restatements of csrc/rf_util.h ----

def _mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def gen_key(seed, entry):
    return int(_mix64(np.array([(seed ^ ((entry + 1) * _K_GEN)) & _M64], dtype=np.uint64))[0])


def gen_limbs(bits, seed, entry):
    """The synthetic magnitude of entry `entry` of a request seeded `seed`
    with exactly `bits` bits, as uint64 limbs (least significant first)."""
    n = (bits + 63) // 64
    key = np.uint64(gen_key(seed, entry))
    i = np.arange(1, n + 1, dtype=np.uint64)
    v = _mix64(key + i * np.uint64(_K_LIMB))
    tb = bits - 64 * (n - 1)
    top = int(v[-1])
    if tb < 64:
        top &= (1 << tb) - 1
    top |= 1 << (tb - 1)
    v[-1] = np.uint64(top)
    return v


def gen_value(bits, sign, seed, entry):
    if bits == 0 or sign == 0:
        return 0
    m = int.from_bytes(gen_limbs(bits, seed, entry).tobytes(), "little")
    return -m if sign < 0 else m


def checksum_limbs(limbs, entry):
    """Checksum share of C entry `entry` given its two's complement limbs."""
    limbs = np.asarray(limbs, dtype=np.uint64)
    if limbs.size == 0:
        return 0
    i = np.arange(1, limbs.size + 1, dtype=np.uint64)
    t = _mix64(limbs + (np.uint64(entry << 40) + i) * np.uint64(_K_LIMB))
    return int(np.sum(t, dtype=np.uint64))


def checksum_values(C, rn):
    """Checksum of a result (list of r*d ints) in its rn-limb two's complement form."""
    h = 0
    for e, c in enumerate(C):
        if rn == 0:
            continue
        x = int(c) % (1 << (64 * rn))
        limbs = np.frombuffer(x.to_bytes(8 * rn, "little"), dtype=np.uint64)
        h = (h + checksum_limbs(limbs, e)) & _M64
    return h


class Server:
    """A request server on the library's GPU.  submit_*() return request ids;
    poll() returns completions as dicts."""

    def __init__(self, hbm_budget, streams=3, max_batch=256, max_batch_bytes=0):
        L = _bind()
        cfg = ServerConfig(int(hbm_budget), streams, max_batch, int(max_batch_bytes))
        h = ctypes.c_void_p()
        check(L.rf_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def _submit(self, r, d, ents, input_mode, seed, output, tag, out=None, out_cap=0, keep=None, left=None):
        """ents: A's r*d entries then B's d*d as (limb pointer, bits, sign);
        with a resident left operand, B's d*d only."""
        L = _bind()
        A = (ServeEntry * (r * d))()
        B = (ServeEntry * (d * d))()
        na = 0 if left is not None else r * d
        for e, (ptr, bits, sign) in enumerate(ents):
            tgt = A[e] if e < na else B[e - na]
            tgt.limbs = ptr
            tgt.bits = bits
            tgt.sign = sign
        req = Request(r, d, input_mode, output, ctypes.cast(A, ctypes.POINTER(ServeEntry)),
                      ctypes.cast(B, ctypes.POINTER(ServeEntry)), seed & _M64,
                      out.ctypes.data if out is not None else None, out_cap, tag,
                      left.handle.value if left is not None else None)
        rid = ctypes.c_uint64()
        check(L.rf_server_submit(self.h, ctypes.byref(req), ctypes.byref(rid)))
        if keep is not None or out is not None:
            self._keep[rid.value] = (keep, out)
        return rid.value

    def submit_synth(self, r, d, bits, signs, seed, tag=0, output=RF_OUT_CHECKSUM, left=None):
        """Entries generated in device memory at admission: entry e has
        bits[e] bits and sign signs[e] (A's r*d entries, then B's d*d; B's
        only when `left`, a ResidentOperand, stands for A)."""
        ents = [(None, int(b), int(s)) for b, s in zip(bits, signs)]
        return self._submit(r, d, ents, RF_IN_SYNTH, seed, output, tag, left=left)

    def submit_host(self, r, d, A, B, tag=0, output=RF_OUT_HOST | RF_OUT_CHECKSUM, left=None):
        """Entries from Python ints, copied at admission (A is ignored when
        `left`, a ResidentOperand, stands for it).  With RF_OUT_HOST the
        result is copied back at completion (host_result(id) decodes it)."""
        limbs = []
        for x in ([] if left is not None else list(A)) + list(B):
            x = int(x)
            m = abs(x)
            n = (m.bit_length() + 63) // 64
            limbs.append(((x > 0) - (x < 0), np.frombuffer(m.to_bytes(8 * max(n, 1), "little"), dtype=np.uint64).copy(),
                          m.bit_length()))
        return self.submit_host_limbs(r, d, limbs, tag, output, left)

    def out_limbs_for(self, r, d, ents, left=None):
        """Limbs per C entry that a host result of these entries needs."""
        na = 0 if left is not None else r * d
        la = left.la_limbs * 64 if left is not None else max([b if sg else 0 for (sg, _, b) in ents[:na]] or [0])
        lb = max([b if sg else 0 for (sg, _, b) in ents[na:]] or [0])
        return max(int(lib().rf_out_limbs((la + 63) // 64, (lb + 63) // 64, d)), 1)

    def submit_host_limbs(self, r, d, ents, tag=0, output=RF_OUT_HOST | RF_OUT_CHECKSUM, left=None, out=None):
        """Entries as (sign, uint64 limb array, bits) triples, copied at
        admission; A's r*d then B's d*d (B's only with `left`).  `out`: a
        caller's uint64 array of r*d*out_limbs_for(...) limbs for the result
        (a new one when None)."""
        bufs = [a for (_, a, _) in ents]
        cents = [(a.ctypes.data, int(bits) if sg else 0, int(sg) if bits else 0) for (sg, a, bits) in ents]
        na = 0 if left is not None else r * d
        la = left.la_limbs * 64 if left is not None else max([e[1] for e in cents[:na]] or [0])
        lb = max([e[1] for e in cents[na:]] or [0])
        cap = max(int(lib().rf_out_limbs((la + 63) // 64, (lb + 63) // 64, d)), 1)
        if not output & RF_OUT_HOST:
            out = None
        elif out is None:
            out = np.zeros(r * d * cap, dtype=np.uint64)
        else:
            assert out.dtype == np.uint64 and out.size >= r * d * cap
        rid = self._submit(r, d, cents, RF_IN_HOST, 0, output, tag, out, cap if out is not None else 0, bufs, left)
        if out is not None:
            self._keep[rid] = (bufs, out, r, d, cap)
        return rid

    def host_result(self, rid, out_limbs):
        """Decodes an RF_OUT_HOST result after its completion."""
        bufs, out, r, d, cap = self._keep.pop(rid)
        res = []
        for e in range(r * d):
            if out_limbs == 0:
                res.append(0)
                continue
            v = int.from_bytes(out[e * cap:e * cap + out_limbs].tobytes(), "little")
            if v >> (64 * out_limbs - 1):
                v -= 1 << (64 * out_limbs)
            res.append(v)
        return res

    def poll(self, max_n=4096, wait_ms=0):
        L = _bind()
        if getattr(self, "_buf", None) is None or len(self._buf) < max_n:
            self._buf = (Completion * max_n)()
        buf = self._buf
        n = L.rf_server_poll(self.h, buf, max_n, int(wait_ms))
        if n < 0:
            raise RfError(n, L.rf_last_error().decode())
        out = []
        for i in range(n):
            c = buf[i]
            out.append({"id": c.id, "tag": c.tag, "status": c.status, "out_limbs": c.out_limbs,
                        "checksum": c.checksum, "t_submit": c.t_submit, "t_admit": c.t_admit, "t_done": c.t_done})
            ent = self._keep.get(c.id)
            if ent is not None and len(ent) == 2:  # inputs held only until completion
                self._keep.pop(c.id)
        return out

    def wait_all(self, n, timeout_s=600):
        """Polls until n completions arrived; returns them."""
        import time
        got, t0 = [], time.time()
        while len(got) < n:
            got += self.poll(4096, 50)
            if time.time() - t0 > timeout_s:
                raise TimeoutError(f"{len(got)} of {n} requests completed")
        return got

    def read(self, rid, r, d, out_limbs):
        """A completed RF_OUT_KEEP result as r*d Python ints."""
        L = _bind()
        n = r * d * out_limbs
        buf = np.zeros(max(n, 1), dtype=np.uint64)
        check(L.rf_server_read(self.h, rid, buf.ctypes.data, n))
        res = []
        for e in range(r * d):
            if out_limbs == 0:
                res.append(0)
                continue
            v = int.from_bytes(buf[e * out_limbs:(e + 1) * out_limbs].tobytes(), "little")
            if v >> (64 * out_limbs - 1):
                v -= 1 << (64 * out_limbs)
            res.append(v)
        return res

    def release(self, rid):
        check(_bind().rf_server_release(self.h, rid))

    def stats(self):
        st = ServerStats()
        check(_bind().rf_server_stats_get(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def clock(self):
        return _bind().rf_server_clock(self.h)

    def close(self):
        if self.h is not None:
            check(_bind().rf_server_destroy(self.h))
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gen_entry_device(buf_ptr, bits, seed, entry):
    check(_bind().rf_gen_entry_device(ctypes.c_void_p(buf_ptr), bits, seed & _M64, entry))


def checksum_device(ptr, n, entry):
    s = ctypes.c_uint64()
    check(_bind().rf_checksum_device(ctypes.c_void_p(ptr), n, entry, ctypes.byref(s)))
    return s.value


def gen_entries_device(ptrs, bits, seeds, entries):
    """rf_gen_entries_device over parallel lists."""
    n = len(ptrs)
    if n == 0:
        return
    a = lambda v, t: np.ascontiguousarray(np.array(v, dtype=t))
    P, B, S, E = a(ptrs, np.uint64), a(bits, np.uint64), a([x & _M64 for x in seeds], np.uint64), a(entries, np.uint32)
    check(_bind().rf_gen_entries_device(P.ctypes.data, B.ctypes.data, S.ctypes.data, E.ctypes.data, n))


def checksums_device(ptrs, nlimbs, entries, slots, nslots):
    """rf_checksums_device over parallel lists; returns nslots sums."""
    a = lambda v, t: np.ascontiguousarray(np.array(v, dtype=t))
    P, N, E, S = a(ptrs, np.uint64), a(nlimbs, np.uint64), a(entries, np.uint32), a(slots, np.uint32)
    out = np.zeros(max(nslots, 1), dtype=np.uint64)
    check(_bind().rf_checksums_device(P.ctypes.data, N.ctypes.data, E.ctypes.data, S.ctypes.data, len(ptrs),
                                      out.ctypes.data, max(nslots, 1)))
    return [int(x) for x in out[:nslots]]
