"""Device-resident batches: many products of one product-tree level in one
call of rf_batch_mult_device, with inputs and outputs in HBM (torch
tensors as plain device memory)."""
import ctypes

from . import DevEntry, DevOut, DevProduct, check, lib


class DeviceBatch:
    """Uploads a list of products [(r, d, A, B), ...] (Python ints) once; run()
    enqueues the whole pipeline over all of them; results() reads back."""

    def __init__(self, products, device="cuda"):
        import numpy as np
        import torch
        self.products = products
        L = lib()
        chunks, descs = [], []
        pos = 0
        plan_in = []
        for (r, d, A, B) in products:
            ents = []
            for x in list(A) + list(B):
                x = int(x)
                n = (abs(x).bit_length() + 63) // 64
                if n:
                    chunks.append(np.frombuffer(abs(x).to_bytes(8 * n, "little"), dtype=np.int64))
                ents.append((pos, n, (x > 0) - (x < 0)))
                pos += n
            la = max([n for (_, n, s) in ents[:r * d] if s] or [0])
            lb = max([n for (_, n, s) in ents[r * d:] if s] or [0])
            plan_in.append((r, d, ents, int(L.rf_out_limbs(la, lb, d))))
        host = np.concatenate(chunks) if chunks else np.zeros(1, dtype=np.int64)
        self.limbs = torch.from_numpy(np.ascontiguousarray(host)).to(device)
        self.out_sizes = []
        total_out = sum(max(rn, 1) * r * d for (r, d, _, rn) in plan_in)
        self.out = torch.zeros(max(total_out, 1), dtype=torch.int64, device=device)
        base = self.limbs.data_ptr()
        obase = self.out.data_ptr()
        self._keep = []
        prods = (DevProduct * len(products))()
        opos = 0
        self.layout = []
        for pi, (r, d, ents, rn) in enumerate(plan_in):
            A = (DevEntry * (r * d))()
            B = (DevEntry * (d * d))()
            C = (DevOut * (r * d))()
            for e, (off, n, s) in enumerate(ents):
                tgt = A[e] if e < r * d else B[e - r * d]
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
            self._keep += [A, B, C]
        self.prods = prods

    def run(self, sync=True):
        check(lib().rf_batch_mult_device(self.prods, len(self.products), None, 0 if sync else 1))

    def results(self):
        """Per product, the r*d entries of C as Python ints."""
        import numpy as np
        import torch
        torch.cuda.synchronize()
        host = self.out.cpu().numpy().view(np.uint64)
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
