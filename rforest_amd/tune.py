"""Autotuner for the planner's free choices.

A product's plan is a chunk width 64 t (t = 1..8), which fixes the transform
length 2^lgN and the prime count k, and, for 2^14 points, a choice between
one-workgroup transforms and two rows of 2^13 (column + row passes).  For
every plan class (r, lgN, t, split) of square r x r products this measures,
on the device, the time per product of a prepared batch of products at the
class's full size (entries of 64 t 2^(lgN-1) bits, so that the product just
fits 2^lgN points), replayed from its graph; each batch's results are
checked against the same products planned by the cost model.  The planner
then picks, for every square product, the measured valid plan of least cost
(rf_tune_set); the table goes to tuned_plans.txt beside librforest_hip.so,
which hw_mpz_setup loads, so the drop-in uses it too.

    python -m rforest_amd.tune [--r 2 3 4] [--lgn 6 16] [--out PATH]
"""
import argparse
import ctypes
import os
import random
import sys
import time

from . import DeviceBatch, RfError, check, device_sync, hw_mpz_setup, lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tuned_plans.txt")
MAX_BITS = 1 << 24  # the largest entries the mixed request stream carries


def _bind():
    L = lib()
    L.rf_tune_set.restype = ctypes.c_int
    L.rf_tune_set.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.rf_tune_clear.restype = None
    L.rf_tune_load.restype = ctypes.c_int
    L.rf_tune_load.argtypes = [ctypes.c_char_p]
    return L


def set_rule(r, lgN, t, split, cost_us):
    check(_bind().rf_tune_set(r, lgN, t, split, float(cost_us)))


def clear():
    _bind().rf_tune_clear()


def reset():
    """Back to the table hw_mpz_setup loads (tuned_plans.txt beside the
    library), or to the cost model when there is none."""
    clear()
    env = os.environ.get("RF_TUNE_FILE")
    path = env if env else DEFAULT_OUT
    if path != "none" and os.path.exists(path):
        return load(path)
    return 0


def force(r, t, split=0):
    """Square r x r products use chunk width 64 t whenever that plan is valid
    (a zero-cost measurement for every transform length)."""
    for lgN in range(1, 18):
        set_rule(r, lgN, t, split if lgN == 14 else 0, 0.0)


def load(path=DEFAULT_OUT):
    n = _bind().rf_tune_load(path.encode())
    if n < 0:
        raise RfError(n, lib().rf_last_error().decode())
    return n


def plan(bits, r):
    t, lg, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(lib().rf_plan_query(bits, bits, r, r, ctypes.byref(t), ctypes.byref(lg), ctypes.byref(k)))
    return t.value, lg.value, k.value


def class_bits(lgN, t):
    return 64 * t << (lgN - 1)


def classes(r, lgn_lo, lgn_hi):
    """Plan classes (lgN, t, split, k) whose full-size entries stay within
    MAX_BITS and whose plan is valid."""
    out = []
    for lgN in range(lgn_lo, lgn_hi + 1):
        for t in range(1, 9):
            bits = class_bits(lgN, t)
            if bits > MAX_BITS:
                continue
            clear()
            set_rule(r, lgN, t, 0, 0.0)
            try:
                pt, pl, k = plan(bits, r)
            except RfError:
                continue
            if (pt, pl) != (t, lgN):
                continue
            out.append((lgN, t, 0, k))
            if lgN == 14:
                out.append((lgN, t, 1, k))
    clear()
    return out


def _specs(r, bits, n, seed):
    rng = random.Random(seed)
    return [(r, r, [bits] * (2 * r * r), [rng.choice([-1, 1]) for _ in range(2 * r * r)],
             rng.getrandbits(64)) for _ in range(n)]


def time_batch(batch, reps):
    batch.prepare()
    for _ in range(2):
        batch.launch(graph=True)
    device_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.launch(graph=True)
    device_sync()
    return (time.perf_counter() - t0) / reps


def measure(r, lgN, t, split, reps=5):
    """Microseconds per product of the class, and whether its results equal
    the cost model's plan of the same products."""
    bits = class_bits(lgN, t)
    n = min(256, max(4, (1 << 27) // (bits * r * r)))
    batch = DeviceBatch.synthetic(_specs(r, bits, n, seed=(r << 20) + (lgN << 8) + t))
    clear()
    batch.run()
    want = batch.checksums()
    set_rule(r, lgN, t, split, 0.0)
    sec = time_batch(batch, reps)
    same = batch.checksums() == want
    batch.release()
    clear()
    return sec / n * 1e6, same, n


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--r", type=int, nargs="+", default=[2, 3, 4])
    ap.add_argument("--lgn", type=int, nargs=2, default=[6, 16])
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args(argv)
    hw_mpz_setup()
    clear()
    lines = ["# r lgN t split us_per_product -- rforest_amd.tune, " + time.strftime("%Y-%m-%d"),
             "# square r x r products at each plan class's full size (entries of 64 t 2^(lgN-1) bits)"]
    bad = 0
    for r in args.r:
        for (lgN, t, split, k) in classes(r, args.lgn[0], args.lgn[1]):
            us, same, n = measure(r, lgN, t, split, args.reps)
            bad += not same
            print(f"r={r} lgN={lgN} t={t} split={split} k={k}: {us:10.2f} us/product "
                  f"({n} products){'' if same else '  RESULTS DIFFER'}", flush=True)
            lines.append(f"{r} {lgN} {t} {split} {us:.3f}")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {args.out}; {bad} classes with differing results", flush=True)
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
