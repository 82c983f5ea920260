#!/usr/bin/env python3
"""Headline benchmark: exact r=2 1-Mbit matrix products per second on one MI355X.

One step = one pass of the whole product pipeline (pack + forward NTT,
pointwise product + inverse NTT, Garner CRT + carry scan) over a batch of
independent products 2x2 . 2x2 whose entries are 1,048,576-bit integers with
the top bit set (GMP generator seeded 42), inputs and outputs resident in
HBM.  Prints one JSON line with the throughput, a roofline entry for the
dominant kernel measured with HIP events on the library's stream, and the
reference CPU path timed on a bounded sample of the same products.
"""
import argparse
import json
import math
import os
import random
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# fp64 VALU: 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz; v_fma_f64, v_mul_f64,
# v_add_f64, v_rndne_f64 and v_mad_u64_u32 all issue at this full rate on
# gfx950 (78.6 TFLOP/s counting an FMA as two flops).
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
WORKLOAD = "2x2 . 2x2 exact integer matrix product, 1,048,576-bit unsigned entries (top bit set), GMP seed 42"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def modcheck(products, results, q):
    """C == A B modulo a random 61-bit prime, for every product (a
    size-independent property of the exact result)."""
    bad = 0
    for (r, d, A, B), C in zip(products, results):
        Am = [x % q for x in A]
        Bm = [x % q for x in B]
        for i in range(r):
            for j in range(d):
                want = sum(Am[i * d + k] * Bm[k * d + j] for k in range(d)) % q
                if C[i * d + j] % q != want:
                    bad += 1
    return bad


def traffic_from_profiles(kernel, batch, plan, field="hbm_bytes_per_launch"):
    """HBM bytes (or another per-launch counter figure) of `kernel` from the
    committed rocprofv3 --pmc summary (profiles/pmc_traffic.json), if it was
    measured for this batch and plan; otherwise None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("batch") != batch or list(rec.get("plan", [])) != list(plan):
        return None
    k = rec.get("kernels", {}).get(kernel)
    return None if k is None else k.get(field)


def cpu_baseline(products, gpu_results, budget_s):
    import oracle
    if oracle.ref_available():
        fn, kind = oracle.ref_rmatrix_mult, "reference"
        desc = "reference CPU mpz_rmatrix_mult_fft (src/lib/hwmpz.c:65-130) built from its sources"
    else:
        def fn(A, r, B, d):
            pl, _ = oracle.ref_plan(max((abs(x).bit_length() + 63) // 64 for x in A + B), d)
            return oracle.rmatrix_mult(A, r, B, d, pl)
        kind = "port"
        desc = "C restatement of the reference path (oracle/rf_oracle.c)"
    n, el, mismatches, lat = 0, 0.0, 0, []
    while n < 5 or el < budget_s:
        i = n % len(products)
        r, d, A, B = products[i]
        t0 = time.perf_counter()
        C = fn(A, r, B, d)
        dt = time.perf_counter() - t0
        el += dt
        lat.append(dt)
        if n < len(products):
            mismatches += int(C != gpu_results[i])
        n += 1
    distinct = min(n, len(products))
    return {"value": n / el, "unit": "products/s", "cores": 1, "kind": kind, "cpu_model": cpu_model(),
            "median_latency_ms": round(sorted(lat)[len(lat) // 2] * 1e3, 3),
            "sample": f"{n} products ({distinct} distinct) of the same workload, one thread, {el:.1f} s; {desc}",
            "gpu_results_identical": mismatches == 0}


def _cpu_fn():
    import oracle
    if oracle.ref_available():
        return oracle.ref_rmatrix_mult
    def fn(A, r, B, d):
        pl, _ = oracle.ref_plan(max((abs(x).bit_length() + 63) // 64 for x in A + B), d)
        return oracle.rmatrix_mult(A, r, B, d, pl)
    return fn


def cpu_worker(seconds, batch):
    """One CPU worker process of the all-core baseline: the same workload,
    products in turn for `seconds`; prints its count as JSON."""
    from rforest_amd import workloads
    products = workloads.headline(min(batch, 8), seed=42)  # the first products of the same stream
    fn = _cpu_fn()
    n, t0 = 0, time.perf_counter()
    while n < 1 or time.perf_counter() - t0 < seconds:
        r, d, A, B = products[n % len(products)]
        fn(A, r, B, d)
        n += 1
    print(json.dumps({"n": n, "s": time.perf_counter() - t0}), flush=True)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_all_cores(seconds, batch):
    """All-core CPU throughput: one worker process per core this process may
    use, at most 16 -- the host CPU share that comes with one GPU on these
    boxes, whatever the machine's own core count -- started before this
    process touches the GPU."""
    import subprocess
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    workers = max(1, min(16, allowed))
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", str(seconds),
                               "--batch", str(batch)], stdout=subprocess.PIPE, text=True)
             for _ in range(workers)]
    n, wall = 0, 0.0
    for p in procs:
        out, _ = p.communicate()
        rec = json.loads(out.strip().splitlines()[-1])
        n += rec["n"]
        wall = max(wall, rec["s"])
    return {"workers": workers, "products_per_s": round(n / wall, 2), "products": n, "seconds": round(wall, 2),
            "host_logical_cpus": os.cpu_count(), "affinity_cpus": allowed, "cpu_model": cpu_model(),
            "workers_rule": "one worker per core allowed to this process, capped at the 16-core host share "
                            "that comes with one GPU"}


def dropin_section(rf, products, seconds):
    """The drop-in as rforest calls it: mpz_rmatrix_mult_fft on host mpz_t,
    one headline product per call (PCIe transfers, staging and the sign
    epilogue included), the mpz arrays made once; per-call latency
    percentiles over `seconds` of back-to-back calls."""
    from rforest_amd.gmp import MpzArray
    n = min(8, len(products))
    arrs = [(MpzArray(A), MpzArray(B), MpzArray.zeros(4)) for (_, _, A, B) in products[:n]]
    w = MpzArray.zeros(1)
    L = rf.lib()
    for i in range(2 * n):  # warm: plans, staging buffers
        mA, mB, mC = arrs[i % n]
        L.mpz_rmatrix_mult_fft(mC.ptr(), mA.ptr(), 2, mB.ptr(), 2, w.ptr())
    lat = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or len(lat) < 100:
        mA, mB, mC = arrs[len(lat) % n]
        t0 = time.perf_counter()
        L.mpz_rmatrix_mult_fft(mC.ptr(), mA.ptr(), 2, mB.ptr(), 2, w.ptr())
        lat.append(time.perf_counter() - t0)
    exact = all(arrs[i][2].values() == rf.mpz_rmatrix_mult_naive(products[i][2], 2, products[i][3], 2)
                for i in range(min(2, n)))
    for a in arrs:
        for m in a:
            m.close()
    w.close()
    return {"workload": "one 2x2 . 2x2 product of 1,048,576-bit entries per mpz_rmatrix_mult_fft call, "
                        "host mpz_t in and out (PCIe included)",
            "calls": len(lat), "products_per_s": round(len(lat) / sum(lat), 1),
            "p50_us": round(_pct(lat, 50) * 1e6, 1), "p99_us": round(_pct(lat, 99) * 1e6, 1),
            "checks": {"exact_vs_gmp": exact}}


def tree_level_section(rf, workloads, reps):
    """1024 independent 2x2 . 2x2 products of uneven sizes (entries
    log-uniform in [2^10, 2^18) bits, each matrix within 2x, random signs):
    one prepared batch of two lanes (its length groups' kernels, each a few
    tens of microseconds, overlap pairwise) replayed from its graph, against
    the plain path (each product alone through the one-product pipeline, and
    the host mpz drop-in on a sample)."""
    prods = workloads.tree_level(1024, seed=1003)
    batch = rf.DeviceBatch(prods)
    batch.prepare(lanes=2)
    batch.launch(graph=True)
    rf.device_sync()
    res = batch.results()
    q = (1 << 61) - 1
    bad = modcheck(prods, res, q)
    exact = all(res[i] == rf.mpz_rmatrix_mult_naive(prods[i][2], 2, prods[i][3], 2) for i in range(0, 1024, 97))
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.launch(graph=True)
    rf.device_sync()
    batched = 1024 * reps / (time.perf_counter() - t0)
    batch.run_one(0)
    t0 = time.perf_counter()
    for i in range(1024):
        batch.run_one(i)
    plain_dev = 1024 / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    ns = 128
    for i in range(ns):
        rf.mpz_rmatrix_mult_fft(prods[i][2], 2, prods[i][3], 2)
    plain_host = ns / (time.perf_counter() - t0)
    batch.release()
    # the same level planned by the cost model alone (no measured plan table)
    from rforest_amd import tune
    nrules = tune.reset()
    tune.clear()
    mb = rf.DeviceBatch(prods)
    mb.prepare(lanes=2)
    mb.launch(graph=True)
    rf.device_sync()
    same_model = mb.results() == res
    t0 = time.perf_counter()
    for _ in range(reps):
        mb.launch(graph=True)
    rf.device_sync()
    model = 1024 * reps / (time.perf_counter() - t0)
    mb.release()
    tune.reset()
    return {"workload": "1024 products 2x2 . 2x2, entries log-uniform in [2^10, 2^18) bits (matrix within 2x), "
                        "random signs, GMP seed 1003; two lanes",
            "autotuned_plans": nrules,
            "cost_model_products_per_s": round(model, 1),
            "speedup_vs_cost_model": round(batched / model, 3),
            "batched_products_per_s": round(batched, 1),
            "plain_device_products_per_s": round(plain_dev, 1),
            "plain_dropin_products_per_s": round(plain_host, 1),
            "speedup_vs_plain_device": round(batched / plain_dev, 2),
            "speedup_vs_plain_dropin": round(batched / plain_host, 2),
            "checks": {"mod_2^61-1_mismatches": bad, "exact_sample": exact, "identical_to_cost_model": same_model}}


def tree_section(rf, workloads, reps):
    """A product tree of 1024 2x2 matrices with 4096-bit signed entries (GMP
    seed 1006) kept in device memory (rf_mpz_tree_product: leaves up, root
    down), against the same tree level by level through the host mpz
    drop-in."""
    import ctypes
    from rforest_amd.gmp import MpzArray
    rng = workloads.GmpRandom(1006)
    mats = [[rng.signed(4096) for _ in range(4)] for _ in range(1024)]
    rng.close()
    L = rf.lib()
    L.rf_mpz_tree_product.restype = ctypes.c_int
    L.rf_mpz_tree_product.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    leaves = MpzArray([x for M in mats for x in M])
    root = MpzArray.zeros(4)
    for _ in range(3):  # warm-up: the library's tree blocks, plans and copy threads
        rf.check(L.rf_mpz_tree_product(root.ptr(), leaves.ptr(), len(mats), 2))
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        rf.check(L.rf_mpz_tree_product(root.ptr(), leaves.ptr(), len(mats), 2))
        ts.append(time.perf_counter() - t0)
    t_dev = sorted(ts)[len(ts) // 2]  # median: one tree is 1-2 ms, so host jitter moves a mean
    got = root.values()
    # level by level through the drop-in
    w = MpzArray.zeros(1)
    t0 = time.perf_counter()
    cur = [MpzArray(M) for M in mats]
    t_conv = time.perf_counter() - t0
    t0 = time.perf_counter()
    while len(cur) > 1:
        nxt = []
        for i in range(0, len(cur) - 1, 2):
            C = MpzArray.zeros(4)
            L.mpz_rmatrix_mult_fft(C.ptr(), cur[i].ptr(), 2, cur[i + 1].ptr(), 2, w.ptr())
            nxt.append(C)
        if len(cur) % 2:
            nxt.append(cur[-1])
        cur = nxt
    t_plain = time.perf_counter() - t0
    same = cur[0].values() == got
    root_bits = max(abs(x).bit_length() for x in got)
    return {"workload": "product tree of 1024 2x2 matrices, 4096-bit signed entries, GMP seed 1006 "
                        f"(1023 products; root entries of {root_bits} bits)",
            "device_tree_s": round(t_dev, 5),
            "device_tree_timing": f"median of {reps} trees after 3 warm-up trees",
            "dropin_levels_s": round(t_plain, 5),
            "speedup": round(t_plain / t_dev, 2),
            "checks": {"identical_to_dropin": same, "mpz_conversion_excluded_s": round(t_conv, 4)}}


def resident_section(rf, workloads, reps):
    """One 3x3 left matrix of 4,194,304-bit entries times 16 right 3x3
    matrices of 1,048,576-bit entries: the left operand transformed once and
    kept in the NTT domain, against transforming it in every product."""
    A, Bs = workloads.resident_case(1004)
    plain = rf.DeviceBatch([(3, 3, A, B) for B in Bs])
    plain.prepare()
    plain.launch(graph=True)
    rf.device_sync()
    want = plain.results()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain.launch(graph=True)
    rf.device_sync()
    t_plain = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    op = rf.ResidentOperand(A, 3, 3, 1 << 20)
    t_create = time.perf_counter() - t0
    res = rf.DeviceBatch([(3, 3, op, B) for B in Bs])
    res.prepare()
    res.launch(graph=True)
    rf.device_sync()
    got = res.results()
    t0 = time.perf_counter()
    for _ in range(reps):
        res.launch(graph=True)
    rf.device_sync()
    t_res = (time.perf_counter() - t0) / reps
    exact = got[0] == rf.mpz_rmatrix_mult_naive(A, 3, Bs[0], 3)
    t, lgN, k = op.plan()
    res.release()
    plain.release()
    op.release()
    return {"workload": "3x3 left of 4,194,304-bit entries x 16 right 3x3 of 1,048,576-bit entries, GMP seed 1004",
            "plan": {"chunk_bits": 64 * t, "ntt_length": 1 << lgN, "primes": k},
            "resident_products_per_s": round(16 / t_res, 1),
            "recompute_products_per_s": round(16 / t_plain, 1),
            "speedup": round(t_plain / t_res, 2),
            "operand_create_ms": round(t_create * 1e3, 3),
            "transforms_per_16": {"resident": 9 + 16 * (9 + 9), "recompute": 16 * (9 + 9 + 9)},
            "checks": {"identical_to_recompute": got == want, "exact_product0": exact}}


def remainder_section(rf, workloads, reps):
    """A remainder-tree level: 256 remainders X mod D of 2,097,152-bit signed
    dividends by 1,048,576-bit moduli, each modulus prepared once (D and its
    inverse kept as resident transforms), all remainders of the level in one
    call, against one remainder per call through mpz_t and GMP's mpz_mod on
    one host core."""
    from rforest_amd.gmp import mod_values
    from rforest_amd.mod import ModBatch, Modulus
    pairs = workloads.remainder_level(256, seed=1007)
    import ctypes
    from rforest_amd.gmp import MpzArray
    from rforest_amd.mod import _bind as mod_bind
    L = mod_bind()
    ds = MpzArray([d for (d, _) in pairs])  # mpz_t made outside the timed calls
    mb = (ctypes.c_uint64 * len(pairs))(*([1 << 21] * len(pairs)))
    hs = (ctypes.c_void_p * len(pairs))()
    rf.check(L.rf_mod_create_batch(mb, ds.ptr(), 4, hs))  # warm (plans, tables)
    for i in range(4):
        L.rf_mod_destroy(hs[i])
    t_one = []
    for i in range(4):
        t0 = time.perf_counter()
        rf.check(L.rf_mod_create_batch(mb, ctypes.c_void_p(ds.ptr().value + i * ctypes.sizeof(ds.arr[0])), 1, hs))
        t_one.append(time.perf_counter() - t0)
        L.rf_mod_destroy(hs[0])
    t_create_one = sorted(t_one)[len(t_one) // 2]
    # the first whole level also grows the library's device arena and pinned
    # staging buffer (kept for later calls): warm once, then time three
    rf.check(L.rf_mod_create_batch(mb, ds.ptr(), len(pairs), hs))
    for i in range(len(pairs)):
        L.rf_mod_destroy(hs[i])
    t_lv = []
    for _ in range(3):
        t0 = time.perf_counter()
        rf.check(L.rf_mod_create_batch(mb, ds.ptr(), len(pairs), hs))
        t_lv.append(time.perf_counter() - t0)
        for i in range(len(pairs)):
            L.rf_mod_destroy(hs[i])
    t_create = sorted(t_lv)[1] / len(pairs)
    ds.close()
    mods = Modulus.batch([d for (d, _) in pairs], 1 << 21)
    batch = ModBatch([(M, x) for M, (_, x) in zip(mods, pairs)])
    batch.run()
    got = batch.results()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.run()
    t_batch = (time.perf_counter() - t0) / reps
    want = mod_values([x for (_, x) in pairs], [d for (d, _) in pairs])
    ns = 16
    t0 = time.perf_counter()
    one = [mods[i].mod(pairs[i][1]) for i in range(ns)]
    t_one = (time.perf_counter() - t0) / ns
    t0 = time.perf_counter()
    nc = 0
    while nc < 8 or time.perf_counter() - t0 < 2.0:
        mod_values([pairs[nc % len(pairs)][1]], [pairs[nc % len(pairs)][0]])
        nc += 1
    t_cpu = (time.perf_counter() - t0) / nc
    batch.release()
    for M in mods:
        M.close()
    return {"workload": "256 remainders X mod D, X 2,097,152-bit signed, D 1,048,576-bit (top bit set), GMP seed 1007",
            "batched_remainders_per_s": round(len(pairs) / t_batch, 1),
            "one_per_call_remainders_per_s": round(1 / t_one, 1),
            "gmp_cpu_remainders_per_s": round(1 / t_cpu, 1),
            "gmp_cpu_sample": f"{nc} calls of mpz_mod on one host core",
            "modulus_create_ms": round(t_create * 1e3, 3),
            "modulus_create_sample": "per modulus, median of 3 calls of rf_mod_create_batch preparing the level's 256 "
                                     "(after one untimed call of the same size) "
                                     "(mpz_t made beforehand); the reciprocals by the device's Newton recursion",
            "modulus_create_one_ms": round(t_create_one * 1e3, 3),
            "checks": {"identical_to_gmp": got == want, "one_per_call_identical": one == want[:ns]}}


def _pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(math.ceil(q / 100.0 * len(xs))) - 1)] if xs else None


def _run_stream(rf, serve, reqs, cfg, rate=None):
    """Submits reqs to a server built with cfg (all at once, or open loop with
    the stream's exponential gaps scaled to `rate` per second); returns
    (completions by tag, wall seconds from first submission to last
    completion, server stats)."""
    srv = serve.Server(**cfg)
    t0 = srv.clock()
    due = 0.0
    comps = []
    for i, (r, bits, signs, seed, gap) in enumerate(reqs):
        if rate:
            due += gap / rate
            while srv.clock() - t0 < due:
                comps += srv.poll(4096, 0)
        srv.submit_synth(r, r, bits, signs, seed, tag=i)
    comps += srv.wait_all(len(reqs) - len(comps))
    t1 = max(c["t_done"] for c in comps)
    st = srv.stats()
    srv.close()
    return {c["tag"]: c for c in comps}, t1 - t0, st


def _run_stream_host(rf, serve, reqs, limbs, outs, cfg, rate=None):
    """As _run_stream, with every request's entries supplied from host
    memory (RF_IN_HOST) and its result copied back (RF_OUT_HOST) into the
    caller's buffers `outs`: PCIe both ways."""
    srv = serve.Server(**cfg)
    t0 = srv.clock()
    due = 0.0
    comps, ids = [], []
    out = serve.RF_OUT_HOST | serve.RF_OUT_CHECKSUM
    for i, (r, bits, signs, seed, gap) in enumerate(reqs):
        if rate:
            due += gap / rate
            while srv.clock() - t0 < due:
                comps += srv.poll(4096, 0)
        ids.append(srv.submit_host_limbs(r, r, limbs[i], tag=i, output=out, out=outs[i]))
    comps += srv.wait_all(len(reqs) - len(comps))
    t1 = max(c["t_done"] for c in comps)
    for rid in ids:  # drop the host result buffers
        srv._keep.pop(rid, None)
    srv.close()
    return {c["tag"]: c for c in comps}, t1 - t0


def serving_host_section(rf, workloads, sat, n_host):
    """The first n_host requests of the served stream with host-memory
    inputs and outputs (the PCIe-inclusive served figure), saturated and
    open loop at 0.7 x the saturated rate; each result's checksum must equal
    the synthetic run's (sat) for the same request."""
    from rforest_amd import serve
    reqs = workloads.served_stream(n_host, seed=1005)
    limbs = [[(sg, serve.gen_limbs(b, seed, e), b) for e, (b, sg) in enumerate(zip(bits, signs))]
             for (r, bits, signs, seed, _) in reqs]
    nbytes = sum(a.nbytes for ents in limbs for (_, a, _) in ents)
    # the caller's result buffers, allocated and touched once (a serving
    # loop reuses its buffers; first-touch page faults are not the server's)
    sizer = serve.Server(1 << 30, streams=1)
    outs = [np.ones(r * r * sizer.out_limbs_for(r, r, ents), dtype=np.uint64)
            for (r, _, _, _, _), ents in zip(reqs, limbs)]
    sizer.close()
    cfg = dict(hbm_budget=64 * 10**9, streams=3, max_batch=256)
    got, t_sat = _run_stream_host(rf, serve, reqs, limbs, outs, cfg)
    same = all(got[i]["checksum"] == sat[i]["checksum"] and got[i]["status"] == rf.RF_OK for i in range(n_host))
    rate = 0.7 * n_host / t_sat
    op, _ = _run_stream_host(rf, serve, reqs, limbs, outs, cfg, rate)
    lat = [c["t_done"] - c["t_submit"] for c in op.values()]
    return {"workload": f"first {n_host} requests of the served stream, entries in host memory (RF_IN_HOST), "
                        f"results copied back (RF_OUT_HOST): {nbytes / 1e9:.2f} GB of input over PCIe",
            "products_per_s": round(n_host / t_sat, 1),
            "open_loop": {"rate_per_s": round(rate, 1), "p50_ms": round(_pct(lat, 50) * 1e3, 3),
                          "p99_ms": round(_pct(lat, 99) * 1e3, 3)},
            "checks": {"identical_to_synthetic_run": same}}


def serving_resident_section(rf, workloads, n_req, n_ops):
    """Requests that reuse left operands kept resident in the NTT domain
    inside the server (rf_request.left), against the same requests carrying
    the left matrix themselves: n_ops resident 3x3 operands of
    1,048,576-bit entries, n_req requests each naming one (seeded), right
    3x3 operands of 1,048,576-bit synthetic entries generated in HBM."""
    from rforest_amd import serve
    rng = workloads.GmpRandom(1008)
    As = [[rng.signed(1 << 20) for _ in range(9)] for _ in range(n_ops)]
    pick = [rng.urandomb(16) % n_ops for _ in range(n_req)]
    rng.close()
    ops = [rf.ResidentOperand(A, 3, 3, 1 << 20) for A in As]
    bits = [1 << 20] * 9
    cfg = dict(hbm_budget=64 * 10**9, streams=3, max_batch=256)

    def run(resident):
        srv = serve.Server(**cfg)
        t0 = srv.clock()
        for i, j in enumerate(pick):
            if resident:
                srv.submit_synth(3, 3, bits, [1] * 9, 1008 * (1 << 20) + i, tag=i, left=ops[j])
            else:  # A generated in HBM too (the same sizes; the operand's values are not needed for timing)
                srv.submit_synth(3, 3, bits + bits, [1] * 18, 1008 * (1 << 20) + i, tag=i)
        comps = srv.wait_all(n_req)
        el = max(c["t_done"] for c in comps) - t0
        ok = all(c["status"] == rf.RF_OK for c in comps)
        srv.close()
        return n_req / el, ok, {c["tag"]: c for c in comps}

    res_rate, ok1, comps = run(True)
    plain_rate, ok2, _ = run(False)
    # exactness: a sample of resident results against GMP on the host-generated right operands
    exact = True
    srv = serve.Server(**cfg)
    for i in range(0, n_req, max(1, n_req // 4)):
        B = [serve.gen_value(1 << 20, 1, 1008 * (1 << 20) + i, e) for e in range(9)]
        rid = srv.submit_host(3, 3, None, B, tag=i, left=ops[pick[i]])
        (c,) = srv.wait_all(1)
        C = srv.host_result(rid, c["out_limbs"])
        exact = exact and C == rf.mpz_rmatrix_mult_naive(As[pick[i]], 3, B, 3)
    srv.close()
    for op in ops:
        op.release()
    return {"workload": f"{n_req} requests 3x3 . 3x3 with 1,048,576-bit entries, each naming one of {n_ops} left "
                        "operands kept resident in the NTT domain (GMP seed 1008); right operands generated in HBM",
            "resident_products_per_s": round(res_rate, 1),
            "left_supplied_products_per_s": round(plain_rate, 1),
            "speedup": round(res_rate / plain_rate, 2),
            "checks": {"all_ok": ok1 and ok2, "exact_sample_vs_gmp": exact}}


def serving_section(rf, workloads, n_req, n_plain, n_dropin, n_open):
    """The mixed request stream through the request server (continuous
    batching, 64 GB HBM budget, synthetic entries generated in device memory
    at admission, results checksummed on the device), against the plain
    per-request path: the same server with one stream slot and one product
    per batch, and the host mpz drop-in on a prefix."""
    from rforest_amd import serve
    reqs = workloads.served_stream(n_req, seed=1005)
    budget = 64 * 10**9
    served_cfg = dict(hbm_budget=budget, streams=3, max_batch=256)
    plain_cfg = dict(hbm_budget=budget, streams=1, max_batch=1)
    sat, t_sat, st = _run_stream(rf, serve, reqs, served_cfg)
    bad = sum(1 for c in sat.values() if c["status"] != rf.RF_OK)
    plain, t_plain, _ = _run_stream(rf, serve, reqs[:n_plain], plain_cfg)
    same = all(plain[i]["checksum"] == sat[i]["checksum"] for i in plain)
    served_rate = len(reqs) / t_sat
    plain_rate = n_plain / t_plain
    # exactness of a few small requests against GMP on host-generated entries
    exact = True
    small = [i for i in range(n_plain) if reqs[i][1][0] < 200000][:4]
    for i in small:
        r, bits, signs, seed, _ = reqs[i]
        vals = [serve.gen_value(b, sg, seed, e) for e, (b, sg) in enumerate(zip(bits, signs))]
        C = rf.mpz_rmatrix_mult_naive(vals[:r * r], r, vals[r * r:], r)
        exact = exact and serve.checksum_values(C, sat[i]["out_limbs"]) == sat[i]["checksum"]
    # the host mpz drop-in (PCIe and mpz conversion included) on a prefix
    from rforest_amd.gmp import MpzArray
    arrs = []
    for (r, bits, signs, seed, _) in reqs[:n_dropin]:
        vals = [serve.gen_value(b, sg, seed, e) for e, (b, sg) in enumerate(zip(bits, signs))]
        arrs.append((r, MpzArray(vals[:r * r]), MpzArray(vals[r * r:]), MpzArray.zeros(r * r)))
    w = MpzArray.zeros(1)
    t0 = time.perf_counter()
    for (r, mA, mB, mC) in arrs:
        rf.lib().mpz_rmatrix_mult_fft(mC.ptr(), mA.ptr(), r, mB.ptr(), r, w.ptr())
    dropin_rate = n_dropin / (time.perf_counter() - t0)
    for a in arrs:
        for m in a[1:]:
            m.close()
    # the reference CPU path on a prefix of the stream (at most 200 requests
    # or 15 s of one host core), extrapolated to the whole stream
    fn = _cpu_fn()
    n_cpu, t_cpu = 0, 0.0
    while n_cpu < min(200, len(reqs)) and t_cpu < 15.0:
        r, bits, signs, seed, _ = reqs[n_cpu]
        vals = [serve.gen_value(b, sg, seed, e) for e, (b, sg) in enumerate(zip(bits, signs))]
        t0 = time.perf_counter()
        fn(vals[:r * r], r, vals[r * r:], r)
        t_cpu += time.perf_counter() - t0
        n_cpu += 1
    cpu_rate = n_cpu / t_cpu
    # open loop at 0.7 x the plain capacity: latency of served vs plain
    rate = 0.7 * plain_rate
    op, _, _ = _run_stream(rf, serve, reqs[:n_open], served_cfg, rate)
    pp, _, _ = _run_stream(rf, serve, reqs[:n_open], plain_cfg, rate)
    lat = [c["t_done"] - c["t_submit"] for c in op.values()]
    plat = [c["t_done"] - c["t_submit"] for c in pp.values()]
    host = serving_host_section(rf, workloads, sat, min(1000, n_req))
    return {"workload": f"{n_req} requests, r in {{2,3,4}} w.p. {{0.6,0.3,0.1}}, entries log-uniform in [2^12, 2^24] bits, "
                        "random signs, GMP seed 1005; entries generated in HBM at admission; 64 GB HBM budget",
            "host_fed": host,
            "served_products_per_s": round(served_rate, 1),
            "plain_products_per_s": round(plain_rate, 1),
            "plain_sample": f"first {n_plain} requests, same server with one slot and one product per batch",
            "dropin_products_per_s": round(dropin_rate, 1),
            "dropin_sample": f"first {n_dropin} requests, host mpz_rmatrix_mult_fft (PCIe included)",
            "cpu_reference_products_per_s": round(cpu_rate, 2),
            "cpu_reference_sample": f"first {n_cpu} requests on one host core ({t_cpu:.1f} s), the reference CPU path; "
                                    f"extrapolated to the whole stream: {len(reqs) / cpu_rate:.0f} s",
            "speedup_vs_plain": round(served_rate / plain_rate, 2),
            "speedup_vs_dropin": round(served_rate / dropin_rate, 2),
            "batches": st["batches"], "mean_batch": round(st["products_in_batches"] / max(st["batches"], 1), 1),
            "peak_pool_gb": round(st["peak_bytes"] / 1e9, 2),
            "open_loop": {"rate_per_s": round(rate, 1), "requests": n_open,
                          "p50_ms": round(_pct(lat, 50) * 1e3, 3), "p99_ms": round(_pct(lat, 99) * 1e3, 3),
                          "plain_p50_ms": round(_pct(plat, 50) * 1e3, 3), "plain_p99_ms": round(_pct(plat, 99) * 1e3, 3)},
            "checks": {"failed_requests": bad, "mix_equals_alone": same, "exact_sample": exact}}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=1024, help="products per step")
    ap.add_argument("--lanes", type=int, default=3, help="overlapping pipelines per prepared batch")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-worker", type=float, default=0.0, help=argparse.SUPPRESS)
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip timing the host mpz drop-in (PCIe-inclusive, one product per call)")
    ap.add_argument("--requests", type=int, default=20000, help="requests in the served stream")
    ap.add_argument("--no-extra", action="store_true",
                    help="headline only (skip the tree-level and resident-operand sections)")
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(args.cpu_worker, args.batch)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the product pipeline runs on a single MI355X; "
              "sharding over several GPUs is a design, not a measured path (--gpus 1 only)",
              file=sys.stderr)
        sys.exit(2)
    all_cores = None
    if not args.no_cpu:
        log("timing the CPU baseline on all cores (before the GPU is used)")
        all_cores = cpu_all_cores(args.cpu_seconds, args.batch)

    import rforest_amd as rf
    from rforest_amd import workloads

    log("generating inputs")
    products = workloads.headline(args.batch, seed=42)
    plan = rf.plan(1 << 20, 1 << 20, 2, 2)
    t, lgN, k = plan
    N = 1 << lgN
    rf.hw_mpz_setup()
    # One device-to-device copy of 256 MiB: under rocprofv3 --pmc its
    # FETCH_SIZE / WRITE_SIZE calibrate the traffic counters (known bytes).
    cal = [rf.DeviceBuffer(256 << 20) for _ in range(2)]
    rf.check(rf.lib().rf_memcpy_d2d(cal[1].ptr, cal[0].ptr, 256 << 20))
    for c in cal:
        c.free()
    batch = rf.DeviceBatch(products)
    log(f"plan: chunk {64 * t} bits, N = 2^{lgN}, k = {k} primes; batch {args.batch}")

    batch.run(sync=True)  # one-shot path once, then the prepared pipeline
    batch.prepare(lanes=args.lanes)
    for _ in range(max(args.warmup, 1)):
        batch.launch(graph=True)
    rf.device_sync()
    results = batch.results()
    exact0 = results[0] == rf.mpz_rmatrix_mult_naive(products[0][2], 2, products[0][3], 2)
    q = (1 << 61) - 1
    bad = modcheck(products, results, q)
    if not exact0 or bad:
        log(f"RESULT MISMATCH: exact product 0 {exact0}, modular mismatches {bad}")
        sys.exit(1)

    rf.device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.launch(graph=True)
    rf.device_sync()
    el = time.perf_counter() - t0
    ms_per_step = el / args.steps * 1e3
    value = args.batch * args.steps / el

    # Per-kernel timing with HIP events on the library's stream, on one lane
    # (kernels not overlapped, whole batch per launch).
    batch.release()
    batch.prepare(lanes=1)
    rf.profile(enable=True)
    rf.profile(reset=True)
    psteps = max(3, min(args.steps, 10))
    for _ in range(psteps):
        batch.launch(graph=False)  # events need eager launches
    rf.device_sync()
    prof = rf.profile()
    rf.profile(enable=False)
    # algorithmic bytes: one read and one write of every residue word per
    # transform (per prime), fp64 words
    per_transform = 2 * N * 8
    fwd_bytes = args.batch * (4 + 4) * k * per_transform
    inv_bytes = args.batch * 4 * k * per_transform
    alg = {"fwd": fwd_bytes, "inv": inv_bytes}
    kern_names = {"fwd": "k_fwd (pack + forward NTT)", "inv": "k_inv (pointwise + inverse NTT)"}
    # algorithmic fp64 VALU operations: 8 per radix-2 butterfly (a modular
    # product is 6: two products and a rounding for the quotient, one FMA, the
    # error term's add; then the sum and the difference), (N/2) lgN
    # butterflies per transform and prime; the pack 8 per 64-bit limb and
    # prime (two conversions, six FMAs); the pointwise product 7 per term
    # (a modular product of two data operands)
    bfly_ops = 8 * (N // 2) * lgN
    alg_ops = {"fwd": args.batch * (4 + 4) * k * (bfly_ops + 8 * (1 << 20) // 64),
               "inv": args.batch * 4 * k * (bfly_ops + 7 * 2 * N)}
    dom = max(("fwd", "inv"), key=lambda s: prof[s][0])
    ms, nl = prof[dom]
    avg_s = ms / nl / 1e3
    achieved = alg[dom] / avg_s / 1e9
    valu_achieved = alg_ops[dom] / avg_s / 1e12
    traffic = traffic_from_profiles(dom, args.batch, plan)
    valu_insts = traffic_from_profiles(dom, args.batch, plan, "valu_wave_insts_per_launch")
    valu_busy = None if valu_insts is None else round(valu_insts * 64 / (VALU_PEAK_TOPS * 1e12) / avg_s, 4)
    ntt_ms = prof["fwd"][0] + prof["inv"][0]
    ntt_launch_s = ntt_ms / psteps / 1e3
    ntt_ach = (fwd_bytes + inv_bytes) / ntt_launch_s / 1e9
    # secondary figure: radix-2 butterflies, (N/2) lgN per transform and prime
    bfly = (4 + 4 + 4) * k * (N // 2) * lgN
    stage_ms = {s: round(prof[s][0] / psteps, 4) for s in prof}

    dropin = None
    if not args.no_dropin:
        log("drop-in: one product per mpz_rmatrix_mult_fft call")
        dropin = dropin_section(rf, products, 2.0)

    extra = {}
    if not args.no_extra:
        log("tree level: batched vs plain")
        extra["tree_level"] = tree_level_section(rf, workloads, max(3, args.steps // 4))
        log("resident operand")
        extra["resident_operand"] = resident_section(rf, workloads, max(3, args.steps // 4))
        log("product tree in device memory")
        extra["product_tree"] = tree_section(rf, workloads, 15)
        log("remainder-tree level")
        extra["remainder_level"] = remainder_section(rf, workloads, 3)
        log("serving: mixed request stream")
        extra["serving"] = serving_section(rf, workloads, args.requests, min(2000, args.requests), 50,
                                           min(5000, args.requests))
        log("serving: resident left operands")
        extra["serving_resident"] = serving_resident_section(rf, workloads, 512, 16)

    cpu = None
    if not args.no_cpu:
        log("timing the CPU baseline")
        cpu = cpu_baseline(products, results, args.cpu_seconds)
        cpu["all_cores"] = all_cores

    out = {
        "metric": "r=2 1-Mbit matmuls/sec; mixed-stream p50/p99 latency, batched vs plain",
        "value": round(value, 2),
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "products_per_step": args.batch, "r": 2, "d": 2,
                   "entry_bits": 1 << 20, "chunk_bits": 64 * t, "ntt_length": N, "primes": k,
                   "prime_bits": 48, "parallelism": "single GPU",
                   "lanes": args.lanes},
        # the dominant kernel issues fp64 VALU work at about half the VALU
        # peak and moves about half the HBM peak; its counters show the VALU
        # busy and the rest waiting on loads, so the binding roof named is
        # the VALU's, with the HBM figures beside it
        "roofline": {"bound": "valu", "kernel": kern_names[dom], "achieved": round(valu_achieved, 3),
                     "peak": round(VALU_PEAK_TOPS, 3), "unit": "Tops/s",
                     "frac": round(valu_achieved / VALU_PEAK_TOPS, 4), "traffic": traffic,
                     "alg_ops_per_launch": alg_ops[dom], "valu_busy": valu_busy,
                     "hbm": {"achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(achieved / HBM_PEAK_GBS, 4), "alg_bytes_per_launch": alg[dom]},
                     "avg_launch_ms": round(avg_s * 1e3, 5)},
        "roofline_ntt_passes": {"achieved": round(ntt_ach, 2), "frac": round(ntt_ach / HBM_PEAK_GBS, 4),
                                "alg_bytes_per_product": (fwd_bytes + inv_bytes) // args.batch,
                                "alg_bytes_per_product_reference_params": 25165824,
                                "frac_reference_params": round(25165824 * args.batch / ntt_launch_s / 1e9 / HBM_PEAK_GBS, 4),
                                "butterflies_per_product": bfly,
                                "gbutterflies_per_s": round(bfly * args.batch / ntt_launch_s / 1e9, 2)},
        "stage_ms_per_step": stage_ms,
        "dropin": dropin,
        "cpu_baseline": cpu,
        "checks": {"exact_product0": exact0, "mod_2^61-1_mismatches": bad},
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
