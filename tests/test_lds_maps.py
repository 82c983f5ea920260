"""CPU checks of the transform kernels' LDS exchange maps (rforest_amd/csrc/
rf_ntt.hip: lds_slot, lds_at, lds_map, with the group and index functions
fwd_group, fwd_index, inv_index, inv_r0 and Pass restated here).  For every
transform length that runs whole waves (2^10 to 2^14 points) and every
exchange of the forward and inverse transforms:

* the chosen slot map is a bijection of the N residues into the transform's
  LDS (N + N/16 doubles from 2^13 points, N below);
* lds_at's split of an address into a per-thread base plus a constant offset
  gives exactly the slot of the full index, for every thread, group and
  register (the index is the OR of the thread's bits and the register
  position's);
* each 32-lane half of every single ds_read_b64 / ds_write_b64 covers 32
  distinct slots modulo 32 (all 64 banks) wherever the XOR map of the
  previous kernels did.

The GPU parity tests show the maps are bijections in effect; bank
distinctness and the base-plus-offset identity are what they cannot see."""
import pytest


def shape(lgn):
    n = 1 << lgn
    e = 16 if lgn >= 4 else n
    t = n // e
    return n, e, t, (lgn + 3) // 4


def pass_of(lgn, p, r0=4):
    l0 = 0 if p == 0 else r0 + 4 * (p - 1)
    r = min(r0, lgn) if p == 0 else min(lgn - l0, 4)
    return l0, r


def npasses(lgn, r0):
    return 1 if lgn <= r0 else 1 + (lgn - r0 + 3) // 4


def inv_r0(lgn):
    _, e, _, _ = shape(lgn)
    return ((lgn - 1) % 4) + 1 if (e == 16 and lgn >= 5 and ((lgn - 1) % 4) < 2) else 4


def fwd_group(lgn, p, tid, u):
    _, e, t, np_ = shape(lgn)
    _, r = pass_of(lgn, p)
    if 0 < p == np_ - 1 and (e >> r) > 1 and t > 64:
        return ((tid >> 6) << (10 - r)) | (u << 6) | (tid & 63)
    return tid + t * u


def fwd_index(lgn, p, g, j):
    l0, r = pass_of(lgn, p)
    nlo = lgn - l0 - r
    return ((g >> nlo) << (lgn - l0)) | (j << nlo) | (g & ((1 << nlo) - 1))


def inv_index(lgn, p, g, j, r0):
    l0, r = pass_of(lgn, p, r0)
    return ((g >> l0) << (l0 + r)) | (j << l0) | (g & ((1 << l0) - 1))


def lds_slot(sw, i):
    if sw == 0:
        return i
    if sw == 2:
        return i + ((i >> 5) << 1)
    if sw == 3:
        return i + (i >> 4)
    if sw == 4:
        return (i + ((i >> 5) << 1)) ^ ((i >> 5) & 1)
    m = i >> 5
    return i ^ (((m & 15) << 1) | (m & 1))


def lds_at(sw, t, o):
    """Slot of t | o as the kernel forms it from the thread's part t and the
    constant o."""
    if sw == 1:
        return ((lds_slot(1, t) << 3) ^ (lds_slot(1, o) << 3)) >> 3
    if sw == 4:
        pt, po = lds_slot(2, t), lds_slot(2, o)
        return (pt & ~1) + (po & ~1) + (((pt | po) ^ ((t ^ o) >> 5)) & 1)
    return lds_slot(sw, t) + lds_slot(sw, o)


def lds_map(pad, wide, r1, r2, a, b):
    if not wide:
        return 1
    a, b = min(a, b), max(a, b)
    if a >= 5:
        return 0
    if not pad:
        return 1
    if r1 >= 3 and r2 >= 3 and a >= 1 and b >= 5:
        return 2
    if r1 == 4 and r2 == 4 and a == 0 and b == 4:
        return 3
    return 4


def exchanges(lgn):
    """(name, map, sides): each side is (index(tid, u, j), groups per
    thread, group size) of the pass before and after the exchange."""
    n, e, t, np_ = shape(lgn)
    pad = lgn >= 13
    out = []
    for p in range(1, np_):
        lp, rp = pass_of(lgn, p - 1)
        lc, rc = pass_of(lgn, p)
        sw = lds_map(pad, t >= 64, rp, rc, lgn - lp - rp, lgn - lc - rc)
        sides = ((lambda tid, u, j, p=p: fwd_index(lgn, p - 1, fwd_group(lgn, p - 1, tid, u), j), e >> rp, 1 << rp),
                 (lambda tid, u, j, p=p: fwd_index(lgn, p, fwd_group(lgn, p, tid, u), j), e >> rc, 1 << rc))
        out.append((f"fwd{p}", sw, sides))
    r0 = inv_r0(lgn)
    for p in range(1, npasses(lgn, r0)):
        lp, rp = pass_of(lgn, p - 1, r0)
        lc, rc = pass_of(lgn, p, r0)
        sw = lds_map(pad, t >= 64, rp, rc, lp, lc)
        sides = ((lambda tid, u, j, p=p: inv_index(lgn, p - 1, tid + t * u, j, r0), e >> rp, 1 << rp),
                 (lambda tid, u, j, p=p: inv_index(lgn, p, tid + t * u, j, r0), e >> rc, 1 << rc))
        out.append((f"inv{p}", sw, sides))
    return out


def worst_conflict(lgn, sides, sw):
    _, _, t, _ = shape(lgn)
    worst = 1
    for idx, ng, gs in sides:
        for w0 in range(0, t, 64):
            for half in (0, 32):
                lanes = range(w0 + half, w0 + half + 32)
                for u in range(ng):
                    for j in range(gs):
                        banks = [lds_slot(sw, idx(tid, u, j)) % 32 for tid in lanes]
                        worst = max(worst, max(banks.count(b) for b in set(banks)))
    return worst


@pytest.mark.parametrize("lgn", [10, 11, 12, 13, 14])
def test_exchange_maps(lgn):
    n, _, t, _ = shape(lgn)
    ldsn = n + (n >> 4) if lgn >= 13 else n
    for name, sw, sides in exchanges(lgn):
        (wi, wng, wgs), (ri, rng, rgs) = sides
        written = set()
        for tid in range(t):
            for u in range(wng):
                for j in range(wgs):
                    s = lds_slot(sw, wi(tid, u, j))
                    assert 0 <= s < ldsn and s not in written, (lgn, name, tid, u, j)
                    written.add(s)
        assert len(written) == n
        read = {lds_slot(sw, ri(tid, u, j)) for tid in range(t) for u in range(rng) for j in range(rgs)}
        assert read == written, (lgn, name)
        for idx, ng, gs in sides:
            for tid in range(t):
                base = idx(tid, 0, 0)
                for u in range(ng):
                    for j in range(gs):
                        o = idx(0, u, j)
                        assert base & o == 0, (lgn, name)
                        assert lds_at(sw, base, o) == lds_slot(sw, idx(tid, u, j)), (lgn, name, tid, u, j)
        assert worst_conflict(lgn, sides, sw) <= worst_conflict(lgn, sides, 1), (lgn, name, sw)


def test_padded_maps_are_conflict_free_at_the_headline_length():
    # 2^13 points (the 1-Mbit products): every exchange single-access
    # conflict-free, and the padded maps where the kernels use them
    maps = {name: sw for name, sw, _ in exchanges(13)}
    assert maps == {"fwd1": 0, "fwd2": 2, "fwd3": 4, "inv1": 4, "inv2": 2, "inv3": 0}
    for name, sw, sides in exchanges(13):
        assert worst_conflict(13, sides, sw) == 1, name
