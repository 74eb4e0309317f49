"""Histogram adds as 64-bit atomics (word pairs, and packed 16-bit fields):
the algebra.

The chaos kernel adds (R, G) and (B, count) as two 64-bit atomics. A 64-bit
atomic is linearizable, so each add's returned old value is the exact 64-bit
prefix; the add that carries out of the low word, or wraps the 64 bits,
counts the event, and the fix-up (hist_fixup in p3_image.hip) sets
  low  = 2^32 - 1 if any carry, else the stored low word,
  high = min(stored high + 2^32 x wraps - carries, 2^32 - 1).
Here that scheme runs in a sequential model of the atomics (any
linearization of concurrent adds is some order of them) and must equal the
reference's per-word saturating sum min(start + sum of adds, 2^32 - 1), over
random orders, increments from 0 to 2^32 - 1, starting words near 2^32, and
several dispatch epochs with a fix-up after each.
"""
import random

import numpy as np

MAX = 0xFFFFFFFF
M64 = (1 << 64) - 1


def run_pair(lo0, hi0, adds):
    """Stored pair after the adds and the fix-up, via the event scheme."""
    v = (hi0 << 32) | lo0
    carries = wraps = 0
    for dlo, dhi in adds:
        inc = dlo | (dhi << 32)
        if inc == 0:
            continue
        o = v
        v = (v + inc) & M64
        if (o & MAX) + dlo > MAX:
            carries += 1
        if o + inc > M64:
            wraps += 1
    lo, hi = v & MAX, v >> 32
    s_hi = hi + (wraps << 32) - carries
    return (MAX if carries else lo), min(s_hi, MAX)


def test_fixup_equals_saturating_sums():
    rng = random.Random(7)
    for trial in range(4000):
        lo, hi = (rng.choice([0, rng.randrange(MAX), MAX - rng.randrange(3000), MAX])
                  for _ in range(2))
        true_lo, true_hi = lo, hi
        for _epoch in range(rng.randrange(1, 4)):
            n = rng.randrange(0, 60)
            big = rng.random() < 0.2
            adds = [(rng.randrange(MAX + 1) if big and rng.random() < 0.3 else rng.randrange(256),
                     rng.randrange(MAX + 1) if big and rng.random() < 0.3 else rng.randrange(256))
                    for _ in range(n)]
            rng.shuffle(adds)
            lo, hi = run_pair(lo, hi, adds)
            true_lo = min(true_lo + sum(a for a, _ in adds), MAX)
            true_hi = min(true_hi + sum(b for _, b in adds), MAX)
            assert (lo, hi) == (true_lo, true_hi), (trial, adds)


def test_fixup_edge_cases():
    # the high word exactly at max - 1 receiving a carry (stored high = max)
    assert run_pair(MAX, MAX - 1, [(1, 0)]) == (MAX, MAX - 1)
    # repeated carries into a saturated low word keep the high word exact
    assert run_pair(MAX, 5, [(255, 1)] * 10) == (MAX, 15)
    # the 64-bit wrap: high saturates, low keeps its own sum
    assert run_pair(10, MAX - 3, [(5, 7)]) == (15, MAX)
    # both: low carries and the pair wraps in one add
    assert run_pair(MAX - 1, MAX, [(2, 0)]) == (MAX, MAX)
    # nothing reaches the limit: the stored words are the sums
    assert run_pair(100, 200, [(1, 2), (3, 4)]) == (104, 206)


# Packed form (add_packed in p3_chaos_kernel.h, hist_merge in p3_image.hip):
# (R, G, B, count), each below 2^16, as four 16-bit fields of one u64 per bin;
# every add's carry out of each field is counted; the merge adds
# S_k = field_k + 2^16 c_k - c_{k-1} to the u32 words, saturating.

def run_packed(words, adds):
    """u32 words after the packed adds and the merge (sequential model)."""
    acc = 0
    c = [0, 0, 0, 0]
    for d in adds:
        assert all(0 <= x < 1 << 16 for x in d)
        v = d[0] | d[1] << 16 | d[2] << 32 | d[3] << 48
        if v == 0:
            continue
        o = acc
        n = (o + v) & M64
        x = o ^ v ^ n
        for k, bit in enumerate((16, 32, 48)):
            c[k] += (x >> bit) & 1
        c[3] += n < o
        acc = n
    out = []
    for k in range(4):
        f = (acc >> (16 * k)) & 0xFFFF
        s = f + (c[k] << 16) - (c[k - 1] if k else 0)
        assert s >= 0
        out.append(min(words[k] + s, MAX))
    return out


def test_packed_merge_equals_saturating_sums():
    rng = random.Random(11)
    for trial in range(3000):
        words = [rng.choice([0, rng.randrange(MAX), MAX - rng.randrange(70000), MAX])
                 for _ in range(4)]
        n = rng.randrange(0, 400)
        top = rng.choice([256, 1 << 16])
        adds = [tuple(rng.randrange(top) for _ in range(4)) for _ in range(n)]
        if rng.random() < 0.2:   # fields driven through many carries
            adds += [(0xFFFF, 0xFFFF, 0xFFFF, 0xFFFF)] * rng.randrange(1, 40)
        rng.shuffle(adds)
        got = run_packed(words, adds)
        want = [min(words[k] + sum(d[k] for d in adds), MAX) for k in range(4)]
        assert got == want, (trial, words)


def test_packed_carry_chain_edge_cases():
    # a carry into a field at 0xFFFF ripples: field 1 adds 0xFFFF + 1
    assert run_packed([0] * 4, [(0xFFFF, 0xFFFF, 0, 0), (1, 0, 0, 0)]) == [0x10000, 0xFFFF, 0, 0]
    # ripple through every field and out of bit 63
    assert run_packed([0] * 4, [(0xFFFF,) * 4, (1, 0, 0, 0)]) == [0x10000, 0xFFFF, 0xFFFF, 0xFFFF]
    # 2^16 adds of one count: the count field wraps exactly once per 2^16
    assert run_packed([7, 0, 0, MAX - 5], [(0, 0, 0, 1)] * (1 << 16)) == [7, 0, 0, MAX]
    # numpy cross-check of the field sums on a large random batch
    rng = np.random.default_rng(3)
    d = rng.integers(0, 1 << 16, (5000, 4), dtype=np.int64)
    assert run_packed([0] * 4, [tuple(map(int, r)) for r in d]) == [int(x) for x in d.sum(0)]
