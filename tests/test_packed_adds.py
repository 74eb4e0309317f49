"""Histogram adds as one 64-bit atomic per splat (packed 16-bit fields):
the algebra.

The chaos kernel adds (R, G, B, count), each below 2^16, as four 16-bit
fields of one u64 accumulator per bin (add_packed in p3_chaos_kernel.h). A
64-bit atomic is linearizable, so each add's returned old value is the exact
prefix of the accumulator; the add that carries out of a field (bits 16, 32,
48, or the 64-bit wrap) counts that carry for the bin, and the merge
(hist_merge in p3_image.hip) adds S_k = field_k + 2^16 c_k - c_{k-1} to the
u32 words, saturating. Here that scheme runs in a sequential model of the
atomics (any linearization of concurrent adds is some order of them) and must
equal the reference's per-word saturating sum min(start + sum of adds,
2^32 - 1), over random orders, increments up to 2^16 - 1 and starting words
near 2^32.
"""
import random

import numpy as np

MAX = 0xFFFFFFFF
M64 = (1 << 64) - 1


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
