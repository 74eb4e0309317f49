"""Saturating histogram adds as 64-bit word-pair atomics: the algebra.

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
