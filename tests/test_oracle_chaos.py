"""Oracle chaos loop vs closed forms and the RNG-consumption contract.

The reference has no vectors for the loop (src/shaders/chaos.wgsl:8269-8732);
these pin it by what the algorithm fixes:
  * identity map, jitter 0: every walker stays at its first two rand_11 draws,
    so the histogram is a brute-force sum over walkers;
  * opacity 0 deposits nothing; at opacity 1 the count channel is 255 x the
    in-bounds post-fuse samples;
  * draw count per accepted turn = pick + variation draws + 2 jitter (+ final
    opacity gate when the final's opacity is not 1); a bad turn adds the two
    reseed draws and retries up to four times without advancing i (except at
    i = 0), the fifth bad turn is accepted (chaos.wgsl:8391-8455);
  * with xform blend on, each turn adds its gate draw, and when the gate
    passes a fallback-row pick and the mix draw (chaos.wgsl:8372-8381);
  * saturation pins at 2^32 - 1 (chaos.wgsl:8120-8135).
"""
import numpy as np
import pytest

from helpers import blobs, genome, sierpinski, xform
from pyr3_amd import pack


def test_identity_map_brute_force(orc):
    W = H = 32
    g = genome([xform(1.0, (1, 0, 0, 0, 1, 0), vars=((0, 1.0),), color=0.3, speed=0.0)],
               scale=W / 2.0)
    walkers, iters = 50, 37
    xf, ds, u, pal = blobs(g, W, H, iters=iters, fuse=3, seed=99, walkers=walkers, jitter=0.0)
    hist, _ = orc.chaos(xf, ds, u, pal)
    want = np.zeros((H, W, 4), dtype=np.uint64)
    st = orc.seed_walkers(99, walkers)
    for w in range(walkers):
        words, _ = orc.isaac_draw(st[w], 3)
        x, y = orc.rand_transforms(words[:2])[1]
        z0 = orc.rand_transforms(words[2:3])[0][0]
        z = np.float32(z0)
        # colour coordinate converges to 0.3 * 0 speed -> stays z0
        px = np.float32(x) * np.float32(W / 2.0) + np.float32(W * 0.5)
        py = np.float32(y) * np.float32(W / 2.0) + np.float32(H * 0.5)
        xi, yi = int(np.floor(px)), int(np.floor(py))
        if 0 <= xi < W and 0 <= yi < H:
            pi = min(int(np.float32(z) * np.float32(255.0)), 255)
            c = pal[pi]
            want[yi, xi, 0] += int(np.float32(c[0]) * np.float32(255.0)) * iters
            want[yi, xi, 1] += int(np.float32(c[1]) * np.float32(255.0)) * iters
            want[yi, xi, 2] += int(np.float32(c[2]) * np.float32(255.0)) * iters
            want[yi, xi, 3] += 255 * iters
    assert np.array_equal(hist.astype(np.uint64), want)


def test_opacity_zero_deposits_nothing(orc):
    g = genome([xform(1.0, (0.5, 0, 0, 0, 0.5, 0), opacity=0.0)])
    xf, ds, u, pal = blobs(g, 16, 16, walkers=10)
    hist, _ = orc.chaos(xf, ds, u, pal)
    assert not hist.any()


def test_count_channel_is_255_per_in_bounds_sample(orc):
    g = genome(sierpinski(), scale=10.0, cx=0.5, cy=0.5)
    W = H = 32
    xf, ds, u, pal = blobs(g, W, H, iters=200, fuse=20, walkers=40)
    hist, _ = orc.chaos(xf, ds, u, pal)
    # attractor lies in [0,1]^2 -> all samples in bounds at scale 10 around (0.5, 0.5)
    assert hist[..., 3].sum() == 255 * 200 * 40
    assert np.all(hist[..., :3] <= hist[..., 3:4])


def _draws_per_turn(trace_w):
    d = trace_w[:, 10].view(np.uint32).astype(np.int64)
    return np.diff(np.concatenate([[3], d]))  # 3 init draws precede turn 0


def test_draw_contract_plain(orc):
    g = genome(sierpinski(), scale=10.0)
    xf, ds, u, pal = blobs(g, 16, 16, iters=30, fuse=5, walkers=4)
    _, tr = orc.chaos(xf, ds, u, pal, trace_shape=(4, 35))
    for w in range(4):
        assert np.all(_draws_per_turn(tr[w]) == 3)  # pick + 2 jitter


def test_draw_contract_final_opacity_gate(orc):
    fin = xform(0.0, (1, 0, 0, 0, 1, 0), opacity=0.5)
    g = genome(sierpinski(), scale=10.0, final=fin)
    xf, ds, u, pal = blobs(g, 16, 16, iters=30, fuse=5, walkers=4)
    _, tr = orc.chaos(xf, ds, u, pal, trace_shape=(4, 35))
    assert np.all(_draws_per_turn(tr[0]) == 4)  # + final opacity gate
    fin.opacity = 1.0
    xf, ds, u, pal = blobs(g, 16, 16, iters=30, fuse=5, walkers=4)
    _, tr = orc.chaos(xf, ds, u, pal, trace_shape=(4, 35))
    assert np.all(_draws_per_turn(tr[0]) == 3)


def test_draw_contract_xform_blend(orc):
    # blend gate draw, then (when it passes) a fallback-row pick draw, the
    # second chain's draws (none: linear maps) and the mix draw
    # (chaos.wgsl:8372-8381)
    g = genome(sierpinski(), scale=10.0, xform_blend=1.0)
    xf, ds, u, pal = blobs(g, 16, 16, iters=30, fuse=5, walkers=4)
    _, tr = orc.chaos(xf, ds, u, pal, trace_shape=(4, 35))
    assert np.all(_draws_per_turn(tr[0]) == 6)
    g.xform_blend = 0.5
    xf, ds, u, pal = blobs(g, 16, 16, iters=30, fuse=5, walkers=4)
    _, tr = orc.chaos(xf, ds, u, pal, trace_shape=(4, 35))
    d = np.concatenate([_draws_per_turn(tr[w]) for w in range(4)])
    assert set(np.unique(d)) == {4, 6}


def test_pre_blur_brute_force(orc):
    """pre_blur (chaos.wgsl:8164-8181, flam3's pre_blur): five draws per turn
    after the pick, r0..r3 giving the radius weight * (r0 + r1 + r2 + r3 - 2)
    and r4 the angle 2 pi r4, added to the affine image before the variation
    sum; with linear at weight 0.8 beside it the sum is 0.8 * (pa + offset).
    Replayed from each walker's own ISAAC words, the affine image from the
    previous turn's position."""
    aff = (0.6, -0.2, 0.1, 0.3, 0.7, -0.2)
    g = genome([xform(1.0, aff, vars=((97, 0.3), (0, 0.8)))], scale=8.0)
    walkers, iters = 6, 20
    xf, ds, u, pal = blobs(g, 16, 16, iters=iters, fuse=0, seed=5, walkers=walkers)
    _, tr = orc.chaos(xf, ds, u, pal, trace_shape=(walkers, iters))
    st = orc.seed_walkers(5, walkers)
    for w in range(walkers):
        assert np.all(_draws_per_turn(tr[w]) == 8)   # pick + 5 pre_blur + 2 jitter
        words, _ = orc.isaac_draw(st[w], 3 + 8 * iters)
        r01 = orc.rand_transforms(words)[0].astype(np.float64)
        x, y = (float(v) for v in orc.rand_transforms(words[:2])[1])
        for i in range(iters):
            pa = (aff[0] * x + aff[1] * y + aff[2], aff[3] * x + aff[4] * y + aff[5])
            assert np.allclose(tr[w, i, 2:4], pa, rtol=1e-5, atol=1e-6), (w, i)
            r = r01[3 + 8 * i + 1:3 + 8 * i + 6]
            rad = 0.3 * (r[0] + r[1] + r[2] + r[3] - 2.0)
            ang = r[4] * 2 * np.pi
            want = (0.8 * (pa[0] + np.cos(ang) * rad), 0.8 * (pa[1] + np.sin(ang) * rad))
            assert np.allclose(tr[w, i, 4:6], want, rtol=1e-5, atol=1e-6), (w, i)
            x, y = float(tr[w, i, 12]), float(tr[w, i, 13])


def test_retry_window_all_bad(orc):
    """Every turn is bad: turns 1..4 retry, the 5th is accepted."""
    g = genome([xform(1.0, (1, 0, 0, 0, 1, 0), vars=((0, 1e12),))])
    iters, fuse = 12, 0
    xf, ds, u, pal = blobs(g, 16, 16, iters=iters, fuse=fuse, walkers=1)
    _, tr = orc.chaos(xf, ds, u, pal, trace_shape=(1, iters))
    # simulate the loop's bookkeeping
    draws, i, consec, last = 3, 0, 0, {}
    while i < iters:
        draws += 3           # pick + 2 reseed
        consec += 1
        if 0 < consec < 5:
            last[i] = draws
            if i > 0:
                i -= 1
            i += 1
            continue
        consec = 0
        draws += 2           # jitter
        last[i] = draws
        i += 1
    got = tr[0, :, 10].view(np.uint32)
    assert [int(got[k]) for k in range(iters)] == [last[k] for k in range(iters)]
    assert np.all(tr[0, :, 8] == 1.0)


def test_saturation_pins_at_u32_max(orc):
    g = genome([xform(1.0, (0, 0, 0, 0, 0, 0))], scale=1.0)  # every sample lands on (0,0)
    W = H = 8
    xf, ds, u, pal = blobs(g, W, H, iters=100, fuse=1, walkers=3)
    hist = np.zeros((H, W, 4), dtype=np.uint32)
    hist[H // 2, W // 2] = 0xFFFFFFFF - 1000
    out, _ = orc.chaos(xf, ds, u, pal, hist=hist)
    assert out[H // 2, W // 2, 3] == 0xFFFFFFFF
    assert out[H // 2, W // 2, 0] == 0xFFFFFFFF
    assert out.sum(dtype=np.uint64) == np.uint64(4 * 0xFFFFFFFF)


def test_mt_equals_single_thread(orc):
    g = genome(sierpinski(), scale=12.0, cx=0.5, cy=0.5)
    xf, ds, u, pal = blobs(g, 24, 20, iters=300, fuse=10, walkers=37)
    a, _ = orc.chaos(xf, ds, u, pal)
    b = orc.chaos_mt(xf, ds, u, pal, 5)
    assert np.array_equal(a, b)


def test_follow_mode_replays_own_trace(orc):
    """Replaying a run from its own recorded turn end points reproduces it."""
    g = genome([xform(0.6, (0.6, 0.3, 0.1, -0.2, 0.7, 0.05), vars=((1, 0.5), (2, 0.5))),
                xform(0.4, (-0.5, 0.2, -0.3, 0.4, 0.5, 0.2), vars=((0, 0.7), (1, 0.3)))],
               scale=40.0)
    xf, ds, u, pal = blobs(g, 32, 32, iters=20, fuse=0, walkers=64)
    _, tr = orc.chaos(xf, ds, u, pal, trace_shape=(64, 20))
    rep = orc.chaos_follow(xf, ds, u, pal, tr)
    clean = ~(tr[..., 8] > 0).any(axis=1)
    assert clean.sum() > 50
    assert np.array_equal(rep[clean].view(np.uint32), tr[clean].view(np.uint32))
