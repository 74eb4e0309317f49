"""HIP render path vs the CPU oracle, through the C ABI (needs an MI355X).

Bars (each asserted below):
  * ISAAC stream and xform-pick sequence bit-exact per walker (trace words);
  * per-walker FP32 trajectories within 1e-5 relative over the first 20 turns;
  * histograms bit-exact where the arithmetic is exact (+,-,*,/ only, no
    transcendentals): the u32 adds are order-independent;
  * DE on an identical histogram within 1e-5 relative; tonemap float output
    within 1e-5 relative and RGBA8 within 1 LSB;
  * whole images by PSNR at matched sample counts.
"""
import numpy as np
import pytest

from helpers import blobs, genome, psnr, sierpinski, xform
from pyr3_amd import pack
from pyr3_amd.genome import Tonemap

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    from pyr3_amd.renderer import Renderer
    return Renderer


def _render_hist_gpu(R, g, W, H, os_=1, iters=100, fuse=20, seed=1, walkers=64, trace=None,
                     jitter=None):
    r = R(W, H, os_)
    r.reset(g)
    if trace:
        r.set_trace(*trace)
    r.iterate(g, seed, walkers, iters, walker_jitter=jitter, fuse=fuse)
    h = r.read_histogram()
    tr = r.read_trace() if trace else None
    return r, h, tr


def test_isaac_stream_and_pick_sequence_bit_exact(orc, R):
    g = genome(sierpinski(), scale=10.0, cx=0.5, cy=0.5)
    W = H = 32
    walkers, iters, fuse = 300, 40, 10
    _, hg, tg = _render_hist_gpu(R, g, W, H, iters=iters, fuse=fuse, seed=0xC0FFEE,
                                 walkers=walkers, trace=(walkers, iters + fuse))
    xf, ds, u, pal = blobs(g, W, H, iters=iters, fuse=fuse, seed=0xC0FFEE, walkers=walkers)
    ho, to = orc.chaos(xf, ds, u, pal, trace_shape=(walkers, iters + fuse))
    # raw pick draw, pick index and draw counter: bit-exact for every walker/turn
    assert np.array_equal(tg[..., 11].view(np.uint32), to[..., 11].view(np.uint32))
    assert np.array_equal(tg[..., 1], to[..., 1])
    assert np.array_equal(tg[..., 10].view(np.uint32), to[..., 10].view(np.uint32))
    # linear maps only: trajectories and histogram are exact
    assert np.array_equal(tg[..., 12:14], to[..., 12:14])
    assert np.array_equal(hg, ho)


def _linear_xforms(n, rng):
    """n contractive linear maps with random weights (attractor near the unit square)."""
    out = []
    for j in range(n):
        a = rng.uniform(0.3, 0.6)
        out.append(xform(float(rng.uniform(0.2, 1.0)),
                         (a, 0.0, float(rng.uniform(0, 0.5)), 0.0, a, float(rng.uniform(0, 0.5))),
                         color=j / max(n - 1, 1)))
    return out


@pytest.mark.parametrize("case", ["lds16_xaos", "hbm20", "unsorted_table"])
def test_pick_paths_bit_exact(orc, R, case):
    """Picks of tables of up to 16 xforms come from step boundaries in LDS,
    others (more xforms, or a caller table whose rows are not sorted) from
    the table itself; both give the table's entry bit for bit."""
    import dataclasses
    from pyr3_amd.renderer import pack_genome
    rng = np.random.default_rng(7)
    n = 20 if case == "hbm20" else (16 if case == "lds16_xaos" else 5)
    xs = _linear_xforms(n, rng)
    if case == "lds16_xaos":
        for x in xs:
            x.xaos = [float(v) for v in rng.choice([0.0, 0.5, 1.0, 3.0], size=n)]
            x.xaos[0] = 1.0
    g = genome(xs, scale=10.0, cx=0.5, cy=0.5)
    W = H = 32
    walkers, iters, fuse, seed = 512, 12, 4, 0xBEEF
    xf, ds, u, pal = blobs(g, W, H, iters=iters, fuse=fuse, seed=seed, walkers=walkers)
    pg = pack_genome(g)
    if case == "unsorted_table":
        ds = np.ascontiguousarray(ds[:, rng.permutation(ds.shape[1])])
        pg = dataclasses.replace(pg, distrib=ds)
    r = R(W, H, 1)
    r.reset(g)
    r.set_trace(walkers, iters + fuse)
    r.iterate(g, seed, walkers, iters, fuse=fuse, packed=pg)
    tg, hg = r.read_trace(), r.read_histogram()
    r.destroy()
    ho, to = orc.chaos(xf, ds, u, pal, trace_shape=(walkers, iters + fuse))
    assert np.array_equal(tg[..., 11].view(np.uint32), to[..., 11].view(np.uint32))
    assert np.array_equal(tg[..., 1], to[..., 1])
    assert len(np.unique(tg[..., 1])) == n
    assert np.array_equal(hg, ho)


def test_xform_blend_bit_exact(orc, R):
    """Blended turns (gate, fallback-row pick, second chain, mix draw) on
    linear maps: draws, picks, trajectories and histogram bit-exact."""
    g = genome(sierpinski(), scale=10.0, cx=0.5, cy=0.5, xform_blend=0.5)
    W = H = 32
    walkers, iters, fuse = 512, 30, 5
    _, hg, tg = _render_hist_gpu(R, g, W, H, iters=iters, fuse=fuse, seed=0xB1E4D,
                                 walkers=walkers, trace=(walkers, iters + fuse))
    xf, ds, u, pal = blobs(g, W, H, iters=iters, fuse=fuse, seed=0xB1E4D, walkers=walkers)
    ho, to = orc.chaos(xf, ds, u, pal, trace_shape=(walkers, iters + fuse))
    assert np.array_equal(tg[..., 10].view(np.uint32), to[..., 10].view(np.uint32))
    assert np.array_equal(tg[..., 1], to[..., 1])
    assert np.array_equal(tg[..., 12:14], to[..., 12:14])
    assert np.array_equal(hg, ho)


def test_isaac_explicit_states_match_device_seeding(orc, R):
    g = genome(sierpinski(), scale=10.0, cx=0.5, cy=0.5)
    W = H = 16
    walkers = 77  # not a multiple of 64
    st = orc.seed_walkers(4242, walkers)
    r = R(W, H, 1)
    r.reset(g)
    r.iterate(g, 4242, walkers, 50, fuse=5)
    a = r.read_histogram()
    r.reset(g)
    r.iterate(g, 4242, walkers, 50, fuse=5, isaac_states=st)
    b = r.read_histogram()
    assert np.array_equal(a, b) and a[..., 3].sum() > 0


def _nonlinear_genome():
    return genome([xform(0.6, (0.6, 0.3, 0.1, -0.2, 0.7, 0.05), vars=((1, 0.5), (2, 0.5)),
                         color=0.1),
                   xform(0.4, (-0.5, 0.2, -0.3, 0.4, 0.5, 0.2), vars=((0, 0.7), (1, 0.3)),
                         color=0.9, post=(0.9, 0.1, 0.0, -0.1, 0.9, 0.0))],
                  scale=40.0)


def _contractive_genome():
    return genome([xform(0.6, (0.45, 0.2, 0.1, -0.15, 0.5, 0.05), vars=((1, 0.4), (0, 0.6)),
                         color=0.1),
                   xform(0.4, (-0.4, 0.1, -0.3, 0.2, 0.45, 0.2), vars=((0, 0.7), (1, 0.3)),
                         color=0.9, post=(0.9, 0.1, 0.0, -0.1, 0.9, 0.0))],
                  scale=40.0)


def _rel(a, b):
    """|a - b| / max(|b|, 1): relative error with absolute error near zero."""
    a, b = a.astype(np.float64), b.astype(np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)


def test_trajectories_within_1e5_over_20_turns(orc, R):
    """Contractive transcendental maps: whole 20-turn trajectories agree."""
    g = _contractive_genome()
    W = H = 64
    walkers, turns = 256, 20
    _, _, tg = _render_hist_gpu(R, g, W, H, iters=turns, fuse=0, seed=7, walkers=walkers,
                                trace=(walkers, turns))
    xf, ds, u, pal = blobs(g, W, H, iters=turns, fuse=0, seed=7, walkers=walkers)
    _, to = orc.chaos(xf, ds, u, pal, trace_shape=(walkers, turns))
    assert np.array_equal(tg[..., 1], to[..., 1])  # same pick sequence
    for col in (2, 3, 4, 5, 6, 7, 9, 12, 13):
        assert _rel(tg[..., col], to[..., col]).max() <= 1e-5, col


def test_one_step_parity_on_expanding_maps(orc, R):
    """Expanding maps amplify 1-ulp transcendental differences turn over turn,
    so each GPU turn is replayed by the oracle from the GPU's own start point."""
    g = _nonlinear_genome()
    W = H = 64
    walkers, turns = 512, 20
    _, _, tg = _render_hist_gpu(R, g, W, H, iters=turns, fuse=0, seed=7, walkers=walkers,
                                trace=(walkers, turns))
    xf, ds, u, pal = blobs(g, W, H, iters=turns, fuse=0, seed=7, walkers=walkers)
    to = orc.chaos_follow(xf, ds, u, pal, tg)
    clean = ~(tg[..., 8] > 0).any(axis=1)          # walkers with no retried turn
    assert clean.mean() > 0.9
    assert np.array_equal(tg[clean][..., 1], to[clean][..., 1])
    assert np.array_equal(tg[clean][..., 10].view(np.uint32), to[clean][..., 10].view(np.uint32))
    for col in (2, 3, 4, 5, 6, 7, 9):
        assert _rel(tg[clean][..., col], to[clean][..., col]).max() <= 1e-5, col


def test_histogram_totals_and_image_psnr(orc, R):
    g = _nonlinear_genome()
    W, H = 96, 80
    walkers, iters = 512, 2000
    r, hg, _ = _render_hist_gpu(R, g, W, H, iters=iters, fuse=20, seed=11, walkers=walkers)
    xf, ds, u, pal = blobs(g, W, H, iters=iters, fuse=20, seed=11, walkers=walkers)
    ho = orc.chaos_mt(xf, ds, u, pal, 8)
    # in-bounds sample count agrees to within chaotic divergence
    cg, co = int(hg[..., 3].sum()), int(ho[..., 3].sum())
    assert abs(cg - co) <= 0.01 * co
    samples = walkers * iters
    r.present(g, samples)
    img_g = r.read_image()
    k1, k2 = pack.calibration(g.scale, samples, 1.0, 1)
    fo = orc.density(ho, pack.de_uniforms(W, H, 9, 0, 0.4, k1, k2, 1))
    _, img_o = orc.tonemap(pack.viz_uniforms(W, H, k1, k2, g.tonemap, 1, 3), pack.spatial_kernel(0.5, 1),
                           filtered=fo)
    assert psnr(img_g, img_o) > 30.0


@pytest.mark.parametrize("os_", [1, 2])
def test_density_and_tonemap_on_identical_histogram(orc, R, os_):
    rng = np.random.default_rng(21 + os_)
    W, H = 70, 45
    SW, SH = W * os_, H * os_
    h = np.zeros((SH, SW, 4), dtype=np.uint32)
    dense = rng.random((SH, SW)) < 0.5
    cnt = (rng.integers(1, 3000, (SH, SW)) * 255).astype(np.uint32)
    h[..., 3] = np.where(dense, cnt, 0)
    for c in range(3):
        h[..., c] = (h[..., 3] * rng.random((SH, SW))).astype(np.uint32)
    g = genome(sierpinski(), scale=30.0)
    g.tonemap = Tonemap(gamma=4.0, vibrancy=1.0, highlight_power=1.0)
    samples = 5e6
    r = R(W, H, os_)
    r.set_image_f32(True)
    r.reset(g)
    r.write_histogram(h)
    r.present(g, samples)
    fg = r.read_filtered()
    k1, k2 = pack.calibration(g.scale, samples, 1.0, os_)
    fo = orc.density(h, pack.de_uniforms(SW, SH, 9, 0, 0.4, k1, k2, os_))
    # every filtered value within 1e-5 of the oracle's, relative to itself
    assert (np.abs(fg.astype(np.float64) - fo) <= 1e-5 * np.abs(fo.astype(np.float64))).all()
    taps = pack.spatial_kernel(0.5, os_)
    of, o8 = orc.tonemap(pack.viz_uniforms(W, H, k1, k2, g.tonemap, os_, len(taps)), taps,
                         filtered=fo)
    gf, g8 = r.read_image_f32(), r.read_image()
    assert np.allclose(gf, of, rtol=1e-5, atol=1e-5)
    assert np.abs(g8.astype(int) - o8.astype(int)).max() <= 1
    # DE off: raw histogram path
    r.present(g, samples, force_de_off=True)
    of2, o82 = orc.tonemap(pack.viz_uniforms(W, H, k1, k2, g.tonemap, os_, len(taps)), taps,
                           hist=h)
    assert np.allclose(r.read_image_f32(), of2, rtol=1e-5, atol=1e-5)
    assert np.abs(r.read_image().astype(int) - o82.astype(int)).max() <= 1


def test_saturation_on_device(orc, R):
    g = genome([xform(1.0, (0, 0, 0, 0, 0, 0))], scale=1.0)
    W = H = 8
    h0 = np.zeros((H, W, 4), dtype=np.uint32)
    h0[H // 2, W // 2] = 0xFFFFFFFF - 1000
    r = R(W, H, 1)
    r.reset(g)
    r.write_histogram(h0)
    r.iterate(g, 5, 3, 100, fuse=1)
    out = r.read_histogram()
    assert out[H // 2, W // 2, 3] == 0xFFFFFFFF and out[H // 2, W // 2, 0] == 0xFFFFFFFF


@pytest.mark.parametrize("case", ["near_max", "low_words_wrap", "mixed"])
def test_saturating_adds_bit_exact_near_the_limit(orc, R, case):
    """The splat (one packed 64-bit add per sample with carry counts, merged
    into the u32 histogram before every read) against the oracle's per-word
    saturating sums, on histograms seeded next to 2^32 - 1 where the
    attractor lands: each word must end at min(seed + sum of adds,
    2^32 - 1) exactly. Two dispatches in a row (the second adds to merged
    words)."""
    MAX = 0xFFFFFFFF
    W, H = 24, 20
    g = genome(sierpinski(), scale=18.0, cx=0.5, cy=0.5)
    rng = np.random.default_rng({"near_max": 1, "low_words_wrap": 2, "mixed": 3}[case])
    h0 = np.zeros((H, W, 4), dtype=np.uint64)
    if case == "near_max":
        h0[...] = MAX - rng.integers(0, 40000, (H, W, 4))
    elif case == "low_words_wrap":          # R and B near the limit, G and count low
        h0[..., 0] = MAX - rng.integers(0, 3000, (H, W))
        h0[..., 2] = MAX - rng.integers(0, 3000, (H, W))
        h0[..., 1] = rng.integers(0, 1000, (H, W))
        h0[..., 3] = MAX - 1 - rng.integers(0, 2, (H, W))   # high word one below the limit
    else:
        h0[...] = rng.integers(0, MAX + 1, (H, W, 4), dtype=np.uint64)
        h0[rng.random((H, W)) < 0.5] = MAX - 300
    h0 = h0.astype(np.uint32)
    r = R(W, H, 1)
    r.reset(g)
    r.write_histogram(h0)
    ho = h0.copy()
    for seed, walkers, iters in ((11, 700, 120), (12, 300, 90)):
        r.iterate(g, seed, walkers, iters, fuse=10)
        xf, ds, u, pal = blobs(g, W, H, iters=iters, fuse=10, seed=seed, walkers=walkers)
        orc.chaos(xf, ds, u, pal, hist=ho)
        out = r.read_histogram()
        assert np.array_equal(out, ho), (case, seed)
    assert (ho == MAX).sum() > 0 and (ho < MAX).sum() > 0


def test_wide_splat_words_saturate(orc, R):
    """A splat with a word of 2^16 or more (here a palette of 300, so the
    colour words are 76,500) bypasses the packed accumulator: four u32 adds
    straight into the histogram, pinned at 2^32 - 1 on wrap, while the count
    word still goes through the packed add. Against the oracle's saturating
    sums on a histogram seeded next to the limit, two dispatches in a row."""
    from pyr3_amd.genome import ColorStop, Palette
    MAX = 0xFFFFFFFF
    W, H = 24, 20
    g = genome(sierpinski(), scale=18.0, cx=0.5, cy=0.5)
    g.palette = Palette("wide", [ColorStop(0, 300.0, 300.0, 300.0),
                                 ColorStop(1, 300.0, 200.0, 300.0)])
    rng = np.random.default_rng(5)
    h0 = (MAX - rng.integers(0, 4_000_000, (H, W, 4), dtype=np.int64)).astype(np.uint32)
    r = R(W, H, 1)
    r.reset(g)
    r.write_histogram(h0)
    ho = h0.copy()
    for seed, walkers, iters in ((21, 512, 80), (22, 256, 60)):
        r.iterate(g, seed, walkers, iters, fuse=10)
        xf, ds, u, pal = blobs(g, W, H, iters=iters, fuse=10, seed=seed, walkers=walkers)
        assert pal[:, :3].min() >= 200.0
        orc.chaos(xf, ds, u, pal, hist=ho)
        assert np.array_equal(r.read_histogram(), ho), seed
    r.destroy()
    assert (ho[..., :3] == MAX).sum() > 0 and (ho[..., 3] < MAX).sum() > 0


def test_retry_path_on_device(orc, R):
    g = genome([xform(1.0, (1, 0, 0, 0, 1, 0), vars=((0, 1e12),))])
    iters = 12
    _, h, tg = _render_hist_gpu(R, g, 16, 16, iters=iters, fuse=0, walkers=3, trace=(3, iters))
    xf, ds, u, pal = blobs(g, 16, 16, iters=iters, fuse=0, walkers=3)
    _, to = orc.chaos(xf, ds, u, pal, trace_shape=(3, iters))
    assert np.array_equal(tg[..., 10].view(np.uint32), to[..., 10].view(np.uint32))
    assert np.array_equal(tg[..., 6:8], to[..., 6:8])


def test_unsupported_variation_is_an_error(R):
    from pyr3_amd._lib import P3Error
    g = genome([xform(1.0, vars=((326, 1.0),))])  # one past the catalog
    r = R(16, 16, 1)
    r.reset(g)
    with pytest.raises(P3Error):
        r.iterate(g, 1, 8, 10)


def _shifted(g, k):
    import copy
    h = copy.deepcopy(g)
    for x in h.xforms:
        x.affine.c += 0.01 * k
        x.affine.f -= 0.005 * k
    return h


def test_batched_iterate_equals_sequential(orc, R):
    """One batched launch of sub-dispatches (own records, seeds, walker counts,
    one of them empty) leaves the histogram of one iterate per sub-dispatch."""
    W, H = 64, 48
    seeds = [(11 + k * 0x9e3779b1) & 0xFFFFFFFF for k in range(4)]
    walkers = [700, 0, 1300, 256]
    for base in (_nonlinear_genome(), genome(sierpinski(), scale=20.0, cx=0.5, cy=0.5)):
        subs = [_shifted(base, k) for k in range(4)]
        rb = R(W, H, 1)
        rb.reset(base)
        rb.iterate_multi(subs, seeds, walkers, 150, fuse=10)
        hb = rb.read_histogram()
        rs = R(W, H, 1)
        rs.reset(base)
        for g, sd, w in zip(subs, seeds, walkers):
            rs.iterate(g, sd, w, 150, fuse=10)
        hs = rs.read_histogram()
        assert np.array_equal(hb, hs) and hb[..., 3].sum() > 0
    # exact-arithmetic shape: the oracle's per-sub-dispatch sum is bit-identical
    ho = np.zeros_like(hb)
    for g, sd, w in zip(subs, seeds, walkers):
        if w:
            xf, ds, u, pal = blobs(g, W, H, iters=150, fuse=10, seed=sd, walkers=w)
            orc.chaos(xf, ds, u, pal, hist=ho)
    assert np.array_equal(hb, ho)


def test_batched_iterate_rejects_mixed_shapes(R):
    from pyr3_amd._lib import P3Error
    a = _nonlinear_genome()
    b = genome(sierpinski()[:2], scale=40.0)
    r = R(32, 32, 1)
    r.reset(a)
    with pytest.raises((P3Error, ValueError)):
        r.iterate_multi([a, b], [1, 2], [64, 64], 10)
