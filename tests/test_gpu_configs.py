"""The benchmark configurations' workloads on the GPU, at their full sizes.

  C0  256x256, os 1, q50, fuse 20, 3 xforms of linear / sinusoidal /
      spherical: the reference's own walker schedule (computeDispatch,
      src/renderer.ts:57-78: 800 walkers x 4,096 iterations) on the GPU and
      on the oracle from the same seeds. The ISAAC streams and picks are the
      same; positions drift apart once a transcendental differs in its last
      ulp, so the histograms are compared as two renders of one attractor
      (in-bounds samples and channel sums within 1 %, 16 x 16-bin block
      densities within 3 %). Then DE and the tonemap run on the GPU's own
      histogram on both sides: every filtered value within 1e-5 relative of
      the oracle's, the float image within 1e-5 and RGBA8 within 1 LSB.
  C1  1024x768, os 2, q200, 6 xforms of 2-4 flam3-core variations with
      post-affines: 157 M samples. Size-independent invariants of the
      histogram (opacity-1 counts are multiples of 255, colour words never
      exceed the count word, in-bounds samples at most the samples issued,
      bit-identical on a second render: saturating integer adds do not
      depend on their order), then DE / tonemap of a 256x192 crop of it
      against the oracle as for C0.
  C2  3840x2160, os 2, q200, 12 xforms + final from the whole catalog:
      1.66 G samples into a 531 MB grid, the same invariants (direct-colour
      words may exceed the count: the dc colours are not clamped), every DE /
      tonemap output finite. The same genome with opacities U(0.05, 1), at
      quality 2 so that most bins take one or two splats, leaves bins with
      counts below 108, whose DE radius round(19 / (1 + c/255)^0.4) exceeds
      16 super-pixels: a crop of that render takes DE's HBM-direct branch
      (tile radius above 16) and is checked against the oracle as for C0.
  C3  a sample of the 512-genome batch (motion-blurred and single frames, at
      their own sizes and qualities) through the batch driver against one
      render per job: images bit for bit.
  C4  a 1920x1080 quality-10 frame through the BF16-MFMA denoiser against the
      same network in PyTorch FP32 on the CPU.
"""
import numpy as np
import pytest

from pyr3_amd import pack, synth

pytestmark = pytest.mark.gpu

THREADS = 16


def _rel_close(a, b, rel=1e-5):
    """Element-wise |a - b| <= rel |b|, zeros matching exactly."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.abs(a - b) <= rel * np.abs(b)


def _blobs(g, seed, walkers, iters, fuse):
    e = pack.expand_genome_for_gpu(g)
    W, H, os_ = g.size["width"], g.size["height"], g.oversample
    u = pack.chaos_uniforms(e, W * os_, H * os_, os_, iters, fuse, seed, walkers)
    return pack.pack_xforms(e), pack.pack_xform_distrib(e), u, pack.pack_palette(e)


def _de_tonemap_parity(orc, R, g, hist, W, H, os_, samples):
    """GPU DE + tonemap of `hist` (super-res, W*os x H*os) against the oracle:
    per-element 1e-5 relative on the filtered grid and the float image,
    1 LSB on RGBA8. Returns the oracle's filtered grid."""
    SW, SH = W * os_, H * os_
    assert hist.shape == (SH, SW, 4)
    r = R(W, H, os_)
    r.set_image_f32(True)
    r.reset(g)
    r.write_histogram(hist)
    r.present(g, samples)
    fg, gf, g8 = r.read_filtered(), r.read_image_f32(), r.read_image()
    r.destroy()
    k1, k2 = pack.calibration(g.scale, samples, g.tonemap.brightness, os_)
    d = g.density
    fo = orc.density(hist, pack.de_uniforms(SW, SH, d.max_rad, d.min_rad, d.curve, k1, k2, os_),
                     THREADS)
    ok = _rel_close(fg, fo)
    assert ok.all(), (int((~ok).sum()), float(np.max(np.abs(fg - fo) / np.maximum(np.abs(fo), 1e-30))))
    taps = pack.spatial_kernel(0.5, os_)
    of, o8 = orc.tonemap(pack.viz_uniforms(W, H, k1, k2, g.tonemap, os_, len(taps)), taps,
                         filtered=fo, threads=THREADS)
    assert np.allclose(gf, of, rtol=1e-5, atol=1e-6)
    assert np.abs(g8.astype(int) - o8.astype(int)).max() <= 1
    return fo


def _invariants(h, samples, opacity_one=True, colour_le_count=True):
    cnt = h[..., 3].astype(np.int64)
    if opacity_one:
        assert (cnt % 255 == 0).all()
        assert cnt.sum() // 255 <= samples
    if colour_le_count:
        assert (h[..., :3].astype(np.int64) <= cnt[..., None]).all()
    assert (cnt > 0).mean() > 0.01


@pytest.fixture(scope="module")
def R():
    from pyr3_amd.renderer import Renderer
    return Renderer


def test_config0_reference_schedule_vs_oracle(orc, R):
    g = synth.config0_genome(0)
    W, H = 256, 256
    walkers, iters, samples = pack.compute_dispatch(g.quality, W, H, None)
    assert (walkers, iters) == (800, 4096)
    fuse, seed = 20, 0x9E3779B9
    r = R(W, H, 1)
    r.reset(g)
    r.iterate(g, seed, walkers, iters, fuse=fuse)
    hg = r.read_histogram()
    r.destroy()
    ho = orc.chaos_mt(*_blobs(g, seed, walkers, iters, fuse), THREADS)
    _invariants(hg, samples)
    sg, so = hg.reshape(-1, 4).sum(0).astype(np.float64), ho.reshape(-1, 4).sum(0).astype(np.float64)
    assert np.all(np.abs(sg - so) <= 0.01 * so), (sg, so)
    bg = hg[..., 3].reshape(16, 16, 16, 16).sum((1, 3)).astype(np.float64)
    bo = ho[..., 3].reshape(16, 16, 16, 16).sum((1, 3)).astype(np.float64)
    assert np.abs(bg - bo).sum() <= 0.03 * bo.sum()
    _de_tonemap_parity(orc, R, g, hg, W, H, 1, samples)


def test_config1_full_size(orc, R):
    g = synth.config1_genome(0)
    W, H, os_ = 1024, 768, 2
    samples_target = int(g.quality * W * H)
    walkers = 131072
    iters = -(-samples_target // walkers)
    r = R(W, H, os_)
    hs = []
    for _ in range(2):
        r.reset(g)
        r.iterate(g, 0x51ED, walkers, iters, fuse=pack.FUSE)
        hs.append(r.read_histogram())
    r.destroy()
    hg = hs[0]
    assert np.array_equal(hs[0], hs[1])
    _invariants(hg, walkers * iters)
    # DE / tonemap of a crop around the densest block (256x192 output pixels)
    cnt = hg[..., 3].astype(np.int64)
    by, bx = np.unravel_index(np.argmax(cnt.reshape(6, 256, 8, 256).sum((1, 3))), (6, 8))
    y0, x0 = min(by * 256, 2 * H - 384), min(bx * 256, 2 * W - 512)
    crop = np.ascontiguousarray(hg[y0:y0 + 384, x0:x0 + 512])
    _de_tonemap_parity(orc, R, g, crop, 256, 192, os_, walkers * iters)


def _render_4k(R, g, walkers=131072, seed=0xC2):
    W, H, os_ = 3840, 2160, 2
    iters = -(-int(g.quality * W * H) // walkers)
    r = R(W, H, os_)
    r.set_image_f32(True)
    r.reset(g)
    s = r.iterate(g, seed, walkers, iters, fuse=pack.FUSE)
    h = r.read_histogram()
    r.present(g, s)
    img = r.read_image_f32()
    filt = r.read_filtered()
    r.destroy()
    return h, img, filt, s


def test_config2_4k_full_catalog(orc, R):
    g = synth.config2_genome(0)
    h, img, filt, s = _render_4k(R, g)
    assert h.shape == (4320, 7680, 4)
    _invariants(h, s, colour_le_count=False)
    assert np.isfinite(img).all() and np.isfinite(filt).all()
    assert (img[..., 3] == 1.0).all()


def test_config2_low_opacity_takes_the_wide_de_branch(orc, R):
    import copy
    g = copy.deepcopy(synth.config2_genome(0))
    rng = np.random.default_rng(2002)
    for x in g.xforms:
        x.opacity = float(rng.uniform(0.05, 1.0))
    g.quality = 2
    h, img, filt, s = _render_4k(R, g)
    assert np.isfinite(img).all() and np.isfinite(filt).all()
    cnt = h[..., 3].astype(np.int64)
    # a 192x128 crop (384x256 super-res) holding bins with counts below 108,
    # whose radius round(19 / (1 + c/255)^0.4) exceeds 16
    lo = (cnt > 0) & (cnt < 108)
    assert lo.sum() > 1000
    blk = lo[:4096, :7680].reshape(16, 256, 20, 384).sum((1, 3))
    by, bx = np.unravel_index(np.argmax(blk), blk.shape)
    crop = np.ascontiguousarray(h[by * 256:by * 256 + 256, bx * 384:bx * 384 + 384])
    c = crop[..., 3].astype(np.float64) / 255.0
    rad = np.rint(np.clip(19.0 / (c + 1.0) ** 0.4, 1.0, 19.0))
    assert (rad[crop[..., 3] > 0] > 16).any()
    _de_tonemap_parity(orc, R, g, crop, 192, 128, 2, s)


def test_config3_batch_sample_matches_single_renders(R):
    """Eight jobs of the 512-genome batch (four motion-blurred with 8
    sub-frames) as one union-kernel launch on a captured graph, at the
    batch's default plan: every frame equals a renderer's render of it."""
    from pyr3_amd import batch
    jobs = batch.batch_workload(512)
    sample = [jobs[i] for i in SAMPLE_JOBS]
    assert sum(j.subframes is not None for j in sample) == 4
    gb = batch.GraphBatch(sample, batch.BatchPlan())
    try:
        gb.launch()
        gb.sync()
        for i, j in enumerate(sample):
            r = R(j.width, j.height, 1)
            r.reset(j.genome)
            walkers, iters = gb.walkers[i], gb.iters[i]
            if j.subframes:
                k = len(j.subframes)
                counts = batch.apportion(walkers, [1.0] * k)
                s = r.iterate_multi(j.subframes,
                                    [(j.seed + q * batch.SUBFRAME_SEED_STEP) & 0xFFFFFFFF
                                     for q in range(k)], counts, iters)
            else:
                s = r.iterate(j.genome, j.seed, walkers, iters)
            r.present(j.genome, s)
            assert gb.samples[i] == s
            assert np.array_equal(gb.read_histogram(i), r.read_histogram()), j.name
            assert np.array_equal(gb.read_image(i), r.read_image()), j.name
            r.destroy()
    finally:
        gb.close()


# jobs of the 512-genome batch in this test: four motion-blurred (8
# sub-frames each) and four single frames
SAMPLE_JOBS = (0, 1, 2, 3, 256, 257, 258, 259)


def test_config4_learned_frame_1080p():
    """The 1080p learned frame (q4 input, short fuse, DE off, the shipped
    weights) against the network in PyTorch FP32 on the same features."""
    import torch
    from pyr3_amd import learned
    from pyr3_amd.denoise.weights import flat_params, load
    from pyr3_amd.renderer import Denoiser, Renderer
    from test_gpu_denoise import reference_output
    if not WEIGHTS.exists():
        pytest.skip("no trained weights in the tree")
    W, H = 1920, 1080
    g = synth.bench_genome(0, W, H)
    r = Renderer(W, H, 1)
    r.set_image_f32(True)
    walkers, iters = learned.dispatch(4, W, H)
    r.reset(g)
    s = r.iterate(g, 0x4C, walkers, iters, fuse=learned.FUSE)
    r.present(g, s, force_de_off=True)
    p = load()
    torch.set_num_threads(THREADS)
    ref, _ = reference_output(p, r, g, s)
    d = Denoiser(flat_params(p))
    d.apply(r)
    got = r.read_image()
    d.destroy()
    r.destroy()
    want = np.floor(ref * 255.0 + 0.5).astype(np.int32)
    diff = np.abs(got[..., :3].astype(np.int32) - want)
    assert diff.max() <= 3 and diff.mean() < 0.35, (diff.max(), diff.mean())


from pyr3_amd.denoise.weights import DEFAULT as WEIGHTS  # noqa: E402
