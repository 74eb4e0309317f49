"""Motion-blurred animation frames on the GPU (pyr3_amd.animate.render_animation_frame).

* The frame's histogram equals the oracle's, sub-frame by sub-frame summed,
  bit for bit: the genomes are linear (exact arithmetic) and every
  sub-frame's genome, seed and walker count come from the frame result.
* Sub-frames that straddle a keyframe (two kernel shapes) go out as batched
  launches per shape; the histogram equals one `iterate` per sub-frame.
* One temporal sample is one dispatch of the centre genome.
"""
import numpy as np
import pytest

from helpers import blobs, genome, sierpinski, xform
from pyr3_amd import animate as A
from pyr3_amd import pack
from pyr3_amd.catalog import VARIATION_INDEX as V
from pyr3_amd.genome import Density

pytestmark = pytest.mark.gpu


def _keyframe(time, shift=0.0, extra=None, quality=40.0):
    xs = sierpinski()
    for x in xs:
        x.affine.c += shift
    if extra is not None:
        xs.append(extra)
    g = genome(xs, scale=20.0, cx=0.5, cy=0.5)
    g.time = time
    g.quality = quality
    g.density = Density(3.0, 0.0, 0.4)
    return g


def test_motion_blurred_frame_matches_oracle(orc):
    from pyr3_amd.renderer import Renderer
    W, H = 48, 40
    an = A.Animation([_keyframe(0.0), _keyframe(1.0, 0.1)], ntemporal_samples=5,
                     temporal_filter_type="gaussian", temporal_filter_width=0.6)
    r = Renderer(W, H, 1)
    res = A.render_animation_frame(r, an, 0.5, seed=2024, walkers=1024)
    h = r.read_histogram()
    r.destroy()
    assert res.subframes == 5 and sum(res.walkers) == 1024 and len(res.genomes) >= 4
    iters = res.total_samples // res.total_walkers
    print(f"frame: walkers {res.walkers} iters {iters}", flush=True)
    ho = np.zeros_like(h)
    for g, sd, w in zip(res.genomes, res.seeds, res.walkers):
        xf, ds, u, pal = blobs(g, W, H, iters=iters, fuse=pack.FUSE, seed=sd, walkers=w)
        orc.chaos(xf, ds, u, pal, hist=ho)
    assert h[..., 3].sum() > 0
    assert np.array_equal(h, ho)


def test_frame_across_a_keyframe_equals_sequential_dispatches():
    from pyr3_amd.renderer import Renderer
    W, H = 40, 40
    extra = xform(0.4, (0.3, 0.0, 0.2, 0.0, 0.3, 0.1),
                  vars=((V["spherical"], 0.6), (V["linear"], 0.4)), color=0.7)
    an = A.Animation([_keyframe(0.0), _keyframe(1.0, 0.05), _keyframe(2.0, -0.05, extra)],
                     ntemporal_samples=8, temporal_filter_width=1.0)
    r = Renderer(W, H, 1)
    res = A.render_animation_frame(r, an, 1.0, seed=77, walkers=2048)
    hb = r.read_histogram()
    r.destroy()
    shapes = {A.packed_shape(pack_g) for pack_g in map(_packed, res.genomes)}
    assert len(shapes) == 2      # sub-frames on both sides of the middle keyframe
    iters = res.total_samples // res.total_walkers
    rs = Renderer(W, H, 1)
    rs.reset(res.center)
    for g, sd, w in zip(res.genomes, res.seeds, res.walkers):
        rs.iterate(g, sd, w, iters)
    hs = rs.read_histogram()
    rs.destroy()
    assert hb[..., 3].sum() > 0
    assert np.array_equal(hb, hs)


def _packed(g):
    from pyr3_amd.renderer import pack_genome
    return pack_genome(g)


def test_single_temporal_sample_is_one_dispatch():
    from pyr3_amd.renderer import Renderer
    W, H = 32, 32
    an = A.Animation([_keyframe(0.0), _keyframe(1.0, 0.1)], ntemporal_samples=1)
    r = Renderer(W, H, 1)
    res = A.render_animation_frame(r, an, 0.25, seed=5, walkers=512)
    img = r.read_image()
    r.destroy()
    assert res.subframes == 1 and res.seeds == [5]
    center = A.interpolate(an, 0.25)
    walkers, iters, samples = pack.compute_dispatch(center.quality, W, H, 512)
    r2 = Renderer(W, H, 1)
    r2.reset(center)
    r2.iterate(center, 5, walkers, iters)
    r2.present(center, samples)
    img2 = r2.read_image()
    r2.destroy()
    assert np.array_equal(img, img2)
