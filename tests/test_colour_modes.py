"""Splat colour modes and colour-index capture in the oracle
(chaos.wgsl:8625-8697, 8720-8728; hsv_to_rgb :1662-1666), pinned by
configurations whose colours are known in closed form:

* flow: two translations x -> x +/- 1 (y -> 0, no jitter): every splat moved
  by exactly +/-1 along x, so its hue is cyan (+x, atan2 = 0) or red (-x,
  atan2 = pi) at value 1 - exp(-2): each splat adds (0, 220, 220) or
  (220, 0, 0) with count 255;
* phase: an attractor on the positive real axis has arg z = 0 (cyan, its
  green one LSB short from the f32 arithmetic of hsv_to_rgb), on the negative
  one arg z = pi (red); ring frequency 0 gives value 1;
* trap distance to the line y = 0.5 is 0.5 for every point of the x axis:
  glow exp(-2 * 0.5) picks palette entry 93, rings fract(0.5 * 3) entry 127;
* capture with colour speed 1: each splat adds u32(colour * 255) of its xform;
* strength 0, or colour mode 0 with any parameters, leaves the histogram of
  the palette path unchanged.
"""
import numpy as np
import pytest

from helpers import genome, sierpinski, xform
from pyr3_amd import pack

W = H = 48


def _run(orc, g, color, iters=60, fuse=10, walkers=256, seed=9, capture=False):
    e = pack.expand_genome_for_gpu(g)
    xf, ds, pal = pack.pack_xforms(e), pack.pack_xform_distrib(e), pack.pack_palette(e)
    u = pack.chaos_uniforms(e, W, H, 1, iters, fuse, seed, walkers, jitter=0.0, color=color)
    if capture:
        return orc.chaos_capture(xf, ds, u, pal) + (pal,)
    h, _ = orc.chaos(xf, ds, u, pal)
    return h, pal


def _axis_genome(sign=1.0):
    """Two contractions onto [0.25, 1] (or its mirror) on the x axis."""
    return genome([xform(1.0, (0.5, 0, 0.5 * sign, 0, 0, 0), color=0.2),
                   xform(1.0, (0.5, 0, 0.25 * sign, 0, 0, 0), color=0.8)],
                  scale=20.0, cx=0.6 * sign)


def test_flow_translations_are_cyan_or_red(orc):
    g = genome([xform(1.0, (1, 0, 1, 0, 0, 0)), xform(1.0, (1, 0, -1, 0, 0, 0))], scale=5.0)
    h, _ = _run(orc, g, pack.ColorOpts(mode="flow"))
    n = h[..., 3] // 255
    assert n.sum() > 1000
    assert np.all(h[..., 3] % 255 == 0)
    r, gg, b = h[..., 0], h[..., 1], h[..., 2]
    assert np.array_equal(gg, b)
    assert np.all(r % 220 == 0) and np.all(gg % 220 == 0)
    assert np.array_equal(r // 220 + gg // 220, n)
    assert (r > 0).any() and (gg > 0).any()


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_phase_on_the_real_axis(orc, sign):
    h, _ = _run(orc, _axis_genome(sign), pack.ColorOpts(mode="phase", phase_freq=0.0))
    n = h[..., 3] // 255
    assert n.sum() > 1000
    if sign > 0:   # cyan at full value (f32: fract(0.5 + 2/3) * 6 - 3 = -1.9999995 -> G 254)
        assert not h[..., 0].any() and np.array_equal(h[..., 2], 255 * n)
        assert np.all((h[..., 1] >= 254 * n) & (h[..., 1] <= 255 * n))
    else:          # red
        assert np.array_equal(h[..., 0], 255 * n) and not h[..., 1:3].any()
    # log-modulus rings only change the value: hue channels stay as they were
    h2, _ = _run(orc, _axis_genome(sign), pack.ColorOpts(mode="phase", phase_freq=1.0))
    assert np.array_equal(h2[..., 3], h[..., 3])
    if sign > 0:
        assert not h2[..., 0].any() and np.all(h2[..., 1] <= h2[..., 2])
        assert (h2[..., 2] < h[..., 2]).any() and (h2[..., 2] >= 127 * n).all()


@pytest.mark.parametrize("mode,entry", [("glow", 93), ("rings", 127)])
def test_trap_line_at_constant_distance(orc, mode, entry):
    c = pack.ColorOpts(mode="trap-distance", trap_kind="line", trap_mode=mode, trap_cy=0.5,
                       trap_angle=0.0, trap_falloff=2.0, trap_freq=3.0)
    h, pal = _run(orc, _axis_genome(), c)
    n = h[..., 3].astype(np.int64) // 255
    assert n.sum() > 1000
    for ch in range(3):
        per = int(np.float32(pal[entry, ch]) * np.float32(255.0))
        assert np.array_equal(h[..., ch], per * n), ch


def test_capture_sums_colour_of_last_xform(orc):
    g = genome([xform(1.0, (0.5, 0, 0, 0, 0.5, 0), color=0.25, speed=1.0),
                xform(1.0, (0.5, 0, 0.5, 0, 0.5, 0.5), color=0.75, speed=1.0)],
               scale=20.0, cx=0.5, cy=0.5)
    h, idx, _ = _run(orc, g, pack.ColorOpts(capture_index=True), capture=True)
    n = h[..., 3].astype(np.int64) // 255
    assert n.sum() > 1000
    idx = idx.astype(np.int64)
    a = (191 * n - idx)          # splats of the 0.25 xform, times 128
    assert np.all(a % 128 == 0)
    a //= 128
    assert np.all((a >= 0) & (a <= n)) and a.sum() > 0 and (n - a).sum() > 0
    h0, _ = _run(orc, g, None)
    assert np.array_equal(h, h0)     # capture leaves the histogram alone


@pytest.mark.parametrize("opts", [
    pack.ColorOpts(mode="flow", flow_strength=0.0),
    pack.ColorOpts(mode="trap-distance", trap_kind="circle", trap_strength=0.0),
    pack.ColorOpts(mode="phase", phase_strength=0.0, phase_freq=3.0),
    pack.ColorOpts(mode="palette", flow_strength=0.3, trap_kind="line", trap_strength=0.7,
                   phase_freq=9.0),
])
def test_no_op_settings_keep_the_palette_histogram(orc, opts):
    g = genome(sierpinski(), scale=20.0, cx=0.5, cy=0.5)
    h, _ = _run(orc, g, opts)
    h0, _ = _run(orc, g, None)
    assert h[..., 3].sum() > 0 and np.array_equal(h, h0)


def test_uniform_slots():
    g = genome(sierpinski())
    u = pack.chaos_uniforms(g, 8, 8, 1, 1, 1, 1, 64, color=pack.ColorOpts(
        mode="trap-distance", trap_kind="line", trap_mode="rings", trap_angle=90.0,
        capture_index=True))
    f = u.view(np.float32)
    assert (u[14], u[15], u[18], u[19]) == (1, 2, 2, 1)
    assert np.isclose(f[23], -1.0) and abs(f[24]) < 1e-7      # normal (-sin, cos)
    d = pack.chaos_uniforms(g, 8, 8, 1, 1, 1, 1, 64)
    assert (d[14], d[15]) == (0, 0) and d.view(np.float32)[17] == 2.0
    with pytest.raises(ValueError):
        pack.chaos_uniforms(g, 8, 8, 1, 1, 1, 1, 64, color=pack.ColorOpts(mode="rainbow"))
