"""Oracle DE and tonemap vs closed forms, invariants and an f64 scatter form.

DE (src/shaders/density.wgsl:51-123) is restated in the oracle as the
reference's per-output gather; here it is checked against
  * an all-zero histogram -> zero output,
  * conservation: an isolated bucket deposits ls * (R, G, B, count) in total
    (each bucket's kernel sums to 1, src/density.test.ts:61-71),
  * an independent f64 *scatter* formulation (each bucket spreads over its own
    disc), which catches using the centre's radius instead of the neighbour's.
The tonemap (src/shaders/visualize_f32.wgsl:96-271, visualize_u32.wgsl) is
checked against empty -> background, vibrancy 0 -> pure per-channel gamma,
and an f64 numpy restatement of flam3's alpha/vibrancy composite.
"""
import math

import numpy as np
import pytest

from pyr3_amd import pack
from pyr3_amd.genome import Tonemap


def _de_u(W, H, max_rad=9.0, min_rad=0.0, curve=0.4, k1=267.0, k2=1e-3, os_=1):
    return pack.de_uniforms(W, H, max_rad, min_rad, curve, k1, k2, os_)


def _scatter_f64(hist, max_rad_s, min_rad_s, curve, k1, k2, lut):
    H, W, _ = hist.shape
    out = np.zeros((H, W, 4))
    R = math.ceil(min(max(max_rad_s, 0), 256))
    for y in range(H):
        for x in range(W):
            c = float(hist[y, x, 3])
            if c <= 0:
                continue
            hits = np.float32(c / 255.0)
            rf = min(max(max_rad_s / (float(hits) + 1.0) ** curve, min_rad_s), max_rad_s)
            r = int(min(max(round(rf), 0), 256))  # Python round() is half-even
            s = max(r / 3.0, 1e-6)
            ls = k1 * math.log(1.0 + c * k2) / c
            for dy in range(-min(r, R), min(r, R) + 1):
                for dx in range(-min(r, R), min(r, R) + 1):
                    d2 = dx * dx + dy * dy
                    if d2 > r * r or not (0 <= y + dy < H and 0 <= x + dx < W):
                        continue
                    w = math.exp(-d2 / (2 * s * s)) / max(float(lut[r]), 1e-6) * ls
                    out[y + dy, x + dx] += hist[y, x].astype(np.float64) * w
    return out


def test_de_zero_histogram(orc):
    h = np.zeros((20, 24, 4), dtype=np.uint32)
    assert not orc.density(h, _de_u(24, 20)).any()


@pytest.mark.parametrize("count", [255, 255 * 7, 255 * 1000])
def test_de_isolated_bucket_conserves_mass(orc, count):
    W = H = 48
    h = np.zeros((H, W, 4), dtype=np.uint32)
    h[24, 24] = (count // 2, count // 3, count // 5, count)
    u = _de_u(W, H, max_rad=9.0, k1=300.0, k2=2e-4)
    f = orc.density(h, u).astype(np.float64)
    ls = 300.0 * math.log(1.0 + count * np.float32(2e-4)) / count
    tot = f.reshape(-1, 4).sum(0)
    want = h[24, 24].astype(np.float64) * ls
    assert np.allclose(tot, want, rtol=2e-5)


def test_de_matches_f64_scatter(orc):
    rng = np.random.default_rng(3)
    H, W = 26, 31
    h = np.zeros((H, W, 4), dtype=np.uint32)
    mask = rng.random((H, W)) < 0.35
    cnt = (rng.integers(1, 400, (H, W)) * 255).astype(np.uint32)
    h[..., 3] = np.where(mask, cnt, 0)
    for c in range(3):
        h[..., c] = (h[..., 3] * rng.random((H, W))).astype(np.uint32)
    max_rad, min_rad, curve, k1, k2 = 5.0, 0.0, 0.4, 273.0, 3e-4
    u = _de_u(W, H, max_rad, min_rad, curve, k1, k2, os_=2)
    f = orc.density(h, u).astype(np.float64)
    lut = orc.kernel_norm_lut()
    ref = _scatter_f64(h, float(np.float32(max_rad * 2 + 1)), float(np.float32(min_rad * 2 + 1)),
                       curve, float(np.float32(k1)), float(np.float32(k2)), lut)
    assert np.allclose(f, ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())


def _viz(W, H, tm, os_, fwidth, k1=300.0, k2=1e-3, bg=(0.1, 0.2, 0.3), transparent=False):
    return pack.viz_uniforms(W, H, k1, k2, tm, os_, fwidth, bg, transparent)


def test_tonemap_empty_is_background(orc):
    W, H = 8, 6
    f = np.zeros((H, W, 4), dtype=np.float32)
    taps = pack.spatial_kernel(0.5, 1)
    of, o8 = orc.tonemap(_viz(W, H, Tonemap(), 1, len(taps)), taps, filtered=f)
    assert np.allclose(of[..., :3], np.float32([0.1, 0.2, 0.3])) and np.all(of[..., 3] == 1)
    of2, _ = orc.tonemap(_viz(W, H, Tonemap(), 1, len(taps), transparent=True), taps, filtered=f)
    assert not of2.any()


def _hsv(r, g, b):
    mx, mn = max(r, g, b), min(r, g, b)
    d = mx - mn
    h = 0.0
    if d > 0:
        h = ((g - b) / d if mx == r else (2 + (b - r) / d if mx == g else 4 + (r - g) / d)) * 60
        h = h + 360 if h < 0 else h
    return h, (d / mx if mx > 0 else 0.0), mx


def _rgb(h, s, v):
    hh = h / 60.0
    i = math.floor(hh)
    f = hh - i
    p, q, t = v * (1 - s), v * (1 - s * f), v * (1 - s * (1 - f))
    return [(v, t, p), (q, v, p), (p, v, t), (p, q, v), (t, p, v), (v, p, q)][int(math.fmod(i, 6))]


def _newrgb(c, ls, hp):
    """flam3 calc_newrgb: highlight-power desaturation above white."""
    if ls == 0 or not any(c):
        return [0.0, 0.0, 0.0]
    cw = [v / 255.0 for v in c]
    k = max(range(3), key=lambda i: (ls * cw[i], -i))
    maxa, maxc = ls * cw[k], cw[k]
    newls = 255.0 / maxc
    if maxa > 255.0 and hp >= 0:
        h, s, v = _hsv(*[newls * x / 255.0 for x in cw])
        return [x * 255.0 for x in _rgb(h, s * (newls / ls) ** hp, v)]
    adj = 1.0 if maxa <= 255.0 else min(-hp, 1.0)
    return [((1 - adj) * newls + adj * ls) * x for x in cw]


def _tonemap_f64(c, tm, bg):
    """flam3 alpha / vibrancy / highlight composite on collapsed RGBA c, f64."""
    g = 1.0 / tm.gamma
    lin = tm.gamma_threshold
    H, W, _ = c.shape
    out = np.zeros((H, W, 3))
    for y in range(H):
        for x in range(W):
            a = c[y, x, 3] / 255.0
            if c[y, x, 3] <= 0:
                out[y, x] = bg
                continue
            al = ((1 - a / lin) * a * (lin ** g / lin) + (a / lin) * a ** g) if a < lin else a ** g
            ls = tm.vibrancy * 256.0 * al / max(a, 1e-12)
            al = min(max(al, 0.0), 1.0)
            n = _newrgb(list(c[y, x, :3]), ls, tm.highlight_power)
            for i in range(3):
                per = (1 - tm.vibrancy) * 256.0 * max(c[y, x, i] / 255.0, 0) ** g
                out[y, x, i] = min(max((n[i] + per + (1 - al) * 256.0 * bg[i]) / 256.0, 0), 1)
    return out


@pytest.mark.parametrize("vib,hp", [(0.0, 1.0), (0.4, 1.0), (1.0, 1.0), (1.0, -0.5), (1.0, 3.0)])
def test_tonemap_matches_f64_composite(orc, vib, hp):
    rng = np.random.default_rng(5)
    W, H = 12, 9
    f = (rng.random((H, W, 4)) * 40).astype(np.float32)
    f[..., :3] *= 0.5
    f[2, 3] = 0
    tm = Tonemap(gamma=4.0, vibrancy=vib, highlight_power=hp, brightness=1.0,
                 gamma_threshold=0.01)
    taps = np.array([0.0, 1.0, 0.0], dtype=np.float32)  # identity collapse
    bg = (0.1, 0.2, 0.3)
    of, o8 = orc.tonemap(_viz(W, H, tm, 1, 3, bg=bg), taps, filtered=f)
    want = _tonemap_f64(f.astype(np.float64), tm, bg)
    assert np.allclose(of[..., :3], want, rtol=1e-5, atol=2e-6)
    assert np.array_equal(o8[..., :3], np.floor(np.clip(of[..., :3], 0, 1) * 255 + 0.5))


def test_tonemap_vibrancy_zero_is_per_channel_gamma(orc):
    W, H = 4, 3
    f = np.zeros((H, W, 4), dtype=np.float32)
    f[...] = (100.0, 50.0, 25.0, 255.0 * 0.5)
    tm = Tonemap(gamma=2.0, vibrancy=0.0, gamma_threshold=0.0)
    taps = np.array([0.0, 1.0, 0.0], dtype=np.float32)
    of, _ = orc.tonemap(_viz(W, H, tm, 1, 3, bg=(0, 0, 0)), taps, filtered=f)
    want = np.sqrt(np.array([100.0, 50.0, 25.0]) / 255.0)
    assert np.allclose(of[0, 0, :3], want, rtol=1e-6)


def test_tonemap_de_off_applies_log_density_per_superpixel(orc):
    """DE-off collapse: S = (R, G, B, count) * ls(count) before the taps."""
    rng = np.random.default_rng(9)
    W, H, os_ = 6, 5, 2
    h = np.zeros((H * os_, W * os_, 4), dtype=np.uint32)
    h[..., 3] = rng.integers(0, 50, (H * os_, W * os_)) * 255
    h[..., 0] = h[..., 3] // 2
    taps = pack.spatial_kernel(0.5, os_)
    tm = Tonemap(gamma=3.0, vibrancy=1.0)
    k1, k2 = 280.0, 2e-3
    u = _viz(W, H, tm, os_, len(taps), k1=k1, k2=k2, bg=(0, 0, 0))
    of_u, _ = orc.tonemap(u, taps, hist=h)
    # same thing through the f32 path on an explicitly log-scaled grid
    c = h[..., 3].astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        ls = np.where(c > 0, np.float32(k1) * np.log1p(c * np.float32(k2)) / c, 0).astype(np.float32)
    f = (h.astype(np.float32) * ls[..., None]).astype(np.float32)
    of_f, _ = orc.tonemap(u, taps, filtered=f)
    assert np.allclose(of_u, of_f, rtol=1e-5, atol=1e-6)
