"""Oracle tonemap grading (channel curves, HSV adjust) vs closed forms.

The grading block (src/shaders/visualize_f32.wgsl:240-266, LUT lookup :63-68)
runs after the composite on covered pixels only. Checked here against:
  * an identity LUT on every channel, and HSV (0, x1, +0), leave the image
    unchanged; hue +360 equals hue 0;
  * a hue turn of +120 degrees cycles the channels (R -> G -> B -> R);
  * saturation 0 gives grey at the pixel's max channel; lightness +1 gives a
    max channel of 1; lightness -1 gives black;
  * a channel curve maps that channel through the spline (within the
    256-entry LUT's interpolation error) and leaves the others alone; the
    composite curve applies after the per-channel ones;
  * a constant luma curve c sets every unclipped pixel's Rec.601 luma to c;
  * empty pixels keep the background.
"""
import numpy as np
import pytest

from pyr3_amd import curves as cv
from pyr3_amd import pack
from pyr3_amd.genome import Tonemap

W, H = 14, 11
TAPS = np.array([0.0, 1.0, 0.0], dtype=np.float32)
BG = (0.1, 0.2, 0.3)


def _grid():
    rng = np.random.default_rng(31)
    f = (rng.random((H, W, 4)) * 60).astype(np.float32)
    f[..., :3] *= 0.5
    f[3, 4] = 0
    f[7, 1] = 0
    return f


def _run(orc, mask=0, hsl=None, lut=None):
    tm = Tonemap(gamma=3.0, vibrancy=0.7, highlight_power=1.0)
    u = pack.viz_uniforms(W, H, 1.0, 1.0, tm, 1, 3, BG, False, mask, hsl)
    return orc.tonemap(u, TAPS, filtered=_grid(), curves=lut)[0]


def _identity_lut():
    return np.tile((np.arange(256) / 255.0).astype(np.float32), (5, 1))


def _covered():
    return _grid()[..., 3] > 0


def test_viz_uniform_words():
    u = pack.viz_uniforms(8, 8, 1, 1, Tonemap(), 1, 3, curves_mask=0b10110,
                          hsl_adjust={"hue": -30.0, "sat": 150.0, "light": -20.0})
    f = u.view(np.float32)
    assert u[16] == 0b10110 and u[17] == 1
    assert (f[18], f[19], f[20]) == pytest.approx((-30.0, 1.5, -0.2))
    assert pack.viz_uniforms(8, 8, 1, 1, Tonemap(), 1, 3)[16:21].tolist() == [0] * 5


def test_mask_without_lut_is_refused(orc):
    with pytest.raises(ValueError):
        _run(orc, mask=1)


def test_identity_grading_leaves_the_image(orc):
    base = _run(orc)
    assert np.allclose(_run(orc, 31, lut=_identity_lut()), base, atol=2e-6)
    assert np.allclose(_run(orc, hsl={"hue": 0.0, "sat": 100.0, "light": 0.0}), base, atol=2e-6)
    h0 = _run(orc, hsl={"hue": 0.0, "sat": 100.0, "light": 0.0})
    assert np.allclose(_run(orc, hsl={"hue": 360.0, "sat": 100.0, "light": 0.0}), h0, atol=2e-5)


def test_hue_turn_cycles_channels(orc):
    base = _run(orc)
    rot = _run(orc, hsl={"hue": 120.0, "sat": 100.0, "light": 0.0})
    cov = _covered()
    assert np.allclose(rot[..., 1][cov], base[..., 0][cov], atol=2e-5)
    assert np.allclose(rot[..., 2][cov], base[..., 1][cov], atol=2e-5)
    assert np.allclose(rot[..., 0][cov], base[..., 2][cov], atol=2e-5)
    assert np.array_equal(rot[~cov], base[~cov])       # background untouched


def test_saturation_and_lightness(orc):
    base = _run(orc)
    cov = _covered()
    grey = _run(orc, hsl={"hue": 0.0, "sat": 0.0, "light": 0.0})
    mx = base[..., :3].max(axis=-1)
    for c in range(3):
        assert np.allclose(grey[..., c][cov], mx[cov], atol=1e-6)
    bright = _run(orc, hsl={"hue": 0.0, "sat": 100.0, "light": 100.0})
    assert np.allclose(bright[..., :3].max(axis=-1)[cov], 1.0)
    dark = _run(orc, hsl={"hue": 0.0, "sat": 100.0, "light": -100.0})
    assert np.all(dark[..., :3][cov] == 0)


def test_channel_and_composite_curves(orc):
    base = _run(orc)
    cov = _covered()
    pts_r = [(0.0, 0.0), (0.3, 0.55), (1.0, 1.0)]
    pts_c = [(0.0, 0.1), (0.6, 0.5), (1.0, 0.9)]
    c = cv.identity_curves()
    c["r"] = pts_r
    lut = cv.bake_curves(c)
    out = _run(orc, cv.active_mask(c), lut=lut)
    want_r = np.array([cv.eval_curve(pts_r, float(v)) for v in base[..., 0][cov]])
    # the spline's slope jumps at a control point (segments of unequal width,
    # each on u in [0, 1]); a 256-entry linear LUT is off there by at most
    # (slope jump) x (1/255) / 4 ~ 1.2e-3 for these points
    assert np.allclose(out[..., 0][cov], want_r, atol=1.5e-3)
    far = np.abs(base[..., 0][cov] - 0.3) > 2 / 255
    assert np.allclose(out[..., 0][cov][far], want_r[far], atol=1e-4)
    assert np.array_equal(out[..., 1:3], base[..., 1:3])
    c["composite"] = pts_c
    out = _run(orc, cv.active_mask(c), lut=cv.bake_curves(c))
    want_r = np.array([cv.eval_curve(pts_c, cv.eval_curve(pts_r, float(v)))
                       for v in base[..., 0][cov]])
    want_g = np.array([cv.eval_curve(pts_c, float(v)) for v in base[..., 1][cov]])
    assert np.allclose(out[..., 0][cov], want_r, atol=3e-3)
    assert np.allclose(out[..., 1][cov], want_g, atol=1.5e-3)
    assert np.array_equal(out[~cov], base[~cov])


def test_constant_luma_curve_fixes_luma(orc):
    base = _run(orc)
    lut = _identity_lut()
    lut[4, :] = 0.4
    out = _run(orc, 16, lut=lut)
    luma = out[..., 0] * 0.299 + out[..., 1] * 0.587 + out[..., 2] * 0.114
    l0 = base[..., 0] * 0.299 + base[..., 1] * 0.587 + base[..., 2] * 0.114
    ok = _covered() & (l0 > 0) & (base[..., :3].max(axis=-1) * 0.4 / np.maximum(l0, 1e-9) <= 1)
    assert ok.sum() > 20
    assert np.allclose(luma[ok], 0.4, atol=1e-5)
    # hue is kept: channel ratios unchanged
    assert np.allclose(out[..., 0][ok] * base[..., 1][ok], out[..., 1][ok] * base[..., 0][ok],
                       atol=1e-5)
