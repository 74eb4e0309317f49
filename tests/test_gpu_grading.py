"""Tonemap grading on the GPU: channel curves and HSV adjust vs the oracle.

Same bars as the plain tonemap (f32 within 1e-5, RGBA8 within 1 LSB), with
DE on and off, on a random histogram uploaded to both sides. Also the C ABI's
refusals: a curves mask without an uploaded LUT, a LUT of the wrong size or
with values outside [0, 1].
"""
import numpy as np
import pytest

from helpers import genome, sierpinski
from pyr3_amd import curves as cv
from pyr3_amd import pack
from pyr3_amd.genome import Tonemap

pytestmark = pytest.mark.gpu


def _hist(W, H, os_, seed):
    rng = np.random.default_rng(seed)
    SW, SH = W * os_, H * os_
    h = np.zeros((SH, SW, 4), dtype=np.uint32)
    dense = rng.random((SH, SW)) < 0.6
    h[..., 3] = np.where(dense, (rng.integers(1, 2000, (SH, SW)) * 255).astype(np.uint32), 0)
    for c in range(3):
        h[..., c] = (h[..., 3] * rng.random((SH, SW))).astype(np.uint32)
    return h


@pytest.mark.parametrize("os_,curves_on,hsl", [
    (1, True, None),
    (1, False, {"hue": 75.0, "sat": 140.0, "light": -10.0}),
    (2, True, {"hue": -150.0, "sat": 60.0, "light": 15.0}),
])
def test_grading_matches_oracle(orc, os_, curves_on, hsl):
    from pyr3_amd.renderer import Renderer
    W, H = 60, 40
    h = _hist(W, H, os_, 40 + os_)
    g = genome(sierpinski(), scale=30.0)
    g.tonemap = Tonemap(gamma=3.5, vibrancy=0.8, highlight_power=0.5)
    g.background = [0.05, 0.1, 0.2]
    if curves_on:
        c = cv.identity_curves()
        c["composite"] = [(0.0, 0.05), (0.5, 0.6), (1.0, 0.95)]
        c["b"] = [(0.0, 0.0), (0.4, 0.2), (0.8, 0.9), (1.0, 1.0)]
        c["luma"] = [(0.0, 0.0), (0.5, 0.45), (1.0, 1.0)]
        g.channel_curves = c
    g.hsl_adjust = hsl
    samples = 4e6
    r = Renderer(W, H, os_)
    r.set_image_f32(True)
    r.reset(g)
    r.write_histogram(h)
    k1, k2 = pack.calibration(g.scale, samples, 1.0, os_)
    taps = pack.spatial_kernel(0.5, os_)
    mask = cv.active_mask(g.channel_curves)
    lut = cv.bake_curves(g.channel_curves)
    viz = pack.viz_uniforms(W, H, k1, k2, g.tonemap, os_, len(taps), g.background, False, mask,
                            hsl)
    for de_on in (True, False):
        r.present(g, samples, force_de_off=not de_on)
        if de_on:
            fo = orc.density(h, pack.de_uniforms(W * os_, H * os_, 9, 0, 0.4, k1, k2, os_))
            of, o8 = orc.tonemap(viz, taps, filtered=fo, curves=lut)
        else:
            of, o8 = orc.tonemap(viz, taps, hist=h, curves=lut)
        gf, g8 = r.read_image_f32(), r.read_image()
        assert np.allclose(gf, of, rtol=1e-5, atol=1e-5), de_on
        assert np.abs(g8.astype(int) - o8.astype(int)).max() <= 1, de_on
    r.destroy()


def test_curves_api_refusals():
    from pyr3_amd._lib import P3Error
    from pyr3_amd.renderer import Renderer
    W = H = 16
    g = genome(sierpinski(), scale=30.0)
    r = Renderer(W, H, 1)
    r.reset(g)
    taps = pack.spatial_kernel(0.5, 1)
    viz = np.ascontiguousarray(pack.viz_uniforms(W, H, 200.0, 1e-3, Tonemap(), 1, len(taps),
                                                 curves_mask=2))
    from pyr3_amd.renderer import _up
    assert r._lib.p3_present(r._h, None, _up(viz)) == -3        # P3_E_STATE: no LUT
    with pytest.raises(P3Error):
        r.set_curves(np.zeros(100, dtype=np.float32))
    bad = np.tile(np.linspace(0, 1, 256, dtype=np.float32), (5, 1))
    bad[2, 7] = 1.5
    with pytest.raises(P3Error):
        r.set_curves(bad)
    r.set_curves(np.tile(np.linspace(0, 1, 256, dtype=np.float32), (5, 1)))
    assert r._lib.p3_present(r._h, None, _up(viz)) == 0
    viz[16] = 64
    assert r._lib.p3_present(r._h, None, _up(viz)) == -1        # P3_E_ARG: unknown bit
    r.set_curves(None)
    viz[16] = 2
    assert r._lib.p3_present(r._h, None, _up(viz)) == -3
    r.destroy()
