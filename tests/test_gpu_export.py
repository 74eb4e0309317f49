"""Output side on the GPU: device linear export, the render command's formats,
auto-exposure.

* `p3_read_linear` (histogramToLinearRgba on the device) is bit-identical to
  the oracle's f64 restatement for os 1, 2, 3 and counts up to 2^32 - 1.
* One render read out as png8, png16, exr and exr-linear: png16 / 257 is
  png8 within 1 LSB (both from the same f32 tonemap output), exr is the
  sRGB-to-linear of the binary16 image, exr-linear decodes to finite
  scene-linear values with alpha <= 1 (its values are the device export
  checked bit for bit above); files decode with the independent readers of
  tests/test_export.py.
* Auto-exposure on a real render lands within the deadband of its probe
  target, or runs out of passes.
"""
from pathlib import Path

import numpy as np
import pytest

from helpers import genome, sierpinski
from pyr3_amd import export as X
from pyr3_amd import pack

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).parent / "golden"


@pytest.mark.parametrize("os_", [1, 2, 3])
def test_device_linear_export_bit_exact(orc, os_):
    from pyr3_amd.renderer import Renderer
    W, H = 37, 23
    rng = np.random.default_rng(70 + os_)
    h = rng.integers(0, 2 ** 32, (H * os_, W * os_, 4), dtype=np.uint64).astype(np.uint32)
    h[0, 0] = 0xFFFFFFFF
    h[1, :] = 0
    g = genome(sierpinski(), scale=30.0)
    r = Renderer(W, H, os_)
    r.reset(g)
    r.write_histogram(h)
    k1, k2 = pack.calibration(g.scale, 3.3e6, 1.7, os_)
    dev = r.read_linear(k1, k2)
    r.destroy()
    ref = orc.histogram_to_linear_rgba(h, W, H, os_, k1, k2)
    assert np.array_equal(dev.view(np.uint32), ref.view(np.uint32))


def test_render_command_formats(tmp_path):
    import sys
    sys.path.insert(0, str(GOLDEN))
    sys.path.insert(0, str(Path(__file__).parent))
    from png8 import read_png
    from test_export import _decode_png16, _read_exr
    from pyr3_amd.render import main
    fid = "244.57686"
    src = GOLDEN / "flam3" / fid / f"{fid}.flam3"
    outs = {}
    for fmt in ("png8", "png16", "exr", "exr-linear"):
        out = tmp_path / f"o-{fmt}.{'exr' if 'exr' in fmt else 'png'}"
        print(f"render {fmt}", flush=True)
        assert main([str(src), str(out), "--seed", "11", "--format", fmt,
                     "--no-auto-exposure"]) == 0
        outs[fmt] = out.read_bytes()
    p8 = read_png(tmp_path / "o-png8.png")
    p16 = _decode_png16(outs["png16"])
    assert p16.shape == p8.shape
    assert np.abs(p16.astype(np.int64) / 257.0 - p8.astype(np.int64)).max() <= 1.0
    _, names, exr = _read_exr(outs["exr"])
    assert names == ["A", "B", "G", "R"] and exr.shape == p8.shape
    # the exr holds the linear light of the same display image the png16 holds
    disp = p16.astype(np.float64) / 65535.0
    assert np.allclose(exr[..., :3], X.srgb_to_linear(disp[..., :3]), atol=2e-4)
    _, _, lin = _read_exr(outs["exr-linear"])
    assert lin.shape == p8.shape and np.isfinite(lin).all()
    assert lin[..., 3].max() <= 1.0 and lin[..., :3].max() > 0


def test_auto_exposure_fit_on_a_render():
    from pyr3_amd.renderer import Renderer
    g = genome(sierpinski(), scale=150.0, cx=0.5, cy=0.5)
    g.size = {"width": 512, "height": 512}
    g.quality = 60.0
    r = Renderer(512, 512, 1)
    samples = r.render(g, seed=3, walkers=131072)
    target = X.measure_probe_luminance(g)
    initial = X.compute_mean_luminance(r.read_image())

    def rerender(b):
        r.present(X.with_brightness(g, b), samples)
        return X.compute_mean_luminance(r.read_image())
    fit = X.fit_brightness_to_target(g.tonemap.brightness, target, initial, rerender)
    r.destroy()
    print(f"target {target:.2f} initial {initial:.2f} fit {fit}", flush=True)
    assert target > 0 and initial > 0
    if fit.corrected:
        assert (abs(fit.final_mean / target - 1) <= X.AE_DEADBAND * 1.5
                or fit.iters == X.AE_MAX_ITERS)
    else:
        assert abs(target / initial - 1) <= X.AE_DEADBAND


def test_atomic_probe_counts_every_add():
    """The scatter-add probe behind the chaos roofline's atomic peak: every
    add lands (the four words of a cell stay equal, the total is exact)."""
    from pyr3_amd.renderer import Renderer
    r = Renderer(64, 48, 1)
    r.reset(genome(sierpinski(), scale=30.0))
    rate, per_splat = r.probe_atomics(blocks=64, per_thread=16)
    h = r.read_histogram().astype(np.int64)
    r.destroy()
    assert rate > 0 and per_splat in (1, 2, 4)
    assert h.sum() == 4 * 64 * 256 * 16
    assert np.array_equal(h[..., 0], h[..., 3]) and np.array_equal(h[..., 1], h[..., 2])
    assert np.array_equal(h[..., 0], h[..., 1])
