"""Host tables: xform-pick table, spatial taps, calibration, DE kernel sums.

Known answers come from the reference's own tests:
  src/chaos-xform-pick.test.ts:39-110  (50/50 flip at 8192, 3:1 split, xaos)
  src/spatial-filter.test.ts:18-53     (tap counts, symmetry, sums)
  src/calibration.test.ts:5-59         (closed forms, oversample invariance)
  src/density.test.ts:61-71            (each bucket scatters weight 1)
The product packer (pyr3_amd.pack) and the oracle (oracle/orc_image.cpp) are
two implementations of each table and must agree exactly.
"""
import math

import numpy as np
import pytest

from pyr3_amd import pack
from pyr3_amd.genome import Affine, ColorStop, Genome, Palette, Variation, Xform

GRAIN = 16384


def _genome(weights, xaos=None):
    xs = []
    for i, w in enumerate(weights):
        x = Xform(weight=w, color=0.0, color_speed=0.0, affine=Affine(),
                  variations=[Variation(0, 1.0)])
        if xaos is not None:
            x.xaos = list(xaos[i])
        xs.append(x)
    return Genome(name="t", xforms=xs, scale=10, cx=0, cy=0,
                  palette=Palette("p", [ColorStop(0, 0, 0, 0), ColorStop(1, 1, 1, 1)]))


@pytest.mark.parametrize("impl", ["pack", "oracle"])
def test_equal_weights_flip_at_half(orc, impl):
    t = pack.pack_xform_distrib(_genome([1, 1])) if impl == "pack" else orc.build_distrib([1, 1])
    half = GRAIN // 2
    for row in (t[0], t[-1]):
        assert row[0] == 0 and row[half - 1] == 0 and row[half] == 1 and row[-1] == 1


@pytest.mark.parametrize("impl", ["pack", "oracle"])
def test_three_to_one_split(orc, impl):
    t = pack.pack_xform_distrib(_genome([3, 1])) if impl == "pack" else orc.build_distrib([3, 1])
    c0 = int(np.sum(t[0] == 0))
    c1 = int(np.sum(t[0] == 1))
    assert abs(c0 - 0.75 * GRAIN) <= 1 and abs(c1 - 0.25 * GRAIN) <= 1 and c0 + c1 == GRAIN


@pytest.mark.parametrize("impl", ["pack", "oracle"])
def test_xaos_zero_row(orc, impl):
    xa = [[1, 0], [1, 1]]
    t = (pack.pack_xform_distrib(_genome([1, 1], xa)) if impl == "pack"
         else orc.build_distrib([1, 1], np.array(xa, dtype=np.float64)))
    assert np.all(t[0] == 0)
    assert t[1][GRAIN // 2 - 1] == 0 and t[1][GRAIN // 2] == 1
    assert t.shape == (3, GRAIN)


def test_pack_matches_oracle_on_random_tables(orc):
    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 6, 13, 40):
        w = rng.uniform(0.0, 2.0, n)
        w[rng.random(n) < 0.2] = 0.0
        if w.sum() == 0:
            w[0] = 1.0
        xa = rng.uniform(0, 1.5, (n, n))
        xa[rng.random((n, n)) < 0.3] = 0.0
        a = pack.pack_xform_distrib(_genome(list(w), xa))
        b = orc.build_distrib(w, xa)
        assert np.array_equal(a, b), n


def test_negative_weight_rejected(orc):
    with pytest.raises(ValueError):
        pack.pack_xform_distrib(_genome([1, -1]))
    with pytest.raises(ValueError):
        orc.build_distrib([1, -1])


# (radius, oversample, taps) from src/spatial-filter.test.ts:45-86
@pytest.mark.parametrize("radius,os_,taps", [
    (0.5, 1, 3), (1, 1, 5), (1.5, 1, 5), (2, 1, 7), (3, 1, 11), (0.1, 1, 3),
    (1, 2, 8), (1, 4, 14), (0.5, 4, 8), (0.5, 2, 4), (0.5, 3, 5)])
def test_spatial_kernel_taps(orc, radius, os_, taps):
    a = pack.spatial_kernel(radius, os_)
    b = orc.spatial_kernel(radius, os_)
    assert len(a) == taps and np.array_equal(a, b)
    assert abs(float(np.sum(a.astype(np.float64))) - 1.0) < 1e-6
    assert np.allclose(a, a[::-1])


def test_spatial_kernel_known_values():
    k1 = pack.spatial_kernel(0.5, 1)
    assert abs(k1[1] - 0.99933) < 1e-5 and abs(k1[0] - 3.35e-4) < 1e-6
    k2 = pack.spatial_kernel(0.5, 2)
    assert abs(k2[0] - 0.0089) < 1e-4 and abs(k2[1] - 0.4911) < 1e-4
    assert len(pack.spatial_kernel(3.0, 2)) == 17  # clamp at 17


def test_calibration_closed_form(orc):
    k1, k2 = pack.calibration(240.0, 1e6, 1.5, 2)
    assert k1 == 1.5 * 255 * 268 / 256
    assert k2 == 4 * 240.0 ** 2 / (255 * 1e6)
    assert (k1, k2) == orc.calibration(240.0, 1e6, 1.5, 2)
    # count * k2 is oversample-invariant when counts scale with os^2 at fixed samples/pixel
    _, k2a = pack.calibration(100.0, 1e6, 1, 1)
    _, k2b = pack.calibration(100.0, 4e6, 1, 2)
    assert math.isclose(k2a, k2b)


def test_kernel_norm_sums_to_one_per_bucket(orc):
    """Sum over the disc of exp(-d^2/2s^2)/LUT[r] is 1 for every radius."""
    lut = orc.kernel_norm_lut()
    assert lut[0] == 1.0
    for r in (1, 2, 3, 5, 9, 10, 19, 40):
        s = r / 3.0
        tot = 0.0
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                d2 = dx * dx + dy * dy
                if d2 <= r * r:
                    tot += math.exp(-d2 / (2 * s * s)) / float(lut[r])
        assert abs(tot - 1.0) < 1e-5, r


def test_compute_dispatch_matches_reference_schedules():
    # (quality, W, H) -> (walkers, iters): the reference's computeDispatch table
    assert pack.compute_dispatch(50, 256, 256)[:2] == (800, 4096)
    assert pack.compute_dispatch(200, 1024, 768)[:2] == (1024, 153600)
    w, i, s = pack.compute_dispatch(200, 3840, 2160)
    assert (w, i) == (1583, 1048576) and s == 1583 * 1048576
    assert pack.compute_dispatch(10, 1920, 1080)[:2] == (1024, 20250)
    assert pack.compute_dispatch(4000, 1920, 1080)[:2] == (7911, 1048576)
    assert pack.compute_dispatch(16, 100, 100, walkers_override=7)[:2] == (7, 22858)
