"""Splat colour modes and colour-index capture on the GPU vs the oracle.

The closed-form configurations of tests/test_colour_modes.py (linear maps,
no jitter, colours fixed per splat) must give the oracle's histogram and
capture buffer bit for bit. On a general linear genome the trajectories are
exact but the mode colours go through atan2/exp/log, whose device and host
results may differ in the last ulp: there the count channel is exact and the
colour channels agree within 1e-3 of their total.
"""
import numpy as np
import pytest

from helpers import blobs, genome, sierpinski, xform
from pyr3_amd import pack
import test_colour_modes as T

pytestmark = pytest.mark.gpu

W = H = T.W


def _gpu(g, color, iters=60, fuse=10, walkers=256, seed=9):
    from pyr3_amd.renderer import Renderer
    r = Renderer(W, H, 1)
    r.reset(g)
    r.iterate(g, seed, walkers, iters, walker_jitter=0.0, fuse=fuse, color=color)
    h = r.read_histogram()
    idx = r.read_index() if color is not None and color.capture_index else None
    r.destroy()
    return h, idx


CASES = [
    ("flow", lambda: genome([xform(1.0, (1, 0, 1, 0, 0, 0)), xform(1.0, (1, 0, -1, 0, 0, 0))],
                            scale=5.0), pack.ColorOpts(mode="flow")),
    ("phase+", lambda: T._axis_genome(1.0), pack.ColorOpts(mode="phase", phase_freq=0.0)),
    ("phase-", lambda: T._axis_genome(-1.0), pack.ColorOpts(mode="phase", phase_freq=0.0)),
    ("trap-glow", lambda: T._axis_genome(), pack.ColorOpts(
        mode="trap-distance", trap_kind="line", trap_cy=0.5, trap_falloff=2.0)),
    ("trap-rings", lambda: T._axis_genome(), pack.ColorOpts(
        mode="trap-distance", trap_kind="line", trap_mode="rings", trap_cy=0.5, trap_freq=3.0)),
]


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_closed_form_configurations_bit_exact(orc, name):
    _, make, opts = next(c for c in CASES if c[0] == name)
    g = make()
    hg, _ = _gpu(g, opts)
    ho, _ = T._run(orc, g, opts)
    assert ho[..., 3].sum() > 0
    assert np.array_equal(hg, ho)


def test_capture_bit_exact(orc):
    g = genome([xform(1.0, (0.5, 0, 0, 0, 0.5, 0), color=0.25, speed=1.0),
                xform(1.0, (0.5, 0, 0.5, 0, 0.5, 0.5), color=0.75, speed=1.0)],
               scale=20.0, cx=0.5, cy=0.5)
    opts = pack.ColorOpts(capture_index=True, mode="phase", phase_freq=2.0)
    hg, ig = _gpu(g, opts)
    ho, io, _ = T._run(orc, g, opts, capture=True)
    assert io.sum() > 0
    assert np.array_equal(ig, io)
    assert np.array_equal(hg[..., 3], ho[..., 3])


@pytest.mark.parametrize("opts", [
    pack.ColorOpts(mode="flow", flow_strength=0.7, flow_scale=3.0),
    pack.ColorOpts(mode="trap-distance", trap_kind="point", trap_cx=0.3, trap_cy=0.6),
    pack.ColorOpts(mode="trap-distance", trap_kind="circle", trap_mode="rings", trap_radius=0.4,
                   trap_freq=5.0, trap_strength=0.8),
    pack.ColorOpts(mode="phase", phase_strength=0.9, phase_freq=1.5),
])
def test_general_modes_within_tolerance(orc, opts):
    g = genome(sierpinski(), scale=20.0, cx=0.5, cy=0.5)
    hg, _ = _gpu(g, opts, iters=80)
    ho, _ = T._run(orc, g, opts, iters=80)
    assert np.array_equal(hg[..., 3], ho[..., 3])
    d = np.abs(hg[..., :3].astype(np.int64) - ho[..., :3].astype(np.int64)).sum()
    assert d <= 1e-3 * ho[..., :3].astype(np.int64).sum(), d
