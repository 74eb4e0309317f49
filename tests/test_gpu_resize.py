"""Renderer resize between jobs: allocations kept, tables re-uploaded only
when they change.

`p3_resize` queues the spatial-filter taps and the DE kernel-norm table on
the stream without waiting for earlier work, and uploads them only when the
taps change (a change waits for the stream before the handle's host copy is
rewritten). A renderer taken through sizes, oversamples and filter radii
must give, job for job, the bytes a fresh renderer of that configuration
gives: the histogram adds are exact integers, so equal seeds give equal
images.
"""
import numpy as np
import pytest

from helpers import genome, sierpinski

pytestmark = pytest.mark.gpu

# (width, height, oversample, filter radius): the same taps twice in a row,
# a new size with the same taps, new taps by radius, new taps by oversample,
# and grids with an odd number of bins (135 x 119, 97 x 61) inside larger
# allocations: every per-bin buffer is cleared to its last word
STEPS = [(96, 64, 1, 1.0), (96, 64, 1, 1.0), (128, 80, 1, 1.0), (128, 80, 1, 2.5),
         (64, 48, 2, 2.5), (96, 64, 1, 1.0), (135, 119, 1, 1.0), (97, 61, 1, 1.0)]


def _render(r, k):
    g = genome(sierpinski(), scale=30.0 + 5 * k)
    g.quality = 40.0
    r.render(g, seed=100 + k, walkers=8192)


def test_resized_renderer_matches_fresh_renderers():
    from pyr3_amd.renderer import Renderer
    ref = []
    for k, (w, h, os_, fr) in enumerate(STEPS):
        r = Renderer(w, h, os_, filter_radius=fr)
        _render(r, k)
        ref.append(r.read_image())
        r.destroy()
    r = Renderer(*STEPS[0][:3], filter_radius=STEPS[0][3])
    for k, (w, h, os_, fr) in enumerate(STEPS):
        print(f"[resize] step {k}: {w}x{h} os{os_} radius {fr}", flush=True)
        # the resize is queued behind the previous job, which the read of
        # the previous step has already drained, except for the first step
        r.resize(w, h, os_, filter_radius=fr)
        _render(r, k)
        img = r.read_image()
        assert img.shape == ref[k].shape and np.array_equal(img, ref[k]), k
    r.destroy()
