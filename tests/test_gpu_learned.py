"""The learned path's per-genome gate on the GPU: the plain render at the
reference quality judges a genome's learned frame once; at or above the
target the genome renders through the denoiser, below it through the plain
path (the fallback)."""
import math

import numpy as np
import pytest

from pyr3_amd import learned, synth
from pyr3_amd.denoise.weights import DEFAULT

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not DEFAULT.exists(), reason="no trained weights in the tree")]

W, H = 320, 180


def test_gate_routes_by_the_plain_judge():
    g = synth.bench_genome(0, W, H)
    lr = learned.LearnedRenderer(W, H, quality=3, ref_quality=400)
    try:
        db = lr.check(g, seed=5)
        assert math.isfinite(db) and db > 20.0
        # the verdict is kept: renders follow it without re-judging
        lr.target_db = db - 1.0
        assert lr.render(g, 9) == "learned"
        img_l = lr.read_image()
        lr.render_learned(g, 9)
        assert np.array_equal(img_l, lr.read_image())
        lr.target_db = db + 1.0
        assert lr.render(g, 9, plain_quality=20) == "plain"
        img_p = lr.read_image()
        lr.render_plain(g, 9, 20)
        assert np.array_equal(img_p, lr.read_image())
        assert not np.array_equal(img_l, img_p)
    finally:
        lr.close()


def test_unjudged_genome_is_judged_on_first_render():
    g = synth.bench_genome(1, W, H)
    lr = learned.LearnedRenderer(W, H, quality=3, ref_quality=400, target_db=0.0)
    try:
        assert lr.render(g, 3) == "learned"
        assert len(lr.verdicts) == 1
    finally:
        lr.close()
