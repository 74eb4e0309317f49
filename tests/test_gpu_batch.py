"""Batch object on the GPU: frames of different genome shapes, sizes and
sub-dispatch counts in union-kernel launches on one captured HIP graph give
the histograms and images of one renderer per frame, bit for bit, on every
replay."""
import numpy as np
import pytest

from pyr3_amd import batch, pack
from pyr3_amd.renderer import Renderer
from pyr3_amd.synth import flam3_core_pool, make_genome

pytestmark = pytest.mark.gpu


def mixed_jobs():
    """Six workload jobs (two motion-blurred, 4 sub-frames), one of 18
    xforms (a group of its own: its picks read the pick table), one at
    oversample 2 with a final xform, and one with DE off."""
    jobs = batch.batch_workload(n=6, seed=77, motion_fraction=0.34, subframes=4)
    for j in jobs:
        j.quality = 4.0
    rng = np.random.default_rng(21)
    wide = make_genome(rng, 18, flam3_core_pool(), [1, 2], 96, 64, 4, 1, name="wide")
    jobs.append(batch.Job(wide, 96, 64, 1, 4.0, seed=31, name="wide"))
    fin = make_genome(rng, 4, flam3_core_pool(), [2], 80, 60, 6, 2, final=True,
                      post_fraction=0.5, name="final-os2")
    jobs.append(batch.Job(fin, 80, 60, 2, 6.0, seed=32, name="final-os2"))
    nde = make_genome(rng, 3, ["linear", "spherical", "swirl"], [2], 72, 72, 5, 1, name="no-de")
    nde.density = None
    jobs.append(batch.Job(nde, 72, 72, 1, 5.0, seed=33, name="no-de"))
    return jobs


def single_render(j, walkers, iters):
    r = Renderer(j.width, j.height, j.oversample)
    r.reset(j.genome)
    if j.subframes:
        k = len(j.subframes)
        counts = batch.apportion(walkers, [1.0] * k)
        s = r.iterate_multi(j.subframes, [(j.seed + i * batch.SUBFRAME_SEED_STEP) & 0xFFFFFFFF
                                          for i in range(k)], counts, iters)
    else:
        s = r.iterate(j.genome, j.seed, walkers, iters)
    r.present(j.genome, s)
    out = r.read_histogram(), r.read_image(), s
    r.destroy()
    return out


def test_batch_graph_matches_single_renders_on_every_replay():
    jobs = mixed_jobs()
    plan = batch.BatchPlan(turns=48, max_shapes=3, lanes=2)
    packed = [[batch.pack_for(g) for g in (j.subframes or [j.genome])] for j in jobs]
    assert len(batch.plan_groups(packed, plan.max_shapes)) >= 4
    gb = batch.GraphBatch(jobs, plan, packed=packed)
    try:
        st = gb.stats()
        print("batch stats", st, flush=True)
        assert st["groups"] == len(batch.plan_groups(packed, plan.max_shapes))
        assert st["sub_dispatches"] == sum(len(j.subframes or [0]) for j in jobs)
        gb.launch()
        first = [(gb.read_histogram(i), gb.read_image(i)) for i in range(len(jobs))]
        for rep in range(2):
            gb.launch()
            gb.sync()
            for i in range(len(jobs)):
                assert np.array_equal(gb.read_histogram(i), first[i][0]), (rep, jobs[i].name)
                assert np.array_equal(gb.read_image(i), first[i][1]), (rep, jobs[i].name)
        for i, j in enumerate(jobs):
            h, img, s = single_render(j, gb.walkers[i], gb.iters[i])
            assert s == gb.samples[i]
            assert h[..., 3].sum() > 0, j.name
            assert np.array_equal(first[i][0], h), j.name
            assert np.array_equal(first[i][1], img), j.name
    finally:
        gb.close()


def test_batch_renderer_reads_every_frame():
    jobs = mixed_jobs()[:4]
    br = batch.BatchRenderer(batch.BatchPlan(turns=64, max_shapes=2, lanes=1), compile_workers=4)
    res = br.render(jobs, read_images=True)
    assert br.timings["groups"] == 2 and br.timings["gpu_ms"] > 0
    for j, r in zip(jobs, res):
        assert r.image.shape == (j.height, j.width, 4)
        assert r.image[..., 3].min() == 255
        assert r.samples >= j.quality * j.width * j.height
