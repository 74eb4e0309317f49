"""The library's generic chaos kernel vs the per-shape specialised kernels.

Both run the same per-walker algorithm in the same arithmetic order (the
generic chain loops over each record's slots and enters the catalog switch;
the specialised one has the slots unrolled), so every dispatch must give the
same histogram and trace bit for bit: final xform with direct colour, 16
catalog entries including randomised and direct-colour ones, xform blend,
more than 16 xforms (pick table read from HBM), batched sub-dispatches, and
the 'auto' policy (generic until the specialised kernel is compiled).

Skipped: the first GPU run of this file stalled (no output for over 180 s)
inside test_generic_equals_specialised[final_dc_preblur], which runs the
specialised and then the generic kernel on that genome. Reading the generic
kernel, its launch and host code, the hipRTC compile cache and the kernel's
resource usage (225 VGPRs, 48 B/lane static scratch, no dynamic stack, no
recursion or indirect calls) showed no cause, so the generic kernel and its
kernel-policy API were taken out of the library. The tests stay as the record
of that case; they call an API the library no longer has.
"""
import numpy as np
import pytest

from helpers import genome, sierpinski, xform
from pyr3_amd import catalog as cat
import test_gpu_variations as TV

pytestmark = [pytest.mark.gpu, pytest.mark.skip(
    reason="generic chaos kernel stalled on the GPU (final_dc_preblur case, cause not found); "
           "kernel and kernel-policy API removed")]

V = cat.VARIATION_INDEX


def _render(g, policy, W=40, H=40, walkers=2048, iters=30, fuse=5, seed=321, trace=True):
    from pyr3_amd.renderer import Renderer
    r = Renderer(W, H, 1)
    r.set_kernel_policy(policy)
    r.reset(g)
    if trace:
        r.set_trace(256, iters + fuse)
    r.iterate(g, seed, walkers, iters, fuse=fuse)
    h = r.read_histogram()
    t = r.read_trace() if trace else None
    r.destroy()
    return h, t


def _genomes():
    dc = [xform(1.0, (0.5, 0, 0, 0, 0.5, 0), vars=((0, 1.0), (V["dc_linear"], 0.3)), color=0.1),
          xform(1.0, (0.5, 0, 0.5, 0, 0.5, 0), vars=((V["julia"], 0.6), (V["swirl"], 0.4)),
                color=0.5, post=(0.9, 0.1, 0.0, -0.1, 0.9, 0.05)),
          xform(0.5, (0.5, 0, 0.25, 0, 0.5, 0.5), vars=((V["pre_blur"], 0.2), (2, 0.8)), color=0.9)]
    fin = xform(1.0, (1, 0, 0, 0, 1, 0), vars=((V["tangent"], 0.5), (V["dc_perlin"], 0.1)),
                params=[[], [3.0, 2.0, 0.25]])
    rng = np.random.default_rng(4)
    many = [xform(float(rng.uniform(0.2, 1)), (0.4, 0.05, float(rng.uniform(-0.5, 0.5)), -0.05, 0.4,
                                               float(rng.uniform(-0.5, 0.5))),
                  vars=((int(rng.choice([0, 1, 2, 3, 5, 13])), 1.0),), color=j / 19)
            for j in range(20)]
    return {
        "final_dc_preblur": genome(dc, scale=20.0, final=fin),
        "group16": genome([TV._xform(i) for i in TV.GROUPS[6]], scale=20.0),
        "blend": genome(sierpinski(), scale=20.0, cx=0.5, cy=0.5, xform_blend=0.4),
        "xforms20": genome(many, scale=20.0),
    }


@pytest.mark.parametrize("name", ["final_dc_preblur", "group16", "blend", "xforms20"])
def test_generic_equals_specialised(name):
    g = _genomes()[name]
    hs, ts = _render(g, "specialised")
    hg, tg = _render(g, "generic")
    assert hs[..., 3].sum() > 0
    assert np.array_equal(tg.view(np.uint32), ts.view(np.uint32))
    assert np.array_equal(hg, hs)


def test_generic_batched_and_auto():
    from pyr3_amd.renderer import Renderer
    g = _genomes()["final_dc_preblur"]
    subs = [g.clone() for _ in range(3)]
    for i, s in enumerate(subs):
        s.xforms[0].affine.c += 0.01 * i
    out = {}
    for policy in ("specialised", "generic"):
        r = Renderer(40, 40, 1)
        r.set_kernel_policy(policy)
        r.reset(g)
        r.iterate_multi(subs, [5, 6, 7], [700, 1100, 300], 25)
        out[policy] = r.read_histogram()
        r.destroy()
    assert out["generic"][..., 3].sum() > 0
    assert np.array_equal(out["generic"], out["specialised"])
    # auto: a new shape runs generic now and specialised once compiled; the
    # histogram is the same either way
    g2 = genome(sierpinski() + [xform(0.3, (0.3, 0, 0, 0, 0.3, 0), vars=((V["bubble"], 1.0),))],
                scale=20.0)
    ha, _ = _render(g2, "auto", trace=False)
    hs, _ = _render(g2, "specialised", trace=False)
    assert np.array_equal(ha, hs)
    with pytest.raises(ValueError):
        Renderer(8, 8, 1).set_kernel_policy("fast")


def test_batch_generic_matches_specialised():
    from pyr3_amd import batch
    jobs = batch.batch_workload(n=4, seed=91, motion_fraction=0.5, subframes=3)
    for j in jobs:
        j.quality = 3.0
    imgs = {}
    for kernels in ("specialised", "generic"):
        br = batch.BatchRenderer(n_streams=2, compile_workers=4, kernels=kernels)
        imgs[kernels] = [r.image for r in br.render(jobs, read_images=True)]
        br.close()
    for a, b in zip(imgs["generic"], imgs["specialised"]):
        assert np.array_equal(a, b)
