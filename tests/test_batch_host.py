"""Host side of the batch object (no GPU): dispatch sizing, launch groups,
argument checks of p3_batch_add and the union kernel's compile."""
import ctypes as C

import numpy as np
import pytest

from pyr3_amd import batch, pack
from pyr3_amd.synth import flam3_core_pool, make_genome


def _jobs():
    return batch.batch_workload(n=12, seed=5, motion_fraction=0.25, subframes=3)


def test_job_dispatch_gives_whole_blocks_per_sub_dispatch():
    for j in _jobs():
        for turns in (64, 1024, 4096):
            walkers, iters = batch.job_dispatch(j, turns)
            k = len(j.subframes or [0])
            per = batch.apportion(walkers, [1.0] * k)
            assert all(c == per[0] and c % pack.CHAOS_BLOCK == 0 and c > 0 for c in per)
            assert walkers * iters >= j.quality * j.width * j.height
            # one walker fewer per sub-dispatch would not reach the samples
            assert (walkers - k) * iters < j.quality * j.width * j.height + walkers * 1.0 \
                or per[0] == pack.CHAOS_BLOCK


def test_plan_groups_matches_the_commit_grouping_rules():
    rng = np.random.default_rng(3)
    jobs = _jobs()
    # one genome of 18 xforms: its picks read the pick table, so it groups apart
    wide = make_genome(rng, 18, flam3_core_pool(), [1, 2], 64, 48, 4, 1, name="wide")
    jobs.insert(4, batch.Job(wide, 64, 48, 1, 4.0, seed=9, name="wide"))
    packed = [[batch.pack_for(g) for g in (j.subframes or [j.genome])] for j in jobs]
    for max_shapes in (1, 2, 5, 64):
        groups = batch.plan_groups(packed, max_shapes)
        assert all(1 <= len(g) <= max_shapes for g in groups)
        firsts = [i for g in groups for i in g]
        # every distinct shape of a group is listed once; the wide job last
        assert firsts[-1] == 4
        assert sorted(set(firsts)) == sorted(firsts)
        keys = {batch.union_key(packed, [i]) for i in range(len(jobs))}
        assert {batch.union_key(packed, [i]) for i in firsts} == keys or max_shapes < 64


def _add(lib, h, j, n_slots=None, viz_w=None):
    pgs = [batch.pack_for(g) for g in (j.subframes or [j.genome])]
    us, walkers, iters = batch._frame_blobs(j, pgs, 256, pack.FUSE)
    de, viz, taps = batch._present_blobs(j, walkers * iters)
    if viz_w is not None:
        viz = viz.copy()
        viz[0] = viz_w
    pal = np.ascontiguousarray(pack.pack_palette(j.genome).reshape(-1))
    n = len(pgs)
    xfp = (C.POINTER(C.c_float) * n)(*[batch._fp(p.xforms) for p in pgs])
    dsp = (C.POINTER(C.c_uint32) * n)(*[batch._up(p.distrib) for p in pgs])
    ufp = (C.POINTER(C.c_uint32) * n)(*[batch._up(u) for u in us])
    return lib.p3_batch_add(h, j.width, j.height, 1, batch._fp(taps), len(taps), batch._fp(pal),
                            n, xfp, n_slots or pgs[0].xforms.shape[0], dsp,
                            pgs[0].distrib.shape[0], ufp, batch._up(de), batch._up(viz))


def test_batch_add_checks_arguments_on_the_host():
    from pyr3_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.p3_batch_create(C.byref(h)) == 0
    try:
        jobs = _jobs()
        assert _add(lib, h, jobs[0]) == 0
        assert _add(lib, h, jobs[5]) == 1
        # tonemap block of another size, and a record count that does not
        # match the pick table, are refused with P3_E_ARG
        assert _add(lib, h, jobs[6], viz_w=jobs[6].width + 1) == -1
        assert b"tonemap uniforms" in lib.p3_batch_last_error(h)
        assert _add(lib, h, jobs[6], n_slots=1) == -1
        n = C.c_int()
        assert lib.p3_batch_frames(h, C.byref(n)) == 0 and n.value == 2
        # launch and reads before commit are state errors
        assert lib.p3_batch_launch(h) == -3
        out = np.zeros(4, np.uint8)
        assert lib.p3_batch_read_image(h, 0, batch._bp(out), 4) == -3
        assert lib.p3_batch_commit(h, 0, 2) == -1
        assert lib.p3_batch_commit(h, 4, 9) == -1
    finally:
        lib.p3_batch_destroy(h)


def test_sub_dispatches_of_a_frame_must_share_a_shape():
    from pyr3_amd import _lib
    lib = _lib.load()
    j = _jobs()[0]
    assert j.subframes
    n0 = len(j.genome.xforms)
    rng = np.random.default_rng(11)
    other = make_genome(rng, n0, ["linear", "swirl"], [2], j.width, j.height, 4, 1)
    j.subframes = [j.subframes[0], other]
    h = C.c_void_p()
    assert lib.p3_batch_create(C.byref(h)) == 0
    try:
        rc = _add(lib, h, j)
        assert rc == -1 and b"different genome shape" in lib.p3_batch_last_error(h)
    finally:
        lib.p3_batch_destroy(h)


def test_union_kernel_compiles(tmp_path, monkeypatch):
    """Two shapes (one with a final xform, one without) in one union kernel;
    compiled with hipRTC on the host, no GPU."""
    monkeypatch.setenv("P3_KERNEL_CACHE", str(tmp_path))
    from pyr3_amd import kcache
    key = "U|1n,0;2p,2,3;F|1n,1;1n,0;F2:1n,5;"
    k, rc, dt = kcache._compile_key(key, str(tmp_path))
    assert rc == 0
    assert len(list(tmp_path.glob("p3chaos-*.co"))) == 1
